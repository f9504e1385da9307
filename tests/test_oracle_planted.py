"""Each pin of the oracle's charged-step primitives catches a plausible
mistake: a copy of the oracle with one mistake planted (a textual change of
oracle/rrea_oracle.cpp, compiled on the CPU into a temporary directory) must
fail the tests that pin that function, while the unchanged oracle passes
them.  The mistakes are the ones the reference's own smokes guard against
(a hoisted refreeze, a loss booked to the substep end, a reused RNG ordinal,
a refused substep counted) and slips of sign, index or formula in the
restated arithmetic."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "rrea_oracle.cpp")

# (name, original text, planted text, tests that must fail)
PLANTS = [
    ("hoisted refreeze",
     "    while (rem < se && k >= cutoff) {\n        if (ord >= max_ordinal) fail(ORC_ERR_INVALID, \"event ordinal overflow\");\n"
     "        double const frozen_e = k;",
     "    double const hoisted = k;\n    while (rem < se && k >= cutoff) {\n"
     "        if (ord >= max_ordinal) fail(ORC_ERR_INVALID, \"event ordinal overflow\");\n"
     "        double const frozen_e = hoisted;",
     "test_oracle_interleave.py::test_interleave_refreezes_at_the_survivor_energy"),
    ("loss booked to the substep end",
     "        SegLimit const sl = continuous(rem, ef, frozen_e);",
     "        SegLimit const sl = continuous(rem, se, frozen_e);",
     "test_oracle_interleave.py::test_interleave_refreezes_at_the_survivor_energy"),
    ("null event reuses its RNG ordinal",
     "            if (!open_fr) fail(ORC_ERR_INVALID, \"selected channel closed before its sampled depth\");\n"
     "            rem = ef;\n            ++ord;",
     "            if (!open_fr) fail(ORC_ERR_INVALID, \"selected channel closed before its sampled depth\");\n"
     "            rem = ef;",
     "test_oracle_interleave.py::test_interleave_drag_closed_channel_is_a_null_event"),
    ("refused substep counted",
     "        if (!advance(se)) { exited = true; break; }",
     "        if (!advance(se)) { exited = true; completed = sc + 1U; break; }",
     "test_oracle_interleave.py::test_substep_skeleton"),
    ("half kick books the whole work",
     "    double const hw = 0.5 * f.work;",
     "    double const hw = f.work;",
     "test_oracle_interleave.py::test_half_kick test_oracle_known_answers.py::test_work_energy_in_a_uniform_field"),
    ("ion pairs not clamped to the segment",
     "        double const lb = std::max(b, it.begin), le = std::min(e, it.end);\n"
     "        if (!(le > lb)) continue;\n        double const raw_i",
     "        double const lb = it.begin, le = it.end;\n"
     "        if (!(le > lb)) continue;\n        double const raw_i",
     "test_oracle_interleave.py::test_ion_pair_partition"),
    ("cutoff bisection turned round",
     "        if (l < to_cut) lo = mid; else hi = mid;\n    }\n    return {0.5 * (lo + hi), true};",
     "        if (l < to_cut) hi = mid; else lo = mid;\n    }\n    return {0.5 * (lo + hi), true};",
     "test_oracle_interleave.py::test_cutoff_limiter_on_a_linear_loss "
     "test_oracle_known_answers.py::test_cutoff_crossing_positron_stops_at_the_closed_form_point"),
    ("soft radiative stopping dropped",
     "    double const total = col + soft;\n    SegLimit const lim",
     "    double const total = col;\n    SegLimit const lim",
     "test_oracle_known_answers.py::test_cutoff_crossing_electron_ion_pairs"),
    ("optical depth from u instead of 1 - u",
     "    h.sampled = -std::log(std::max(1.0e-30, 1.0 - u1));",
     "    h.sampled = -std::log(std::max(1.0e-30, u1));",
     "test_oracle_interleave.py::test_competing_hazard_selection_closed_forms"),
    ("event depth against one channel's rate",
     "                double const total_stp = pos ? stp[0] + stp[1] + stp[2] : stp[0] + stp[1];",
     "                double const total_stp = stp[0];",
     "test_oracle_known_answers.py::test_event_fraction_channel_split_and_depth"),
    ("Moller groups interpolated linearly",
     "        return std::exp(std::log(y0) + t * (std::log(y1) - std::log(y0)));\n    }\n\n"
     "    // Load :2146-2179",
     "        return y0 + t * (y1 - y0);\n    }\n\n    // Load :2146-2179",
     "test_oracle_samplers.py::test_moller_and_brems_samplers_restated"),
    ("hard-Moller mean on the left Riemann sum",
     "                mean += sample_moller(r.x[i], (static_cast<double>(s) + 0.5) / samples);",
     "                mean += sample_moller(r.x[i], static_cast<double>(s) / samples);",
     "test_oracle_samplers.py::test_hard_moller_mean_table_against_quadrature"),
    ("GS angles interpolated linearly in u",
     "        one_minus_cos = d0 * std::exp(fu * std::log(d1 / d0));",
     "        one_minus_cos = d0 + fu * (d1 - d0);",
     "test_oracle_samplers.py::test_gs_sampler_nodes_interpolation_and_mean"),
]


def build(tmp, name, src_text):
    d = os.path.join(tmp, name.replace(" ", "_"))
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, "rrea_oracle.cpp")
    with open(path, "w") as f:
        f.write(src_text)
    lib = os.path.join(d, "liboracle.so")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-shared",
                           "-I" + os.path.join(ROOT, "oracle"), "-o", lib, path,
                           os.path.join(ROOT, "oracle", "rrea_maxwell_oracle.cpp")])
    return lib


def run_tests(lib, targets):
    env = dict(os.environ, RREA_ORACLE_LIB=lib)
    args = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"]
    args += [os.path.join(ROOT, "tests", t) for t in targets.split()]
    return subprocess.run(args, cwd=ROOT, env=env, capture_output=True, text=True)


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("planted"))
    text = open(SRC).read()
    jobs = []
    for name, old, new, targets in PLANTS:
        assert text.count(old) == 1, f"{name}: the planted site is not unique in the oracle"
        jobs.append((name, text.replace(old, new), targets))
    jobs.append(("unchanged", text, " ".join(sorted({t for *_, t in PLANTS}))))
    with ThreadPoolExecutor(max_workers=8) as ex:
        libs = list(ex.map(lambda j: build(tmp, j[0], j[1]), jobs))
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = list(ex.map(lambda a: run_tests(a[0], a[1][2]), zip(libs, jobs)))
    return {j[0]: r for j, r in zip(jobs, results)}


def test_unchanged_oracle_passes_the_pins(planted):
    r = planted["unchanged"]
    assert r.returncode == 0, r.stdout[-3000:]


@pytest.mark.parametrize("name", [p[0] for p in PLANTS])
def test_planted_mistake_is_caught(planted, name):
    r = planted[name]
    assert r.returncode != 0 and "failed" in r.stdout, f"{name} was not caught:\n{r.stdout[-2000:]}"
