"""The secondary-group conservation closure (RreaValidateSecondaryGroup
Conservation, OV/RreaParticleInteraction.H:1099-1252) in the oracle, pinned
by the reference's own smoke cases (native_amrex_rrea/src/
RreaSecondaryGroupSmoke.cpp:342-523: every valid group accepted, every
planted bad group rejected with the same result bits), by the tolerance rule
(64 eps times the charge throughput |q_in| + |q_out + q_escaped| + the
particles' and sources' |q|, and times the larger energy side), and by
deletions: dropping any one nonzero term of a closed group must break it."""
import copy

import numpy as np
import pytest

import group_cases as gc
from oracle import binding as ob

ALL = 1 | 2 | 4 | 8


def oracle(g, cut=gc.CUT):
    return ob.validate_group(gc.build(ob.Group, ob.GroupEntry, g), cut)


@pytest.mark.parametrize("name", sorted(gc.VALID))
def test_reference_valid_groups_close(name):
    bits, dq, de = oracle(gc.VALID[name])
    assert bits == ALL, (name, bits, dq, de)


@pytest.mark.parametrize("name", sorted(gc.INVALID))
def test_reference_invalid_groups_rejected(name):
    g, expected = gc.INVALID[name]
    bits, _, _ = oracle(g)
    assert bits == expected, (name, bits, expected)


@pytest.mark.parametrize("process", [gc.CONTINUOUS, gc.SEED])
def test_processes_without_closure_pass(process):
    # RreaProcessRequiresConservationClosure: transport and seeding carry none
    g = gc.group(process, mode=gc.MISSING)
    assert oracle(g)[0] == ALL


def test_energy_tolerance_is_64_eps_of_the_larger_side():
    g = copy.deepcopy(gc.VALID["compton"])
    # input: w (E_gamma) + w m_e (the target electron); output: w E' + w (T + m_e)
    scale = gc.W * (1.0e6 + gc.ME)
    tol = 64.0 * gc.EPS * scale
    g["local_energy_eV"] = 0.9 * tol
    assert oracle(g)[0] == ALL
    g["local_energy_eV"] = 1.1 * tol
    assert oracle(g)[0] == 1 | 2 | 4


def test_charge_tolerance_counts_escaped_charge_in_the_output_side():
    # electron escape, w = 2: throughput |-2| + |0 + (-2)| = 4, tolerance 256 eps;
    # a rule that left the escaped charge out (throughput 2) would reject this
    g = copy.deepcopy(gc.VALID["electron_escape"])
    g["escaped_charge_e"] = -gc.W - 200.0 * gc.EPS
    bits, dq, _ = oracle(g)
    assert abs(dq) > 128.0 * gc.EPS and bits == ALL, (bits, dq)
    g["escaped_charge_e"] = -gc.W - 300.0 * gc.EPS
    assert oracle(g)[0] == 1 | 2 | 4


def test_sub_cutoff_ionization_returns_the_electron_to_the_fluid():
    # a Compton recoil below the cutoff is deposited (kinetic energy to the
    # local term) and its ionization source emits the electron into the fluid
    g = gc.group(gc.COMPTON, gc.G, w=gc.W, before=1.0e6, after=1.0e6 - 500.0, local=500.0,
                 sources=[(gc.ION, gc.W, 500.0)])
    assert oracle(g)[0] == ALL
    g["sources"] = [(gc.ION, gc.W, 500.0)]
    assert oracle(g, cut=400.0)[0] == 1 | 2 | 4  # above that cutoff the electron must be a particle


@pytest.mark.parametrize("name", sorted(k for k, g in gc.VALID.items() if g["closure_mode"] == gc.DETAILED))
def test_dropping_any_term_breaks_the_closure(name):
    g0 = gc.VALID[name]
    variants = []
    for i in range(len(g0["particles"])):
        g = copy.deepcopy(g0)
        del g["particles"][i]
        variants.append(("particle", i, g))
    for i in range(len(g0["sources"])):
        g = copy.deepcopy(g0)
        del g["sources"][i]
        variants.append(("source", i, g))
    for k in ("local_energy_eV", "escaped_energy_eV", "escaped_charge_e"):
        if g0[k] != 0.0:
            g = copy.deepcopy(g0)
            g[k] = 0.0
            variants.append((k, 0, g))
    for i, (s, w, e) in enumerate(g0["particles"]):
        for other in {0, 1, 2} - {s}:
            g = copy.deepcopy(g0)
            g["particles"][i] = (other, w, e)
            variants.append((f"species {s}->{other}", i, g))
    assert variants
    for what, i, g in variants:
        bits, dq, de = oracle(g)
        assert bits == 1 | 2 | 4, (name, what, i, bits, dq, de)


def test_random_closed_groups_on_either_side_of_the_tolerance():
    # closed in exact arithmetic: valid; a planted energy excess of 1-40 ulps
    # of the input side stays inside the 64-ulp tolerance, 100-400 does not
    rng = np.random.default_rng(5)
    counts = {"closed": 0, "inside": 0, "outside": 0}
    for g, ulps in gc.random_groups(rng, 4000):
        bits, dq, de = oracle(g)
        if ulps == 0 or ulps <= 40:
            counts["closed" if ulps == 0 else "inside"] += 1
            assert bits == ALL, (g, ulps, dq, de)
        else:
            counts["outside"] += 1
            assert bits == 1 | 2 | 4, (g, ulps, dq, de)
            assert de > 0.0
    assert min(counts.values()) > 400, counts
