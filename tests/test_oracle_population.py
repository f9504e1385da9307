"""The population-control fixtures (tests/golden/resample.npz, written by the
reference's own thinning plan compiled from the checkout) checked by what
the scheme guarantees, independently of any implementation: every support's
four deposited CIC charges are unchanged, weights stay non-negative, the
removed count is the number of zeroed weights and meets the nominal request,
nothing is removed below the 1.25 x target band, and the request is
min(count - target, capacity).  Where the checkout is present, the fixtures
are regenerated and must be reproduced exactly."""
import os

import numpy as np
import pytest

from golden import make_resample_fixtures as mk

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "resample.npz"))


def case_arrays(case):
    return {k[len(f"{case}_"):]: FIX[k] for k in FIX.files if k.startswith(f"{case}_")}


@pytest.mark.parametrize("case", range(len(FIX["names"])))
def test_fixture_invariants(case):
    a = case_arrays(case)
    z_lo, dr, dz = float(FIX["z_lo"]), float(FIX["dr"]), float(FIX["dz"])
    sr, sz, phase, fr, fz = mk.records(a["r"], a["z"], a["theta"], a["ux"], a["uy"], a["uz"], z_lo, dr, dz)
    assert np.all((fr >= 0) & (fr < 1) & (fz >= 0) & (fz < 1))
    capacity, floor, requested, removed = (int(x) for x in a["counts"])
    n, target = len(a["w"]), int(a["target"])
    w0, w1 = a["w"], a["out_w"]
    assert capacity + floor == n
    due = n > target + target // 4
    assert requested == (min(n - target, capacity) if due else 0)
    assert np.all(w1 >= 0.0)
    assert int((w1 == 0.0).sum()) == removed >= requested
    if not due:
        assert np.array_equal(w0, w1)
    key = sr.astype(np.int64) * 1_000_003 + sz
    for k in np.unique(key):
        m = key == k
        cells = lambda w: np.array([np.sum(w[m] * (1 - fr[m]) * (1 - fz[m])), np.sum(w[m] * fr[m] * (1 - fz[m])),  # noqa: E731
                                    np.sum(w[m] * (1 - fr[m]) * fz[m]), np.sum(w[m] * fr[m] * fz[m])])
        c0, c1 = cells(w0), cells(w1)
        assert np.allclose(c1, c0, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(c0).max())), k
        # at most four records survive where a support was thinned to its floor
        assert int((w1[m] > 0).sum()) >= min(4, int((w0[m] > 0).sum())) or int((w1[m] > 0).sum()) > 0


def test_phase_bins_follow_the_octave_and_signs():
    """phase = (ilogb|u| + 1024) * 8 + [u_r >= 0] + 2 [u_theta >= 0] + 4 [u_z >= 0]."""
    theta = np.array([0.0, 0.0, np.pi / 2, 0.0])
    ux = np.array([3.0, -3.0, 0.0, 0.0])
    uy = np.array([0.0, 0.0, 5.0, 0.0])
    uz = np.array([0.0, 1.0, -1.0, 0.0])
    _, _, ph, _, _ = mk.records(np.full(4, 2.3), np.full(4, 0.4), theta, ux, uy, uz, 0.0, 1.0, 1.0)
    assert ph[0] == (1 + 1024) * 8 + 1 + 2 + 4     # |u| = 3: ilogb 1; u_r > 0, u_theta = 0, u_z = 0
    assert ph[1] == (1 + 1024) * 8 + 0 + 2 + 4     # u_r < 0
    assert ph[2] == (2 + 1024) * 8 + 1 + 2 + 0     # |u| = sqrt(26): ilogb 2; u along theta-hat -> u_r = 0, u_t > 0
    assert ph[3] == 0 * 8 + 1 + 2 + 4              # |u| = 0: exponent clamped to -1024


def test_fixtures_reproduce_from_reference(tmp_path):
    from oracle import binding as ob
    if ob.ref() is None:
        pytest.skip("reference checkout absent")
    import shutil
    import subprocess
    import sys
    script = os.path.join(HERE, "golden", "make_resample_fixtures.py")
    work = tmp_path / "golden"
    work.mkdir()
    shutil.copy(script, work / "make_resample_fixtures.py")
    env = dict(os.environ, PYTHONPATH=os.path.dirname(HERE))
    # run the copy so it writes next to itself, then compare with the committed file
    src = open(work / "make_resample_fixtures.py").read().replace(
        'sys.path.insert(0, os.path.join(HERE, "..", ".."))', f'sys.path.insert(0, {os.path.dirname(HERE)!r})')
    (work / "make_resample_fixtures.py").write_text(src)
    subprocess.check_call([sys.executable, str(work / "make_resample_fixtures.py")], env=env,
                          stdout=subprocess.DEVNULL)
    new = np.load(work / "resample.npz")
    assert sorted(new.files) == sorted(FIX.files)
    for k in FIX.files:
        assert np.array_equal(new[k], FIX[k]), k
