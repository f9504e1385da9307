"""Pins the CPU oracle (oracle/) against the reference's own golden vectors,
the reference-compiled pins in oracle/_ref, the bundle's validation
holdouts, and closed forms.  CPU only."""
import os

import numpy as np
import pytest

from oracle import binding as ob

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# native_amrex_rrea/src/RreaRngEndpointSmoke.cpp:38-52
RNG_PINS = [
    ((0x0, 0x0, 0x0, 0x0, 0x0), 0x3db5adc86dfdd12d, float.fromhex("0x1.edad6e436feecp-3")),
    ((0x1, 0x2, 0x3, 0x4, 0x5), 0x5d3617c86eb12224, float.fromhex("0x1.74d85f21bac4ap-2")),
    ((0xdeadbeef, 0x2bdc545d6b4b87, 0x2a, 0x7, 0x33), 0xeb74a9e550130b92, float.fromhex("0x1.d6e953caa0261p-1")),
    ((0x7ea, 0x1, 0x186a0, 0x3, 0x63), 0xaceff5afeec79e3c, float.fromhex("0x1.59dfeb5fdd8f3p-1")),
    ((0xffffffffffffffff,) * 5, 0xb22a341e3093f92f, float.fromhex("0x1.6454683c6127fp-1")),
    ((0x75bcd15, 0x3ade68b1, 0x37, 0x0, 0x6b), 0x9633b556ab3a3a1f, float.fromhex("0x1.2c676aad56747p-1")),
]


@pytest.mark.parametrize("key,counter,uniform", RNG_PINS)
def test_rng_golden_pins(key, counter, uniform):
    L = ob.lib()
    assert L.orc_rng_counter(*key) == counter
    assert L.orc_rng_uniform(*key) == uniform


def test_rng_matches_reference_build(ref_available):
    R = ob.ref()
    if R is None:
        pytest.skip("reference checkout absent")
    L = ob.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        key = [int(v) for v in rng.integers(0, 2**63, 5, dtype=np.uint64)]
        assert L.orc_rng_counter(*key) == R.ref_rng_counter(*key)
        assert L.orc_rng_uniform(*key) == R.ref_rng_uniform(*key)
    a = np.zeros(64)
    b = np.zeros(64)
    for u1, u2 in [(0.3, 0.7), (1e-9, 0.999999), (0.5, 0.5)]:
        L.orc_uniform_stream(u1, u2, 64, a.ctypes.data_as(ob.P_D))
        R.ref_uniform_stream(u1, u2, 64, b.ctypes.data_as(ob.P_D))
        assert np.array_equal(a, b)


def test_uniform_midpoint_range():
    # RreaRng.cpp:35-44: midpoint sampling keeps every draw in [2^-53, 1 - 2^-53]
    L = ob.lib()
    vals = np.array([L.orc_rng_uniform(7, i, 3, 1, 107) for i in range(20000)])
    assert vals.min() >= 2.0**-53 and vals.max() <= 1.0 - 2.0**-53
    assert abs(vals.mean() - 0.5) < 0.01


def test_interpolation_smoke_pins(otables):
    # native_amrex_rrea/src/RreaInteractionTablesSmoke.cpp:29-46, restated on a
    # three-node zero-threshold table through the oracle's rate slot 14.
    L = ob.lib()
    h = L.orc_tables_create()
    x = np.array([1.0e3, 1.0e4, 1.0e5])
    y = np.array([0.0, 1.0, 100.0])
    L.orc_tables_set_rate(h, 0, x.ctypes.data_as(ob.P_D), y.ctypes.data_as(ob.P_D), 3, 1)
    import ctypes as C
    err = C.c_int32(0)
    assert L.orc_interpolate(h, 0, 1.0e3, C.byref(err)) == 0.0
    assert L.orc_interpolate(h, 0, 1.0e5, C.byref(err)) == 100.0
    mid = L.orc_interpolate(h, 0, np.sqrt(1.0e4 * 1.0e5), C.byref(err))
    assert abs(mid - 10.0) <= 1e-12 * 10.0
    # zero-threshold bracket is linear in log x: halfway in log between 1e3 and 1e4
    assert abs(L.orc_interpolate(h, 0, np.sqrt(1.0e3 * 1.0e4), C.byref(err)) - 0.5) < 1e-12
    # extrapolation is forbidden
    L.orc_interpolate(h, 0, 2.0e5, C.byref(err))
    assert err.value == 1
    L.orc_tables_destroy(h)


def test_quarterpoint_holdout(otables, bundle):
    """The bundle's own validation holdout: the reference runtime interpolation
    at 58k energies across all 15 rate tables."""
    q = np.load(os.path.join(GOLDEN, "quarterpoints.npz"))
    keys = [str(k) for k in q["keys"]]
    scale = {k: float(np.max(bundle[k + "_y"])) for k in keys}
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(len(q["energy"]), 6000, replace=False),
                          np.arange(len(q["energy"]) - 776, len(q["energy"]))])  # every photon row
    exact = 0
    worst = 0.0
    for i in idx:
        key = keys[q["table"][i]]
        got = otables.interpolate(key, float(q["energy"][i]))
        want = float(q["runtime"][i])
        if got == want:
            exact += 1
        elif abs(want) < 1e-9 * scale[key]:
            # rows a few 1e-8 eV above a zero-rate threshold node: log(q) - log(x_lo)
            # cancels to a handful of bits there, so they pin only to the table scale
            assert abs(got - want) <= 1e-12 * scale[key]
        else:
            # log-log interpolation loses about eps / log(x_hi / x_lo) in t; the
            # bundle clusters nodes 1e-8 apart just above its thresholds
            x = bundle[key + "_x"]
            j = int(np.searchsorted(x, q["energy"][i]))
            cond = 8.0 * np.finfo(float).eps / max(np.log(x[j] / x[j - 1]), 1e-300)
            worst = max(worst, abs(got - want) / abs(want) / (2e-10 + cond))
    assert worst < 1.0, worst
    assert exact > 0.5 * len(idx)


def test_hard_moller_mean_is_inside_kinematic_range(otables, bundle):
    hm = otables.hm_mean()
    x = bundle["e_moller_x"]
    y = bundle["e_moller_y"]
    on = y > 0
    # the mean delta ray lies in [T_cut, K/2]
    assert np.all(hm[on] >= 5.0e4 - 1e-6)
    assert np.all(hm[on] <= 0.5 * x[on] * (1 + 1e-12))
    assert np.all(hm[~on] == 0.0)


def test_screened_rutherford_closed_form(otables, bundle):
    # RreaScatteringSamplerSmoke.cpp:1101-1110: cos = 1 - 2 eta u / (1 + eta - u)
    e = bundle["el_e_e"]
    eta = bundle["el_e_eta"]
    for i in (0, 50, 120, 225):
        for u in (0.1, 0.9, 0.99):
            want = 1.0 - 2.0 * eta[i] * u / (1.0 + eta[i] - u)
            got = otables.sample_single_elastic_cos(False, float(e[i]), u)
            assert abs(got - want) <= 1e-12 * max(1.0, abs(want))


def test_moller_sampler_spans_kinematic_limits(otables):
    # Moller secondary spans [T_cut, K/2] (RreaInteractionTables.cpp:2575-2583)
    for k in (1.0e5 * 1.0001, 1.0e6, 7.2e6, 1.0e9):
        lo = otables.sample_moller(k, 0.0)
        hi = otables.sample_moller(k, 1.0)
        assert abs(lo - 5.0e4) < 1e-6 * 5e4
        assert hi <= 0.5 * k * (1 + 1e-9)
        us = np.linspace(0, 1, 33)
        v = [otables.sample_moller(k, float(u)) for u in us]
        assert np.all(np.diff(v) >= 0)


def test_modified_tsai_against_geant4_rejection(otables):
    """Bounded ModifiedTsai (G4ModifiedTsai::SampleCosTheta): sample the
    published rejection form independently and compare distributions."""
    rng = np.random.default_rng(11)
    k = 2.0e6
    gamma = 1.0 + k / 510998.95
    a, d = 0.625, 27.0
    umax = 2.0 * gamma
    draws = []
    while len(draws) < 20000:
        r1, r2, r3 = rng.random(3)
        uu = -np.log(r1 * r2) / (a if r3 < 9.0 / (9.0 + d) else 3.0 * a)
        if uu <= umax:
            draws.append(1.0 - 2.0 * (uu / umax) ** 2)
    draws = np.sort(draws)
    ours = np.sort([otables.sample_brems_cos(k, float(u)) for u in rng.random(20000)])
    # two-sample KS
    grid = np.linspace(-1, 1, 4001)
    cdf_a = np.searchsorted(draws, grid) / len(draws)
    cdf_b = np.searchsorted(ours, grid) / len(ours)
    assert np.max(np.abs(cdf_a - cdf_b)) < 0.02


def test_gs_sampler_bounds_and_isotropic(otables):
    for e in (1.0e4, 1.0e6, 1.0e8, 1.0e9):
        for n in (10.0, 100.0, 1.0e4):
            v = [otables.sample_gs_cos(False, e, n, float(u)) for u in np.linspace(0, 1, 17)]
            assert all(-1.0 <= c <= 1.0 for c in v)
            assert np.all(np.diff(v) >= -1e-15)  # the cdf runs from backward (-1) to forward
    # n >= isotropic threshold: cos = 2u - 1 exactly
    assert otables.sample_gs_cos(False, 1.0e6, 1.0e5, 0.25) == -0.5


def test_gl2_and_deposit_match_reference_build():
    R = ob.ref()
    if R is None:
        pytest.skip("reference checkout absent")
    rng = np.random.default_rng(3)
    L = ob.lib()
    # GL2 is exact for cubics (RreaGaussLegendre8Smoke)
    c = np.array([0.3, -1.2, 2.5, 0.7])
    a, b = 0.1, 0.9
    exact = sum(c[i] * (b ** (i + 1) - a ** (i + 1)) / (i + 1) for i in range(4))
    assert abs(R.ref_gl2_cubic(a, b, c.ctypes.data_as(ob.P_D)) - exact) < 1e-14
    nr, nz = 12, 16
    for trial in range(300):
        rho1, jr1, jz1 = np.zeros(nr * nz), np.zeros((nr + 1) * nz), np.zeros(nr * (nz + 1))
        rho2, jr2, jz2 = rho1.copy(), jr1.copy(), jz1.copy()
        r0, r1 = rng.uniform(0, 9, 2)
        z0, z1 = rng.uniform(2, 14, 2)
        if trial % 3 == 0:
            r1 = r0 + rng.uniform(-0.3, 0.3)
            z1 = z0 + rng.uniform(-0.3, 0.3)
        q = rng.uniform(-2, 2)
        args = (nr, nz, 1.0, 1.0, 0.0, 0.0)
        L.orc_deposit_segment(*args, *(x.ctypes.data_as(ob.P_D) for x in (rho1, jr1, jz1)), q, 1e-12,
                              abs(r0), z0, abs(r1), z1)
        R.ref_deposit_segment(*args, *(x.ctypes.data_as(ob.P_D) for x in (rho2, jr2, jz2)), q, 1e-12,
                              abs(r0), z0, abs(r1), z1)
        assert np.array_equal(jr1, jr2) and np.array_equal(jz1, jz2)


def test_deposit_continuity_identity():
    """RreaSegmentDepositionRZ.H:4-16: rho(end) - rho(start) = -dt div J in index
    space, at every cell, to round-off."""
    L = ob.lib()
    rng = np.random.default_rng(9)
    nr, nz = 10, 12
    for _ in range(200):
        rho, jr, jz = np.zeros(nr * nz), np.zeros((nr + 1) * nz), np.zeros(nr * (nz + 1))
        r0, r1 = rng.uniform(0.6, 8.5, 2)
        z0, z1 = rng.uniform(1.6, 10.4, 2)
        q, dt = 1.0, 1.0
        L.orc_deposit_segment(nr, nz, 1.0, 1.0, 0.0, 0.0, *(x.ctypes.data_as(ob.P_D) for x in (rho, jr, jz)),
                              q, dt, r0, z0, r1, z1)
        cic0, cic1 = np.zeros(nr * nz), np.zeros(nr * nz)
        for r, z, tgt in ((r0, z0, cic0), (r1, z1, cic1)):
            xr, xz = r - 0.5, z - 0.5
            i0, j0 = int(np.floor(xr)), int(np.floor(xz))
            fr, fz = xr - i0, xz - j0
            for a, wa in ((0, 1 - fr), (1, fr)):
                for bb, wb in ((0, 1 - fz), (1, fz)):
                    i, j = i0 + a, j0 + bb
                    if i < 0:
                        i = -1 - i
                    tgt[i + nr * j] += q * wa * wb
        jr2 = jr.reshape(nz, nr + 1)
        jz2 = jz.reshape(nz + 1, nr)
        div = (jr2[:, 1:] - jr2[:, :-1]) + (jz2[1:, :] - jz2[:-1, :])
        assert np.allclose((cic1 - cic0).reshape(nz, nr), -dt * div, atol=1e-13)
