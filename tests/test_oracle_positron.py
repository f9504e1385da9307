"""Pins the oracle's positron physics against what the mathematics fixes and
against the reference's own sampler checks
(native_amrex_rrea/src/RreaScatteringSamplerSmoke.cpp:121-172, 437-449,
579-655), with independent quadratures in place of the formulas under test.
CPU only."""
import numpy as np
import pytest
from scipy import integrate

from oracle import binding as ob
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile, StepConfig

ME = 510998.95


def heitler_density(gamma):
    """Two-photon annihilation in flight, Heitler's spectrum in the photon
    energy fraction e (Geant4 Physics Reference Manual, e+e- -> 2 gamma):
    (1/e) (1 + 2g/(g+1)^2 - e - 1/((g+1)^2 e)) on [(1-b)/2, (1+b)/2],
    b = sqrt((g-1)/(g+1))."""
    gp1 = gamma + 1.0
    b = np.sqrt((gamma - 1.0) / gp1)
    lo, hi = 0.5 * (1 - b), 0.5 * (1 + b)

    def f(e):
        return (1.0 + 2.0 * gamma / gp1 ** 2 - e - 1.0 / (gp1 ** 2 * e)) / e
    return f, lo, hi


@pytest.mark.parametrize("k_mev", [0.3, 1.0, 5.0, 50.0])
def test_annihilation_inverse_cdf_against_quadrature(k_mev):
    """eps(u) is the inverse of the Heitler CDF: integrating the density from
    the lower support edge to eps(u) returns u (a dropped term or a wrong
    coefficient in the antiderivative fails this)."""
    k = k_mev * 1e6
    gamma = 1.0 + k / ME
    f, lo, hi = heitler_density(gamma)
    total = integrate.quad(f, lo, hi, epsabs=0, epsrel=1e-13, limit=200)[0]
    for u in (0.013, 0.1, 0.37, 0.5, 0.81, 0.97):
        s = ob.sample_annihilation(k, u, 0.25)
        eps = s["e1"] / (2 * ME + k)
        assert lo <= eps <= hi
        got = integrate.quad(f, lo, eps, epsabs=0, epsrel=1e-13, limit=200)[0] / total
        assert abs(got - u) < 1e-9, (k_mev, u, got)


@pytest.mark.parametrize("k_mev", [0.3, 1.0, 5.0])
def test_annihilation_in_flight_closure_and_kinematics(k_mev):
    """Energy closure, momentum closure |p - E1 n1| = E2 (the photon-1 cosine
    follows from two-body kinematics), support, and the mean fraction against
    the Heitler quadrature (RreaScatteringSamplerSmoke.cpp:579-619)."""
    k = k_mev * 1e6
    gamma = 1.0 + k / ME
    f, lo, hi = heitler_density(gamma)
    total = integrate.quad(f, lo, hi, epsrel=1e-12)[0]
    mean_ref = integrate.quad(lambda e: e * f(e), lo, hi, epsrel=1e-12)[0] / total
    e_tot = k + 2 * ME
    p = np.sqrt(k * (k + 2 * ME))
    eps = []
    for i in range(601):
        s = ob.sample_annihilation(k, (i + 0.5) / 601, 0.3)
        assert s["valid"] and not s["at_rest"]
        assert abs((s["e1"] + s["e2"]) / e_tot - 1.0) <= 5e-9
        lhs = p * p - 2 * p * s["e1"] * s["cos1"] + s["e1"] ** 2
        assert abs(np.sqrt(lhs) - s["e2"]) <= 1e-9 * e_tot
        eps.append(s["e1"] / e_tot)
    eps = np.array(eps)
    assert eps.min() >= lo - 1e-15 and eps.max() <= hi + 1e-15
    assert abs(eps.mean() - mean_ref) / mean_ref < 0.01


def test_annihilation_at_rest():
    """Two 511 keV photons; the polar draw maps onto an isotropic cosine
    (RreaScatteringSamplerSmoke.cpp:621-655)."""
    s = ob.sample_annihilation(0.0, 0.5, 0.25)
    assert s["at_rest"] and s["e1"] == ME and s["e2"] == ME
    cos = np.array([ob.sample_annihilation(0.0, (i + 0.5) / 1000, 0.25)["cos1"] for i in range(1000)])
    assert np.allclose(cos, 2 * (np.arange(1000) + 0.5) / 1000 - 1, rtol=0, atol=1e-15)


def bhabha_reference_cdf(k, thr):
    """Full Bhabha DCS in the transfer fraction, normalised over [thr/K, 1]
    (RreaScatteringSamplerSmoke.cpp:121-172), by adaptive quadrature."""
    tau = k / ME
    gamma = tau + 1.0
    beta2 = 1.0 - 1.0 / gamma ** 2
    y = 1.0 / (gamma + 1.0)
    one2y = 1.0 - 2.0 * y
    b1 = 2.0 - y * y
    b2 = one2y * (3.0 + y * y)
    b3 = one2y ** 2 + one2y ** 3
    b4 = one2y ** 3

    def dcs(e):
        return max(1.0 / (beta2 * e * e) - b1 / e + b2 - b3 * e + b4 * e * e, 0.0)
    ec = thr / k
    total = integrate.quad(dcs, ec, 1.0, limit=400, epsrel=1e-10)[0]
    return lambda e: integrate.quad(dcs, ec, e, limit=400, epsrel=1e-10)[0] / total


@pytest.mark.parametrize("k_mev", [0.3, 1.0, 5.0, 20.0, 100.0])
def test_bhabha_sampler_against_full_dcs(otables, k_mev):
    """The positron hard-Bhabha CDF reproduces the full Bhabha cross section
    to 0.01 in u (the reference's own tolerance) and spans [T_cut, K]."""
    thr = float(otables.scalars[0])
    k = k_mev * 1e6
    cdf = bhabha_reference_cdf(k, thr)
    worst = 0.0
    for i in range(41):
        u = (i + 0.5) / 41
        t = otables.sample_bhabha(k, u)
        assert thr * (1 - 1e-12) <= t <= k * (1 + 1e-12)
        worst = max(worst, abs(u - cdf(t / k)))
    assert worst <= 0.01, worst


def test_bhabha_and_positron_brems_hit_table_nodes(otables, bundle):
    """At a tabulated primary and a tabulated CDF abscissa the samplers return
    the tabulated secondary energy exactly (the bundle's own vectors)."""
    for key, fn in (("bhabha", otables.sample_bhabha), ("pbrems", otables.sample_positron_brems_energy)):
        prim = bundle[key + "_primary"]
        start = bundle[key + "_start"]
        for g in (1, len(prim) // 2, len(prim) - 2):
            u = bundle[key + "_u"][start[g]:start[g + 1]]
            v = bundle[key + "_v"][start[g]:start[g + 1]]
            for j in (1, len(u) // 3, len(u) - 2):
                assert fn(float(prim[g]), float(u[j])) == v[j], (key, g, j)


def test_positron_brems_spectrum_bounds(otables):
    for k in (1e5, 1e6, 1e7, 1e8):
        vals = np.array([otables.sample_positron_brems_energy(k, (i + 0.5) / 200) for i in range(200)])
        assert vals.min() >= 0.0 and vals.max() <= k * (1 + 1e-6)
        assert np.all(np.diff(vals) >= 0.0)  # inverse CDF is monotone in u


def test_positron_at_rest_step_known_answer(otables):
    """Sub-cutoff positrons annihilate where they stand: status annihilated,
    photon 1 on the cone cos = 2 u_80 - 1 about the momentum with u_80 drawn
    from the particle's key at interaction index 0, photon 2 opposite, the
    residual kinetic energy as K / W_air ion pairs, one positive ion each,
    and no net current (RreaPositronTransport.cpp:129-170)."""
    fields = wl.c2_fields()
    n = 64
    rng = np.random.default_rng(9)
    k = rng.uniform(1e3, 9e3, n)
    tile = wl.c2_electrons(n, fields, seed=12, charge_sign=1.0)
    u = np.sqrt(tile.ux ** 2 + tile.uy ** 2 + tile.uz ** 2)
    s_ = wl.proper_velocity(k) / u
    tile.ux, tile.uy, tile.uz = tile.ux * s_, tile.uy * s_, tile.uz * s_
    tile.cpu = np.arange(n, dtype=np.int32)
    cfg = StepConfig(rng_seed=41, step=3, dt_s=1e-9)
    o = ob.positron_step(otables, fields, cfg, tile.as_dict())
    assert o["code"] == 0 and np.all(o["status"] == 6)
    sec = o["secondaries"]
    assert len(sec) == 2 * n and np.all(sec["process"] == 9) and np.all(sec["energy_eV"] == ME)
    L = ob.lib()
    for i in range(n):
        key = int(tile.id[i]) ^ (int(tile.cpu[i]) << 40)
        u80 = L.orc_rng_uniform(41, key, 3, 0, 80)
        a = np.array([tile.ux[i], tile.uy[i], tile.uz[i]])
        a /= np.linalg.norm(a)
        d1 = np.array([sec["dir_x"][2 * i], sec["dir_y"][2 * i], sec["dir_z"][2 * i]])
        d2 = np.array([sec["dir_x"][2 * i + 1], sec["dir_y"][2 * i + 1], sec["dir_z"][2 * i + 1]])
        assert abs(d1 @ a - (2 * u80 - 1)) < 1e-12
        assert np.array_equal(d2, -d1)
        assert sec["x"][2 * i] == tile.r[i] * np.cos(tile.theta[i])
    assert np.allclose(o["ion_pairs"], k / 33.97, rtol=1e-12)
    assert o["tallies"][10] == n and o["tallies"][12] == n
    assert np.abs(o["jr"]).max() == 0.0 and np.abs(o["jz"]).max() == 0.0


def test_positron_charge_sign_on_the_net_chord(otables):
    """A surviving positron deposits +w on its net chord where an electron
    of the same state deposits -w: with no events and the same chord the two
    z-currents are equal and opposite (field-free, thin air, 1 ps)."""
    from rrea_feedback_warpx_amrex_geant4_amd.config import RzFields
    fields = RzFields.uniform(32, 64, 10.0, 10.0, density_ratio=1e-6)
    k = np.full(4, 5e6)
    u = wl.proper_velocity(k)
    tile = ParticleTile(r=np.array([50.0, 100, 150, 200]), z=np.array([100.0, 200, 300, 400]), theta=np.zeros(4),
                        ux=np.zeros(4), uy=np.zeros(4), uz=u, w=np.ones(4), id=np.arange(1, 5),
                        cpu=np.zeros(4, np.int32))
    cfg = StepConfig(rng_seed=1, dt_s=1e-12)
    oe = ob.electron_step(otables, fields, cfg, tile.as_dict())
    op = ob.positron_step(otables, fields, cfg, tile.as_dict())
    assert np.all(oe["status"] == 0) and np.all(op["status"] == 0)
    assert np.array_equal(oe["z"], op["z"])
    assert np.allclose(oe["jz"], -op["jz"], rtol=1e-15, atol=0)
    assert np.abs(op["jz"]).max() > 0
