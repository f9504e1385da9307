"""The oracle's field side -- the conversion of the net-chord deposit to
physical face currents and the TM Maxwell leapfrog -- pinned bit for bit
against the reference's own header compiled from the checkout (oracle/_ref,
where it builds) and, independently of any implementation, by the discrete
identities the scheme is built on: exact mixed-time energy conservation in a
closed domain, the open-step energy ledger dU_E + dU_B + dt (Joule +
Poynting) = 0, Gauss's law under Ampere's update (div curl = 0), and the
physical divergence of a converted deposit equalling its index-space
divergence over the cell volume."""
import numpy as np
import pytest

from oracle import binding as ob
from oracle import maxwell as M

EPS0 = 8.8541878128e-12


def random_field(rng, nr, nz, dr, dz, scale_e=1.0e3, scale_b=3.0e-6, pin=True):
    f = M.Field(nr, nz, dr, dz)
    f.er[:] = scale_e * rng.standard_normal(f.er.size)
    f.ez[:] = scale_e * rng.standard_normal(f.ez.size)
    f.bt[:] = scale_b * rng.standard_normal(f.bt.size)
    if pin:
        f.er.reshape(nz, nr + 1)[:, 0] = 0.0
        f.bt.reshape(nz + 1, nr + 1)[:, 0] = 0.0
    return f


def padded_index_current(rng, nr, nz, pad=2, density=0.3):
    jr = rng.standard_normal((nr + pad + 1) * (nz + 2 * pad))
    jz = rng.standard_normal((nr + pad) * (nz + 2 * pad + 1))
    jr[rng.random(jr.size) > density] = 0.0
    jz[rng.random(jz.size) > density] = 0.0
    return jr, jz


needs_ref = pytest.mark.skipif(not ob.ref_available(), reason="reference checkout absent")


@needs_ref
@pytest.mark.parametrize("nr,nz", [(1, 1), (3, 7), (16, 12), (40, 64)])
def test_current_conversion_matches_reference(nr, nz):
    rng = np.random.default_rng(nr * 100 + nz)
    a, b = padded_index_current(rng, nr, nz)
    got = M.current_to_physical(nr, nz, 0.7, 1.3, a, b)
    ref = M.current_to_physical(nr, nz, 0.7, 1.3, a, b, use_ref=True)
    for g, r in zip(got, ref):
        assert np.array_equal(g, r)


@needs_ref
@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("extras", ["none", "current", "all"])
def test_maxwell_ops_match_reference(op, extras):
    rng = np.random.default_rng(7 + op)
    nr, nz, dr, dz, dt = 11, 17, 5.0, 7.0, 2.5e-9
    f = random_field(rng, nr, nz, dr, dz, pin=False)
    n_er, n_ez, _ = M.shapes(nr, nz)
    kw = {}
    if extras in ("current", "all"):
        kw.update(jr=0.1 * rng.standard_normal(n_er), jz=0.1 * rng.standard_normal(n_ez))
    if extras == "all":
        kw.update(sigma_r=1e-3 * rng.random(n_er), sigma_z=1e-3 * rng.random(n_ez),
                  amb_er=1e5 * rng.standard_normal(n_er), amb_ez=1e5 * rng.standard_normal(n_ez))
    g = f.copy().apply(op, dt, **kw)
    r = f.copy().apply(op, dt, use_ref=True, **kw)
    for k in ("er", "ez", "bt"):
        assert np.array_equal(getattr(g, k), getattr(r, k)), k


@needs_ref
def test_maxwell_diagnostics_match_reference():
    rng = np.random.default_rng(3)
    nr, nz = 9, 13
    f = random_field(rng, nr, nz, 5.0, 7.0, pin=False)
    n_er, n_ez, _ = M.shapes(nr, nz)
    aer, aez = 1e5 * rng.standard_normal(n_er), 1e5 * rng.standard_normal(n_ez)
    jr, jz = 0.1 * rng.standard_normal(n_er), 0.1 * rng.standard_normal(n_ez)
    ref = f.ref_diagnostics(aer, aez, jr, jz)
    joule, moment = f.joule(jr, jz, aer, aez)
    assert f.energy(aer, aez) == ref[0]
    assert joule == ref[1]
    assert f.poynting(aer, aez) == ref[2]
    assert moment == ref[3]
    prev = 3e-6 * rng.standard_normal(f.bt.size)
    assert f.mixed_magnetic_energy(prev) == ob.ref().ref_maxwell_mixed_magnetic_energy(
        nr, nz, 5.0, 7.0, M._p(f.bt), M._p(prev))


def test_closed_domain_conserves_mixed_time_energy():
    """No absorber, no current, boundary corners at zero: the leapfrog keeps
    eps0/2 sum V_f E^2 + 1/(2 mu0) sum V_c B^{n-1/2} B^{n+1/2} constant.  A
    wrong dual volume, axis term or r-weight breaks this at O(1)."""
    rng = np.random.default_rng(11)
    nr, nz, dr, dz, dt = 12, 16, 5.0, 7.0, 2.5e-9
    f = random_field(rng, nr, nz, dr, dz)
    bt = f.bt.reshape(nz + 1, nr + 1)
    bt[0, :] = bt[-1, :] = 0.0
    bt[:, -1] = 0.0
    energies = []
    for _ in range(200):
        prev = f.bt.copy()
        f.apply(1, dt)
        energies.append(f.electric_energy() + f.mixed_magnetic_energy(prev))
        f.apply(3, dt)
    energies = np.array(energies)
    assert np.max(np.abs(energies - energies[0])) <= 1e-12 * abs(energies[0])
    # a planted mistake in the r-weighting (Ez driven by dB/dr alone) breaks it
    g = random_field(np.random.default_rng(11), nr, nz, dr, dz)
    gb = g.bt.reshape(nz + 1, nr + 1)
    gb[0, :] = gb[-1, :] = 0.0
    gb[:, -1] = 0.0
    bad = []
    for _ in range(50):
        prev = g.bt.copy()
        g.apply(1, dt)
        bad.append(g.electric_energy() + g.mixed_magnetic_energy(prev))
        b2 = g.bt.reshape(nz + 1, nr + 1)
        ez = g.ez.reshape(nz + 1, nr)
        g.apply(3, dt)
        # undo the correct Ez update's curl and apply a plain d/dr instead
        c2dt = dt / (EPS0 * 1.25663706212e-6)
        i = np.arange(nr)
        right = c2dt * ((i + 1) * b2[:, 1:] - i * b2[:, :-1]) / ((i + 0.5) * dr)
        wrong = c2dt * (b2[:, 1:] - b2[:, :-1]) / dr
        ez += wrong - right
    bad = np.array(bad)
    assert np.max(np.abs(bad - bad[0])) > 1e-6 * abs(bad[0])


def test_open_step_energy_ledger_closes():
    """One Ampere substep with current, the Silver-Mueller absorber and a
    curl-free ambient: dU_E(ambient-relative) + dU_B(mixed) + dt (Joule +
    Poynting) = 0, Joule and Poynting the mean over E^n and E^{n+1}."""
    rng = np.random.default_rng(5)
    nr, nz, dr, dz, dt = 10, 14, 5.0, 7.0, 2.5e-9
    n_er, n_ez, _ = M.shapes(nr, nz)
    f = random_field(rng, nr, nz, dr, dz)
    # a curl-free ambient: minus the gradient of a potential P on the cell
    # centres (one ghost row below and above, one ghost column past the wall)
    P = rng.standard_normal((nz + 2, nr + 1)) * 1e5
    amb_er = np.zeros((nz, nr + 1))
    amb_er[:, 1:] = -(P[1:nz + 1, 1:] - P[1:nz + 1, :-1]) / dr
    amb_ez = -(P[1:, :nr] - P[:-1, :nr]) / dz
    amb_er, amb_ez = amb_er.ravel(), amb_ez.ravel()
    # its discrete curl vanishes at every interior corner
    ar, az = amb_er.reshape(nz, nr + 1), amb_ez.reshape(nz + 1, nr)
    curl = (az[1:-1, 1:] - az[1:-1, :-1]) / dr - (ar[1:, 1:nr] - ar[:-1, 1:nr]) / dz
    assert np.max(np.abs(curl)) <= 1e-9 * np.max(np.abs(ar)) / dr
    jr = 0.1 * rng.standard_normal(n_er)
    jr.reshape(nz, nr + 1)[:, 0] = 0.0
    jz = 0.1 * rng.standard_normal(n_ez)
    b_before = f.bt.copy()
    f.apply(1, dt).apply(2)
    ue0 = f.electric_energy(amb_er, amb_ez)
    ub0 = f.mixed_magnetic_energy(b_before)
    p0 = f.poynting(amb_er, amb_ez)
    j0, _ = f.joule(jr, jz, amb_er, amb_ez)
    f.apply(3, dt, jr=jr, jz=jz)
    ue1 = f.electric_energy(amb_er, amb_ez)
    p1 = f.poynting(amb_er, amb_ez)
    j1, _ = f.joule(jr, jz, amb_er, amb_ez)
    after = f.copy().apply(1, dt)
    ub1 = after.mixed_magnetic_energy(f.bt)
    joule, poynting = 0.5 * dt * (j0 + j1), 0.5 * dt * (p0 + p1)
    residual = (ue1 - ue0) + (ub1 - ub0) + joule + poynting
    scale = abs(ue1 - ue0) + abs(ub1 - ub0) + abs(joule) + abs(poynting)
    assert abs(joule) > 0 and abs(poynting) > 0
    assert abs(residual) <= 1e-12 * scale, residual / scale


def test_ampere_update_keeps_gauss_law():
    """div(E^{n+1}) - div(E^n) = -(dt/eps0) div J on every cell: the curl of
    B has no divergence in the r-weighted metric."""
    rng = np.random.default_rng(9)
    nr, nz, dr, dz, dt = 13, 9, 3.0, 4.0, 1e-9
    n_er, n_ez, _ = M.shapes(nr, nz)
    f = random_field(rng, nr, nz, dr, dz)
    jr = 0.1 * rng.standard_normal(n_er)
    jz = 0.1 * rng.standard_normal(n_ez)
    d0 = f.divergence(f.er, f.ez)
    f.apply(3, dt, jr=jr, jz=jz)
    d1 = f.divergence(f.er, f.ez)
    want = -(dt / EPS0) * f.divergence(jr, jz)
    scale = np.max(np.abs(d1)) + np.max(np.abs(want))
    assert np.max(np.abs((d1 - d0) - want)) <= 1e-12 * scale


def test_converted_deposit_divergence_is_index_divergence_over_volume():
    """A converted index-space current's physical RZ divergence equals q
    times its index-space divergence over the cell volume, cell by cell:
    the conversion's face areas are the mesh's own metric."""
    rng = np.random.default_rng(13)
    nr, nz, dr, dz, pad, q = 8, 11, 2.0, 3.0, 2, 1.602176634e-19
    a, b = padded_index_current(rng, nr, nz, pad, density=1.0)
    a.reshape(nz + 2 * pad, nr + pad + 1)[:, 0] = 0.0  # the axis face carries no flux
    jr, jz = M.current_to_physical(nr, nz, dr, dz, a, b, pad=pad, q=q)
    f = M.Field(nr, nz, dr, dz)
    div = f.divergence(jr, jz)
    A = a.reshape(nz + 2 * pad, nr + pad + 1)[pad:pad + nz, :nr + 1]
    B = b.reshape(nz + 2 * pad + 1, nr + pad)[pad:pad + nz + 1, :nr]
    idx_div = (A[:, 1:] - A[:, :-1]) + (B[1:, :] - B[:-1, :])
    i = np.arange(nr)
    vol = np.pi * (((i + 1) * dr) ** 2 - (i * dr) ** 2) * dz
    want = q * idx_div / vol
    assert np.allclose(div, want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))
