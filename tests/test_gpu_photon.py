"""GPU parity of the HIP photon step against the CPU oracle: status, event
counts and secondary order exact; positions, u, secondary payloads and ion
pairs to 1e-12 (lengths along the pushed chord with its round-trip
conditioning); tallies to 1e-10."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from parity import run_photons_and_compare  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile, StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def coupling_for(fields, tables):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling
    return Coupling(fields, tables)


@pytest.mark.parametrize("dt,n,steps", [(1e-12, 1 << 15, 3), (1e-8, 1 << 14, 2), (1e-7, 1 << 13, 2),
                                        (1e-6, 1 << 12, 1)])
def test_c2_photon_parity(tables, otables, dt, n, steps):
    """From no interaction at 1 ps to several Compton continuations, pair and
    photoelectric absorption, and boundary escapes at 1 us."""
    fields = wl.c2_fields()
    tile = wl.c2_photons(n, fields, dt_s=dt, seed=41)
    print(f"case c2 {n} photons dt={dt} x {steps} steps", flush=True)
    run_photons_and_compare(otables, tables, coupling_for(fields, tables), fields, tile,
                            StepConfig(rng_seed=51, dt_s=dt), steps=steps, label=f"photon dt={dt}")


def test_photon_edge_cases(tables, otables):
    """Below the photon cutoff, zero weight, an invalid id, exits through
    r_hi / z_hi / z_lo, the axis, and the top of the photon tables."""
    fields = wl.c2_fields()
    k = np.array([3e4, 1e6, 1e6, 5e6, 2e6, 1e6, 9e10, 1e6, 6e4])
    n = len(k)
    dirs = np.array([[0, 0, 1], [0, 0, 1], [1, 0, 0], [1, 0, 0], [0, 0, 1], [0, 0, -1], [0.6, 0, 0.8],
                     [0, 1, 0], [0, 0, -1]], dtype=float)
    r = np.array([10.0, 20.0, 30.0, fields.r_hi - 1e-3, 0.0, 0.0, 100.0, 5.0, 40.0])
    z = np.array([100.0, 200.0, 300.0, 400.0, fields.z_hi - 1e-3, 1e-3, 500.0, 600.0, 700.0])
    w = np.ones(n)
    w[1] = 0.0
    ids = np.arange(1, n + 1, dtype=np.int64)
    ids[7] = np.iinfo(np.int64).min
    dt = 1e-6
    theta = np.linspace(-3, 3, n)
    ux, uy, uz = k * dirs[:, 0], k * dirs[:, 1], k * dirs[:, 2]
    rc, thc, zc = wl.push_photons(r, theta, z, ux, uy, uz, dt)
    tile = ParticleTile(r=r, z=z, theta=theta, ux=ux, uy=uy, uz=uz, w=w, id=ids, cpu=np.arange(n, dtype=np.int32))
    tile.extra = {"r_cur": rc, "theta_cur": thc, "z_cur": zc}
    print("case photon edge 9 photons dt=1us", flush=True)
    run_photons_and_compare(otables, tables, coupling_for(fields, tables), fields, tile,
                            StepConfig(rng_seed=23, dt_s=dt), steps=1, label="photon edge")


def test_photon_push_and_determinism(tables):
    """The device push moves every valid photon by c dt; two runs of the same
    step give identical photons and secondaries."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import PhotonTile, RreaParticleInteraction, StepOutputs
    fields = wl.c2_fields()
    tile = wl.c2_photons(1 << 15, fields, dt_s=1e-7, seed=42)
    # start from the unpushed point: the device push makes it the captured one
    tile.extra = {"r_cur": tile.r.copy(), "theta_cur": tile.theta.copy(), "z_cur": tile.z.copy()}
    cp = coupling_for(fields, tables)
    results = []
    for _ in range(2):
        dev = PhotonTile.from_host(tile)
        dev.push(1e-7)
        out = StepOutputs(len(tile), fields, 1 << 17)
        RreaParticleInteraction.apply_photons(tables, cp, dev, StepConfig(rng_seed=9, dt_s=1e-7), out)
        torch.cuda.synchronize()
        results.append((dev.cur.r.cpu().numpy(), dev.cur.uz.cpu().numpy(), out.secondaries_host()))
    assert np.array_equal(results[0][0], results[1][0]) and np.array_equal(results[0][1], results[1][1])
    assert results[0][2].tobytes() == results[1][2].tobytes()
    # the push: prev is the old pushed point, the new one is c dt further along u
    dev = PhotonTile.from_host(tile)
    dev.push(1e-7)
    torch.cuda.synchronize()
    pr, pth, pz = (t.cpu().numpy() for t in (dev.prev_r, dev.prev_theta, dev.prev_z))
    assert np.array_equal(pr, tile.r) and np.array_equal(pz, tile.z)
    r, th, z = (t.cpu().numpy() for t in (dev.cur.r, dev.cur.theta, dev.cur.z))
    step = np.sqrt((r * np.cos(th) - pr * np.cos(pth)) ** 2 + (r * np.sin(th) - pr * np.sin(pth)) ** 2 + (z - pz) ** 2)
    assert np.allclose(step, 299792458.0 * 1e-7, rtol=1e-9)


@pytest.mark.parametrize("dt", [1e-12, 1e-8])
def test_photon_fast_path_matches_complete_path(tables, otables, dt):
    """The fast path (no interaction on the chord) and the complete step give
    bit-identical results on the same photons; the complete path alone also
    matches the oracle."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import PhotonTile, RreaParticleInteraction, StepOutputs
    fields = wl.c2_fields()
    cp = coupling_for(fields, tables)
    tile = wl.c2_photons(1 << 14, fields, dt_s=dt, seed=43)
    results = []
    for complete in (False, True):
        cfg = StepConfig(rng_seed=57, dt_s=dt, complete_path_only=complete)
        dev = PhotonTile.from_host(tile)
        out = StepOutputs(len(tile), fields, secondary_capacity=1 << 14)
        RreaParticleInteraction.apply_photons(tables, cp, dev, cfg, out)
        torch.cuda.synchronize()
        results.append((dev, out))
    (a, oa), (b, ob_) = results
    for k in ("r", "z", "theta", "ux", "uy", "uz", "id"):
        assert torch.equal(getattr(a.cur, k), getattr(b.cur, k)), k
    assert torch.equal(oa.status, ob_.status) and torch.equal(oa.substeps, ob_.substeps)
    assert torch.equal(oa.ion_pairs, ob_.ion_pairs)
    sa, sb = oa.secondaries_host(), ob_.secondaries_host()
    assert len(sa) == len(sb) and (sa == sb).all()
    assert torch.allclose(oa.tallies, ob_.tallies, rtol=1e-12, atol=0.0)
    print(f"photon fast == complete at dt={dt}: {int((oa.status == 0).sum())} survivors, {len(sa)} secondaries",
          flush=True)
    run_photons_and_compare(otables, tables, cp, fields, tile,
                            StepConfig(rng_seed=57, dt_s=dt, complete_path_only=True), steps=1,
                            label=f"photon complete dt={dt}")
