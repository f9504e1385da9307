"""One step of the full at-scale tile against the oracle: 2^24 macroparticles
(11,744,051 e-, 4,194,304 gamma, 838,861 e+) on the 256 x 1024 RZ mesh at
dt = 1 ps, each tile cell-ordered on the device first (rrea_tile_sort, as the
benchmark keeps them), so the fast kernels' LDS deposit window, the hand-over
lists and the atomics run at their real size and contention.  The oracle
steps the same ordered tiles; the comparison is the plain parity bar
(tests/parity.py).  The share of live particle-steps within 1e-12 of |u|
outright is printed."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding as ob  # noqa: E402
from parity import compare_photon_step, compare_step, host_tile_from_dev  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu

THREADS = 16


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


@pytest.mark.parametrize("species", ["electron", "positron"])
def test_full_c2_charged_tile(tables, otables, species):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DeviceTile, RreaParticleInteraction,
                                                                StepOutputs)
    fields = wl.c2_fields()
    n_e, _, n_p = wl.c2_species_counts(1 << 24)
    n = n_e if species == "electron" else n_p
    host0 = (wl.c2_electrons(n, fields, seed=2) if species == "electron"
             else wl.c2_electrons(n, fields, seed=3, first_id=1 << 40, charge_sign=1.0))
    print(f"case full c2 {species}: {n} particles, one 1 ps step", flush=True)
    coupling = Coupling(fields, tables)
    dev = DeviceTile.from_host(host0)
    dev.sort_by_cell(coupling)
    torch.cuda.synchronize()
    host = host_tile_from_dev(dev)
    del host0
    cfg = StepConfig(rng_seed=20240623, dt_s=1e-12, step=7)
    out = StepOutputs(n, fields, secondary_capacity=1 << 16)
    apply = RreaParticleInteraction.apply_positrons if species == "positron" else RreaParticleInteraction.apply_electrons
    apply(tables, coupling, dev, cfg, out)
    torch.cuda.synchronize()
    o = ob.charged_step(species, otables, fields, cfg, host.as_dict(), nthreads=THREADS, sec_cap=1 << 16)
    assert o["code"] == 0
    stats = {}
    compare_step(fields, host, o, dev, out, label=f"full c2 {species}", dt=cfg.dt_s, stats=stats)
    st = o["status"]
    print(f"full c2 {species}: {n} particles, {int((st == 0).sum())} survive, {len(o['secondaries'])} secondaries, "
          f"{stats['within_1e-12']} of {stats['n']} live particle-steps ({stats['within_1e-12'] / stats['n']:.6f}) "
          f"within 1e-12 of |u| outright, max {stats['max_rel_u']:.3e}", flush=True)


def test_full_c2_photon_tile(tables, otables):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, PhotonTile, RreaParticleInteraction,
                                                                StepOutputs)
    fields = wl.c2_fields()
    _, n, _ = wl.c2_species_counts(1 << 24)
    dt = 1e-12
    print(f"case full c2 photons: {n} photons, one 1 ps step", flush=True)
    tile = wl.c2_photons(n, fields, dt_s=dt, seed=4, first_id=1 << 41)
    coupling = Coupling(fields, tables)
    dev = PhotonTile.from_host(tile)
    dev.sort_by_cell(coupling)
    torch.cuda.synchronize()
    p = {k: getattr(dev.cur, k).cpu().numpy().copy() for k in ("ux", "uy", "uz", "w", "id", "cpu")}
    p.update({"r": dev.prev_r.cpu().numpy().copy(), "z": dev.prev_z.cpu().numpy().copy(),
              "theta": dev.prev_theta.cpu().numpy().copy(), "r_cur": dev.cur.r.cpu().numpy().copy(),
              "z_cur": dev.cur.z.cpu().numpy().copy(), "theta_cur": dev.cur.theta.cpu().numpy().copy()})
    del tile
    cfg = StepConfig(rng_seed=20240623, dt_s=dt, step=7)
    out = StepOutputs(n, fields, secondary_capacity=1 << 16)
    RreaParticleInteraction.apply_photons(tables, coupling, dev, cfg, out)
    torch.cuda.synchronize()
    o = ob.photon_step(otables, fields, cfg, p, nthreads=THREADS, sec_cap=1 << 16)
    assert o["code"] == 0
    compare_photon_step(fields, p, o, dev, out, label="full c2 photons", dt=dt)
    print(f"full c2 photons: {n} photons, {int((o['status'] == 0).sum())} survive, "
          f"{len(o['secondaries'])} secondaries -- equal to the oracle", flush=True)
