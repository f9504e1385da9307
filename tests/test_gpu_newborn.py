"""GPU parity of the newborn residual advance (rrea_newborn_drain) against the
oracle's transient FIFO from identical staged secondaries: the FIFO's order,
lineages, generations, parents and statuses exact; payloads, survivors and
ion pairs to 1e-12 plus the residual chords' round-trip conditioning;
newborn counts exact, other tallies to 1e-9; the J grid as the step's."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding as ob  # noqa: E402
from parity import compare_drain  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def oracle_staged(otables, fields, cfg, n):
    """The oracle's ordered step outputs of a C2 sample, in Apply's order."""
    e = wl.c2_electrons(n, fields, seed=2).as_dict()
    g = wl.c2_photons(n // 2, fields, dt_s=cfg.dt_s, seed=4, first_id=1 << 41)
    gp = g.as_dict()
    gp.update(g.extra)
    p = wl.c2_electrons(n // 4, fields, seed=3, first_id=1 << 40, charge_sign=1.0).as_dict()
    staged = []
    for name, t in (("electron", e), ("photon", gp), ("positron", p)):
        if name == "photon":
            o = ob.photon_step(otables, fields, cfg, t, nthreads=8, sec_cap=1 << 18)
        else:
            o = ob.charged_step(name, otables, fields, cfg, t, nthreads=8, sec_cap=1 << 18)
        assert o["code"] == 0
        staged.append(o["secondaries"])
    return np.concatenate(staged)


def gpu_drain(tables, fields, cfg, initial, capacity):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DrainOutputs, RreaParticleInteraction,
                                                                StepOutputs)
    grid = StepOutputs(1, fields, secondary_capacity=1, per_particle=False)
    out = DrainOutputs(capacity, grid)
    RreaParticleInteraction.drain_secondaries(tables, Coupling(fields, tables), cfg, initial, out, check=False)
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("dt,n", [(1e-9, 1 << 14), (1e-8, 1 << 12), (1e-7, 1500)])
def test_drain_parity(tables, otables, dt, n):
    """From one generation of mostly surviving Moller electrons at 1 ns to
    five generations of Moller cascades, brems photons, annihilations and
    demotions at 100 ns."""
    fields = wl.c2_fields()
    cfg = StepConfig(rng_seed=61, dt_s=dt)
    ini = oracle_staged(otables, fields, cfg, n)
    print(f"case drain dt={dt}: {len(ini)} staged secondaries", flush=True)
    o = ob.newborn_drain(otables, fields, cfg, ini, cap=1 << 16)
    assert o["code"] == 0
    out = gpu_drain(tables, fields, cfg, ini, 1 << 16)
    out.grid.check_error()
    stats = {}
    compare_drain(fields, cfg, o, out, label=f"drain dt={dt}", stats=stats)
    gens = int(o["work"]["generation"].max())
    print(f"drain dt={dt}: {stats['entries']} FIFO entries over {gens + 1} generations, {stats['survivors']} "
          f"survivors, {stats['within_1e-12']} within 1e-12 of E; max |dE|/E = {stats['max_rel_e']:.3e}", flush=True)


def test_drain_after_gpu_steps(tables, otables):
    """The whole Apply of one step on the device -- electron, photon and
    positron steps, then the drain of their staged outputs -- against the
    oracle doing the same."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DeviceTile, DrainOutputs, PhotonTile,
                                                                RreaParticleInteraction, StepOutputs)
    fields = wl.c2_fields()
    cfg = StepConfig(rng_seed=62, dt_s=1e-9)
    n = 1 << 13
    coupling = Coupling(fields, tables)
    e = wl.c2_electrons(n, fields, seed=2)
    g = wl.c2_photons(n // 2, fields, dt_s=cfg.dt_s, seed=4, first_id=1 << 41)
    p = wl.c2_electrons(n // 4, fields, seed=3, first_id=1 << 40, charge_sign=1.0)
    out_e = StepOutputs(n, fields, secondary_capacity=1 << 14)
    out_g = StepOutputs(n // 2, fields, secondary_capacity=1 << 14, share_grid=out_e)
    out_p = StepOutputs(n // 4, fields, secondary_capacity=1 << 14, share_grid=out_e)
    RreaParticleInteraction.apply_electrons(tables, coupling, DeviceTile.from_host(e), cfg, out_e)
    RreaParticleInteraction.apply_photons(tables, coupling, PhotonTile.from_host(g), cfg, out_g)
    RreaParticleInteraction.apply_positrons(tables, coupling, DeviceTile.from_host(p), cfg, out_p)
    drain = DrainOutputs(1 << 16, out_e)
    RreaParticleInteraction.drain_secondaries(tables, coupling, cfg, [out_e, out_g, out_p], drain)
    torch.cuda.synchronize()
    ini = oracle_staged(otables, fields, cfg, n)
    o = ob.newborn_drain(otables, fields, cfg, ini, cap=1 << 16)
    w = drain.work_host()
    assert len(w) == len(o["work"])
    for f in ("parent_index", "species", "process", "root_key", "lineage", "root_species", "generation"):
        assert np.array_equal(w[f], o["work"][f]), f
    assert np.array_equal(drain.status[: len(w)].cpu().numpy(), o["status"].astype(np.uint8))
    s = drain.survivors_host()
    assert len(s) == len(o["survivors"])
    assert np.allclose(s["energy_eV"], o["survivors"]["energy_eV"], rtol=1e-10)
    print(f"apply + drain: {len(w)} FIFO entries, {len(s)} survivors", flush=True)


def test_drain_edges_and_determinism(tables, otables):
    """An empty FIFO and a non-positive dt drain nothing; a record already at
    the end of the step survives untouched; an invalid birth payload and an
    exhausted FIFO capacity are reported; two drains are bit-identical."""
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    fields = wl.c2_fields()
    cfg = StepConfig(rng_seed=63, dt_s=1e-7)
    ini = oracle_staged(otables, fields, cfg, 600)
    empty = gpu_drain(tables, fields, cfg, ini[:0], 64)
    assert int(empty.work_count.item()) == 0 and int(empty.survivor_count.item()) == 0
    zero_dt = gpu_drain(tables, fields, StepConfig(rng_seed=63, dt_s=0.0), ini, 1 << 14)
    assert int(zero_dt.work_count.item()) == 0
    done = ini[:3].copy()
    done["birth_fraction"] = 1.0
    kept = gpu_drain(tables, fields, cfg, done, 64)
    assert (kept.survivors_host() == done).all()
    bad = ini[:3].copy()
    bad["birth_fraction"][1] = 1.5
    b = gpu_drain(tables, fields, cfg, bad, 64)
    assert b.grid.error.cpu().tolist()[0] == 4
    small = gpu_drain(tables, fields, cfg, ini, len(ini) + 1)
    assert small.grid.error.cpu().tolist()[0] == 5
    with pytest.raises(_lib.RreaHipError):
        small.grid.check_error()
    a1 = gpu_drain(tables, fields, cfg, ini, 1 << 14)
    a2 = gpu_drain(tables, fields, cfg, ini, 1 << 14)
    n = int(a1.work_count.item())
    assert n > len(ini)
    assert torch.equal(a1.work[: n * 112], a2.work[: n * 112])
    assert torch.equal(a1.survivors, a2.survivors)
    # FP64 atomics may add the J contributions in another order
    assert torch.allclose(a1.grid.jz, a2.grid.jz, rtol=1e-12, atol=1e-9 * float(a1.grid.jz.abs().max()))
