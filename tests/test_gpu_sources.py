"""GPU parity of the material sources (rrea_step_out.sources /
rrea_drain_io.sources, rrea_sources_normalize) against the oracle's ledger
replay: the five source grids of the electron, photon and positron steps
and of the newborn drain of a C2 sample, field by field; the normalised
grids; and a batch's per-instance grids against plain per-instance runs."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding as ob  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def compare(label, fields, gpu, orc, rtol=1e-9):
    """Each field to rtol of its largest value: the records' positions agree
    to ~1e-12 relative, and a cell's shape weight moves with position / cell."""
    n = fields.nr * fields.nz
    for f in range(5):
        a, b = gpu[f * n:(f + 1) * n], orc[f * n:(f + 1) * n]
        scale = float(np.abs(b).max())
        if scale == 0.0:
            assert not a.any(), (label, f)
            continue
        err = np.abs(a - b)
        bad = np.flatnonzero(err > rtol * scale)
        assert bad.size == 0, (label, f, bad[:5], a[bad[:5]], b[bad[:5]], scale)


def test_step_and_drain_sources(tables, otables):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DeviceTile, DrainOutputs, PhotonTile,
                                                                RreaParticleInteraction, StepOutputs)
    fields = wl.c2_fields()
    cfg = StepConfig(rng_seed=71, dt_s=1e-9)
    n = 1 << 13
    coupling = Coupling(fields, tables)
    e = wl.c2_electrons(n, fields, seed=2)
    g = wl.c2_photons(n // 2, fields, dt_s=cfg.dt_s, seed=4, first_id=1 << 41)
    p = wl.c2_electrons(n // 4, fields, seed=3, first_id=1 << 40, charge_sign=1.0)
    print(f"case sources: {n} e-, {n // 2} gamma, {n // 4} e+, dt=1e-9", flush=True)
    out_e = StepOutputs(n, fields, secondary_capacity=1 << 14, sources=True)
    out_g = StepOutputs(n // 2, fields, secondary_capacity=1 << 14, share_grid=out_e)
    out_p = StepOutputs(n // 4, fields, secondary_capacity=1 << 14, share_grid=out_e)
    RreaParticleInteraction.apply_electrons(tables, coupling, DeviceTile.from_host(e), cfg, out_e)
    RreaParticleInteraction.apply_photons(tables, coupling, PhotonTile.from_host(g), cfg, out_g)
    RreaParticleInteraction.apply_positrons(tables, coupling, DeviceTile.from_host(p), cfg, out_p)
    torch.cuda.synchronize()
    step_src = out_e.sources.cpu().numpy().copy()
    drain = DrainOutputs(1 << 16, out_e)
    RreaParticleInteraction.drain_secondaries(tables, coupling, cfg, [out_e, out_g, out_p], drain)
    torch.cuda.synchronize()
    all_src = out_e.sources.cpu().numpy()
    # the oracle, in Apply's order, into one ledger
    src = ob.new_sources(fields)
    staged = []
    gp = g.as_dict()
    gp.update(g.extra)
    for name, t in (("electron", e.as_dict()), ("photon", gp), ("positron", p.as_dict())):
        if name == "photon":
            o = ob.photon_step(otables, fields, cfg, t, nthreads=8, sec_cap=1 << 16, sources=src)
        else:
            o = ob.charged_step(name, otables, fields, cfg, t, nthreads=8, sec_cap=1 << 16, sources=src)
        assert o["code"] == 0
        staged.append(o["secondaries"])
    compare("steps", fields, step_src, src)
    d = ob.newborn_drain(otables, fields, cfg, np.concatenate(staged), cap=1 << 16, sources=src)
    assert d["code"] == 0
    compare("steps + drain", fields, all_src, src)
    nz = fields.nr * fields.nz
    totals = [float(all_src[f * nz:(f + 1) * nz].sum()) for f in range(5)]
    print("sources: totals " + ", ".join(f"{x:.6g}" for x in totals) + " match the oracle", flush=True)
    # normalisation
    coupling.normalize_sources(out_e.sources, cfg.dt_s)
    ob.sources_normalize(fields, src, cfg.dt_s, True)
    torch.cuda.synchronize()
    compare("normalised", fields, out_e.sources.cpu().numpy(), src)


def test_batched_sources_match_plain(tables):
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Batch, Coupling, DeviceTile, DrainOutputs,
                                                                RreaParticleInteraction, StepOutputs)
    import ctypes as C
    fields, tiles, seeds = wl.c3_batch(4, seed=7, scale=2e-3)
    dt = 1e-9
    couplings = [Coupling(f, tables) for f in fields]
    plain = []
    for f, t, sd, cp in zip(fields, tiles, seeds, couplings):
        cfg = StepConfig(rng_seed=sd, dt_s=dt)
        out = StepOutputs(len(t.r), f, secondary_capacity=1 << 14, sources=True)
        RreaParticleInteraction.apply_electrons(tables, cp, DeviceTile.from_host(t), cfg, out)
        d = DrainOutputs(1 << 15, out)
        RreaParticleInteraction.drain_secondaries(tables, cp, cfg, [out], d)
        plain.append(out.sources)
    grids = [StepOutputs(1, f, secondary_capacity=1, per_particle=False, sources=True) for f in fields]
    batch = Batch(couplings, grids, seeds)
    arr = (C.c_void_p * len(grids))(*[gr.sources.data_ptr() for gr in grids])
    _lib.check(_lib.lib().rrea_batch_bind_sources(batch.h, arr), "bind sources")
    KEYS = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
    host = type(tiles[0])(*(np.concatenate([getattr(t, k) for t in tiles]) for k in KEYS))
    sizes = [len(t.r) for t in tiles]
    inst = torch.as_tensor(np.repeat(np.arange(len(tiles), dtype=np.int32), sizes), device="cuda")
    out = StepOutputs(len(host.r), fields[0], secondary_capacity=1 << 15)
    cfg = StepConfig(rng_seed=0, dt_s=dt)
    batch.step(tables, "electron", DeviceTile.from_host(host), inst, cfg, out)
    bd = DrainOutputs(1 << 16, out)
    batch.drain(tables, cfg, [out], bd)
    torch.cuda.synchronize()
    for i, (gr, pl) in enumerate(zip(grids, plain)):
        a, b = gr.sources, pl
        assert float(b.abs().max()) > 0.0
        assert torch.allclose(a, b, rtol=0.0, atol=1e-12 * float(b.abs().max())), i
    batch.close()
