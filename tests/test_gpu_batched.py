"""GPU checks of the batched execution model (rrea_transport_step_batched,
rrea_newborn_drain_batched): instances of uneven size, field and density
stepped in one launch give each instance exactly what a plain step of that
instance alone gives -- the particle states bit for bit, the staged
secondaries record for record (parent indices into the batched tile), the J
grid to summation order -- and the batched drain matches the per-instance
drains."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile, StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def concat(tiles):
    return ParticleTile(*(np.concatenate([getattr(t, k) for t in tiles]) for k in KEYS))


@pytest.mark.parametrize("dt", [1e-11, 1e-9])
def test_batched_matches_plain_per_instance(tables, dt):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Batch, Coupling, DeviceTile, DrainOutputs,
                                                                RreaParticleInteraction, StepOutputs)
    fields, tiles, seeds = wl.c3_batch(7, seed=5, scale=2e-3)
    sizes = [len(t.r) for t in tiles]
    print(f"case batched dt={dt}: instances of {sizes} electrons", flush=True)
    couplings = [Coupling(f, tables) for f in fields]
    plain = []
    for f, t, sd, cp in zip(fields, tiles, seeds, couplings):
        cfg = StepConfig(rng_seed=sd, dt_s=dt)
        dev = DeviceTile.from_host(t)
        out = StepOutputs(len(t.r), f, secondary_capacity=1 << 14)
        RreaParticleInteraction.apply_electrons(tables, cp, dev, cfg, out)
        drain = DrainOutputs(1 << 15, out)
        RreaParticleInteraction.drain_secondaries(tables, cp, cfg, [out], drain)
        torch.cuda.synchronize()
        plain.append((dev, out, drain))
    grids = [StepOutputs(1, f, secondary_capacity=1, per_particle=False) for f in fields]
    batch = Batch(couplings, grids, seeds)
    host = concat(tiles)
    dev = DeviceTile.from_host(host)
    inst = torch.as_tensor(np.repeat(np.arange(len(tiles), dtype=np.int32), sizes), device="cuda")
    out = StepOutputs(len(host.r), fields[0], secondary_capacity=1 << 15)
    cfg = StepConfig(rng_seed=0, dt_s=dt)
    batch.step(tables, "electron", dev, inst, cfg, out)
    bdrain = DrainOutputs(1 << 16, out)
    batch.drain(tables, cfg, [out], bdrain)
    torch.cuda.synchronize()
    offs = np.concatenate([[0], np.cumsum(sizes)])
    sec_b = out.secondaries_host()
    st_b = out.status.cpu().numpy()
    work_b, surv_b = bdrain.work_host(), bdrain.survivors_host()
    for i, (pdev, pout, pdrain) in enumerate(plain):
        a, b = offs[i], offs[i + 1]
        for k in KEYS:
            assert torch.equal(getattr(dev, k)[a:b], getattr(pdev, k)), (i, k)
        assert np.array_equal(st_b[a:b], pout.status.cpu().numpy()), i
        # the instance's J grid, to FP64 summation order
        for k in ("jr", "jz", "rho"):
            g, p = getattr(grids[i], k), getattr(pout, k)
            assert torch.allclose(g, p, rtol=0.0, atol=1e-12 * float(p.abs().max()) + 1e-300), (i, k)
        # its secondaries: the same records, parents offset into the batched tile
        sp = pout.secondaries_host()
        sb = sec_b[sec_b["instance"] == i]
        assert len(sb) == len(sp), i
        assert np.array_equal(sb["parent_index"] - a, sp["parent_index"])
        for f in ("species", "process", "x", "y", "z", "dir_x", "dir_y", "dir_z", "energy_eV", "birth_fraction",
                  "root_key", "lineage", "root_species"):
            assert np.array_equal(sb[f], sp[f]), (i, f)
        # its newborn drain: the same FIFO, in the same order, with the same outcome
        wp, wb = pdrain.work_host(), work_b[work_b["instance"] == i]
        assert len(wb) == len(wp), i
        for f in ("species", "process", "root_key", "lineage", "generation", "energy_eV", "birth_fraction"):
            assert np.array_equal(wb[f], wp[f]), (i, f)
        svp, svb = pdrain.survivors_host(), surv_b[surv_b["instance"] == i]
        assert len(svb) == len(svp), i
        for f in ("x", "y", "z", "energy_eV", "root_key", "lineage"):
            assert np.array_equal(svb[f], svp[f]), (i, f)
    print(f"batched dt={dt}: {len(sec_b)} secondaries, {len(work_b)} FIFO entries, {len(surv_b)} survivors match",
          flush=True)
    batch.close()


def concat_photons(tiles):
    t = concat(tiles)
    t.extra = {k: np.concatenate([x.extra[k] for x in tiles]) for k in ("r_cur", "theta_cur", "z_cur")}
    return t


def test_batched_full_mix(tables):
    """Electron, photon and positron steps of every instance in Apply's order,
    then the drain, batched against plain per instance."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Batch, Coupling, DeviceTile, DrainOutputs,
                                                                PhotonTile, RreaParticleInteraction, StepOutputs)
    dt = 1e-9
    fields, etiles, seeds = wl.c3_batch(5, seed=9, scale=1e-3)
    gtiles, ptiles = [], []
    for i, f in enumerate(fields):
        n = len(etiles[i].r)
        gtiles.append(wl.c2_photons(max(1, n // 2), f, dt_s=dt, seed=100 + i, first_id=(1 << 41) + 4096 * i))
        ptiles.append(wl.c2_electrons(max(1, n // 4), f, seed=200 + i, first_id=(1 << 40) + 4096 * i,
                                      charge_sign=1.0))
    print(f"case batched mix: e- {[len(t.r) for t in etiles]}, gamma {[len(t.r) for t in gtiles]}, "
          f"e+ {[len(t.r) for t in ptiles]}", flush=True)
    couplings = [Coupling(f, tables) for f in fields]
    plain = []
    for i, f in enumerate(fields):
        cfg = StepConfig(rng_seed=seeds[i], dt_s=dt)
        e, g, p = DeviceTile.from_host(etiles[i]), PhotonTile.from_host(gtiles[i]), DeviceTile.from_host(ptiles[i])
        oe = StepOutputs(len(e), f, secondary_capacity=1 << 13)
        og = StepOutputs(len(g), f, secondary_capacity=1 << 13, share_grid=oe)
        op = StepOutputs(len(p), f, secondary_capacity=1 << 13, share_grid=oe)
        RreaParticleInteraction.apply_electrons(tables, couplings[i], e, cfg, oe)
        RreaParticleInteraction.apply_photons(tables, couplings[i], g, cfg, og)
        RreaParticleInteraction.apply_positrons(tables, couplings[i], p, cfg, op)
        d = DrainOutputs(1 << 15, oe)
        RreaParticleInteraction.drain_secondaries(tables, couplings[i], cfg, [oe, og, op], d)
        torch.cuda.synchronize()
        plain.append(((e, g, p), (oe, og, op), d))
    grids = [StepOutputs(1, f, secondary_capacity=1, per_particle=False) for f in fields]
    batch = Batch(couplings, grids, seeds)
    cfg = StepConfig(rng_seed=0, dt_s=dt)
    dev, outs, insts = [], [], []
    for tiles, cls, cat in ((etiles, DeviceTile, concat), (gtiles, PhotonTile, concat_photons),
                            (ptiles, DeviceTile, concat)):
        sizes = [len(t.r) for t in tiles]
        t = cls.from_host(cat(tiles))
        inst = torch.as_tensor(np.repeat(np.arange(len(tiles), dtype=np.int32), sizes), device="cuda")
        o = StepOutputs(len(t), fields[0], secondary_capacity=1 << 15, share_grid=outs[0] if outs else None)
        dev.append(t)
        outs.append(o)
        insts.append((inst, np.concatenate([[0], np.cumsum(sizes)])))
    for s, name in enumerate(("electron", "photon", "positron")):
        batch.step(tables, name, dev[s], insts[s][0], cfg, outs[s])
    bd = DrainOutputs(1 << 16, outs[0])
    batch.drain(tables, cfg, outs, bd)
    torch.cuda.synchronize()
    work_b, surv_b = bd.work_host(), bd.survivors_host()
    for i, (pt, po, pd) in enumerate(plain):
        for s in range(3):
            a, b = insts[s][1][i], insts[s][1][i + 1]
            tb, tp = (dev[s].cur, pt[s].cur) if s == 1 else (dev[s], pt[s])
            for k in KEYS:
                assert torch.equal(getattr(tb, k)[a:b], getattr(tp, k)), (i, s, k)
            assert torch.equal(outs[s].status[a:b], po[s].status), (i, s)
            sb = outs[s].secondaries_host()
            sb = sb[sb["instance"] == i]
            sp = po[s].secondaries_host()
            assert len(sb) == len(sp), (i, s)
            assert np.array_equal(sb["parent_index"] - a, sp["parent_index"]), (i, s)
            assert np.array_equal(sb["energy_eV"], sp["energy_eV"]), (i, s)
        for k in ("jr", "jz", "rho"):
            g, p = getattr(grids[i], k), getattr(po[0], k)
            assert torch.allclose(g, p, rtol=0.0, atol=1e-12 * float(p.abs().max()) + 1e-300), (i, k)
        wp, wb = pd.work_host(), work_b[work_b["instance"] == i]
        assert len(wb) == len(wp), i
        for f in ("species", "process", "root_key", "lineage", "generation", "energy_eV"):
            assert np.array_equal(wb[f], wp[f]), (i, f)
        svp, svb = pd.survivors_host(), surv_b[surv_b["instance"] == i]
        assert len(svb) == len(svp), i
        for f in ("x", "y", "z", "energy_eV", "lineage"):
            assert np.array_equal(svb[f], svp[f]), (i, f)
    print(f"batched mix: {len(work_b)} FIFO entries, {len(surv_b)} survivors match", flush=True)
    batch.close()


def test_batch_tile_sort_orders_and_keeps_each_particle(tables):
    """rrea_batch_tile_sort: the batched tile comes back ordered by instance,
    then by the cell of the captured point on that instance's own mesh, with
    invalidated particles last and every array (and the instance index)
    permuted together; a batched step of the sorted tile gives every particle
    (found by id) the state the unsorted tile's step gives it, bit for bit."""
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Batch, Coupling, DeviceTile, StepOutputs
    fields, tiles, seeds = wl.c3_batch(6, seed=8, scale=1e-3)
    sizes = [len(t.r) for t in tiles]
    print(f"case batch tile sort: instances of {sizes} electrons", flush=True)
    couplings = [Coupling(f, tables) for f in fields]
    grids = [StepOutputs(1, f, secondary_capacity=1, per_particle=False) for f in fields]
    batch = Batch(couplings, grids, seeds)
    host = concat(tiles)
    rng = np.random.default_rng(3)
    # interleave the instances and invalidate a few particles, as an
    # avalanche's appended newborns and escapes leave a store
    perm = rng.permutation(len(host.r))
    host = ParticleTile(*(getattr(host, k)[perm] for k in KEYS))
    inst_h = np.repeat(np.arange(len(tiles), dtype=np.int32), sizes)[perm]
    dead = rng.random(len(host.r)) < 0.05
    host.id[dead] = _lib.INVALID_ID
    cfg = StepConfig(rng_seed=0, dt_s=1e-11)

    def run(sort):
        dev = DeviceTile.from_host(host)
        inst = torch.as_tensor(inst_h.copy(), device="cuda")
        pre = None
        if sort:
            batch.sort(dev, inst, energy_bits=3)
            pre = (dev.id.clone(), inst.clone())
        out = StepOutputs(len(host.r), fields[0], secondary_capacity=1 << 14)
        batch.step(tables, "electron", dev, inst, cfg, out)
        torch.cuda.synchronize()
        return dev, pre

    ref, _ = run(False)
    srt, (pre_id, pre_inst) = run(True)
    ids = pre_id.cpu().numpy()
    live = ids != _lib.INVALID_ID
    n_live = int(live.sum())
    assert n_live == int((~dead).sum()) and live[:n_live].all() and not live[n_live:].any()
    # the order of the sort's key over the live particles (captured point =
    # the step's input, taken from the unsorted host arrays by id)
    pos = {int(i): k for k, i in enumerate(host.id)}
    src = np.array([pos[int(i)] for i in ids[:n_live]])
    si = pre_inst.cpu().numpy()[:n_live]
    assert np.array_equal(si, inst_h[src])
    cell = np.array([int(np.clip(host.r[p] / fields[inst_h[p]].dr, 0, fields[inst_h[p]].nr - 1))
                     + fields[inst_h[p]].nr * int(np.clip(host.z[p] / fields[inst_h[p]].dz, 0, fields[inst_h[p]].nz - 1))
                     for p in src])
    key = si.astype(np.int64) * (1 << 20) + cell
    assert np.all(np.diff(key) >= 0)
    # every particle's outcome is the same wherever it sits
    ids = srt.id.cpu().numpy()
    ref_ids = ref.id.cpu().numpy()
    where = {int(i): k for k, i in enumerate(ref_ids) if i != _lib.INVALID_ID}
    for k in ("r", "z", "theta", "ux", "uy", "uz", "w"):
        a = getattr(srt, k).cpu().numpy()
        b = getattr(ref, k).cpu().numpy()
        for j in range(len(ids)):
            if ids[j] != _lib.INVALID_ID and int(ids[j]) in where:
                assert a[j] == b[where[int(ids[j])]], (k, j)
    batch.close()
