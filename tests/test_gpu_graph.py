"""GPU checks of graph-captured batched steps (rrea_batch_graph_*): whole
steps of uneven instances -- photon push, the three species steps in Apply's
order, the newborn drain -- captured once and replayed, against the same
steps made by the host-driven batched entry points on the same inputs.  The
particle states, the FIFO (ordered step records, then every generation) and
the survivors must be identical, record for record; the J grids equal to
FP64 summation order.  The graph is launched twice, so a replay after the
first is checked too.  In the avalanche mode every captured step also
commits its survivors into the batched stores, as a WarpX step adds the
newborns after the hook."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile, StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu

KEYS = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
FIFO_FIELDS = ("parent_index", "species", "process", "x", "y", "z", "dir_x", "dir_y", "dir_z", "weight", "energy_eV",
               "birth_fraction", "root_key", "lineage", "root_species", "generation", "instance")


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def concat(tiles, photon=False):
    t = ParticleTile(*(np.concatenate([getattr(x, k) for x in tiles]) for k in KEYS))
    if photon:
        t.extra = {k: np.concatenate([x.extra[k] for x in tiles]) for k in ("r_cur", "theta_cur", "z_cur")}
    return t


def mixed_batch(n_inst, seed, scale, dt):
    fields, etiles, seeds = wl.c3_batch(n_inst, seed=seed, scale=scale)
    gtiles, ptiles = [], []
    for i, f in enumerate(fields):
        n = len(etiles[i].r)
        g = wl.c2_photons(max(1, n // 2), f, dt_s=dt, seed=300 + i, first_id=(1 << 41) + 4096 * i)
        # the steps open with the photon push, so the current point is the
        # in-domain start (a pushed point outside the domain would have been
        # removed by the boundaries before the next push)
        g.extra = {"r_cur": g.r.copy(), "theta_cur": g.theta.copy(), "z_cur": g.z.copy()}
        gtiles.append(g)
        ptiles.append(wl.c2_electrons(max(1, n // 4), f, seed=400 + i, first_id=(1 << 40) + 4096 * i,
                                      charge_sign=1.0))
    return fields, seeds, {"electron": etiles, "photon": gtiles, "positron": ptiles}


def device_tiles(host):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile, PhotonTile
    out = {}
    for name, tiles in host.items():
        sizes = [len(t.r) for t in tiles]
        inst = torch.as_tensor(np.repeat(np.arange(len(tiles), dtype=np.int32), sizes), device="cuda")
        tile = PhotonTile.from_host(concat(tiles, True)) if name == "photon" else DeviceTile.from_host(concat(tiles))
        out[name] = (tile, inst)
    return out


def state(tile):
    t = tile.cur if hasattr(tile, "cur") else tile
    d = {k: getattr(t, k).clone() for k in KEYS}
    if hasattr(tile, "cur"):
        d.update(prev_r=tile.prev_r.clone(), prev_z=tile.prev_z.clone(), prev_theta=tile.prev_theta.clone())
    return d


# (population scale, dt): few records; a few hundred; thousands over several
# generations, past the LDS size of the one-workgroup record sort
@pytest.mark.parametrize("scale,dt", [(1.5e-3, 1e-10), (0.03, 1e-9), (0.1, 1e-8)])
def test_graph_steps_match_host_driven_steps(tables, scale, dt):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Batch, BatchGraph, Coupling, DrainOutputs,
                                                                InstanceGrids, StepOutputs)
    fields, seeds, host = mixed_batch(5, seed=11, scale=scale, dt=dt)
    print(f"case graph dt={dt}: e- {[len(t.r) for t in host['electron']]}, "
          f"gamma {[len(t.r) for t in host['photon']]}, e+ {[len(t.r) for t in host['positron']]}", flush=True)
    couplings = [Coupling(f, tables) for f in fields]
    cap = 1 << 16
    # the host-driven path
    grids_a = InstanceGrids(fields, sources=True)
    batch_a = Batch(couplings, grids_a.grids, seeds)
    tiles_a = device_tiles(host)
    outs = {}
    for name in ("electron", "photon", "positron"):
        tile = tiles_a[name][0]
        outs[name] = StepOutputs(len(tile), fields[0], secondary_capacity=cap, per_particle=False,
                                 share_grid=outs.get("electron"))
    drain_a = DrainOutputs(cap, outs["electron"])

    def host_steps(first, count):
        for s in range(first, first + count):
            grids_a.zero_()
            for o in outs.values():
                o.secondary_count.zero_()
            cfg = StepConfig(rng_seed=0, step=s, dt_s=dt)
            tiles_a["photon"][0].push(dt)
            for name in ("electron", "photon", "positron"):
                tile, inst = tiles_a[name]
                batch_a.step(tables, name, tile, inst, cfg, outs[name], check=False)
            batch_a.drain(tables, cfg, [outs[n] for n in ("electron", "photon", "positron")], drain_a, check=False)
        torch.cuda.synchronize()
        outs["electron"].check_error()

    # the graph path on its own copy of the same inputs
    grids_b = InstanceGrids(fields, sources=True)
    batch_b = Batch(couplings, grids_b.grids, seeds)
    tiles_b = device_tiles(host)
    graph = BatchGraph(tables, batch_b, tiles_b, StepConfig(rng_seed=0, dt_s=dt), record_capacity=cap,
                       clear=(grids_b.jr, grids_b.jz, grids_b.rho, grids_b.sources), steps=2, max_generations=12,
                       push_photons=True)
    for launch in range(2):
        first = 2 * launch
        print(f"graph launch {launch}: steps {first}, {first + 1}", flush=True)
        host_steps(first, 2)
        graph.launch(first)
        torch.cuda.synchronize()
        graph.check_error()
        for name in ("electron", "photon", "positron"):
            sa, sb = state(tiles_a[name][0]), state(tiles_b[name][0])
            for k in sa:
                assert torch.equal(sa[k], sb[k]), (launch, name, k)
        fa, fb = drain_a.work_host(), graph.records_host()
        assert len(fa) == len(fb), (launch, len(fa), len(fb))
        for f in FIFO_FIELDS:
            assert np.array_equal(fa[f], fb[f]), (launch, f)
        va, vb = drain_a.survivors_host(), graph.survivors_host()
        assert len(va) == len(vb), launch
        for f in FIFO_FIELDS:
            assert np.array_equal(va[f], vb[f]), (launch, f)
        # the last step's grids, to summation order
        for k in ("jr", "jz", "rho", "sources"):
            a, b = getattr(grids_a, k), getattr(grids_b, k)
            assert torch.allclose(a, b, rtol=0.0, atol=1e-12 * float(a.abs().max()) + 1e-300), (launch, k)
        gens = int(fb["generation"].max()) if len(fb) else 0
        print(f"graph launch {launch}: FIFO {len(fb)} records over {gens + 1} generations, {len(vb)} survivors "
              f"match", flush=True)
        if dt >= 1e-8:
            assert len(fb) > 2048 and gens >= 2, (len(fb), gens)
    # tallies accumulate over all four steps in both paths
    ta, tb = outs["electron"].tallies, graph.tallies
    assert torch.allclose(ta, tb, rtol=1e-12, atol=1e-9), (ta, tb)
    graph.close()
    batch_a.close()
    batch_b.close()


def test_graph_generation_limit_is_reported(tables):
    """A FIFO still holding children after the last captured generation is a
    capacity error (index -2), not silently dropped work."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Batch, BatchGraph, Coupling, InstanceGrids
    dt = 1e-8
    fields, seeds, host = mixed_batch(3, seed=12, scale=0.05, dt=dt)
    print("case graph generation limit: 1 generation at dt=10 ns", flush=True)
    couplings = [Coupling(f, tables) for f in fields]
    grids = InstanceGrids(fields)
    batch = Batch(couplings, grids.grids, seeds)
    tiles = device_tiles(host)
    graph = BatchGraph(tables, batch, tiles, StepConfig(rng_seed=0, dt_s=dt), record_capacity=1 << 15,
                       clear=(grids.jr, grids.jz, grids.rho), steps=1, max_generations=1)
    graph.launch(0)
    torch.cuda.synchronize()
    recs = graph.records_host()
    code, index = (int(v) for v in graph.error.cpu().tolist())
    # generation-1 children were made but never advanced
    assert (recs["generation"] > 0).any()
    assert (code, index) == (5, -2)
    assert int(graph.survivor_count.item()) == 0
    graph.close()
    batch.close()


@pytest.mark.parametrize("dt", [1e-9, 1e-8])
def test_graph_avalanche_matches_host_driven_steps(tables, dt):
    """The avalanche: each captured step ends with the commit of the drain's
    survivors into the batched stores (rrea_batch_graph_desc.counts), so
    newborn electrons, positrons and photons are transported from the next
    step on.  Two launches of a 3-step graph against six host-driven steps
    (BatchedAvalanche.step: the same batched entry points plus
    rrea_commit_survivors): the stores' lengths, the next id and every row of
    every store (state, weight, id, instance) identical; the last step's J
    grids equal to FP64 summation order."""
    from rrea_feedback_warpx_amrex_geant4_amd.simulation import BatchedAvalanche
    fields, etiles, seeds = wl.c3_batch(6, seed=21, scale=2e-3)
    cfg = StepConfig(rng_seed=0, dt_s=dt)
    caps = (1 << 15, 1 << 12, 1 << 14)
    host = BatchedAvalanche(tables, fields, etiles, seeds, cfg, capacity=caps, secondary_capacity=1 << 14)
    dev = BatchedAvalanche(tables, fields, etiles, seeds, cfg, capacity=caps, secondary_capacity=1 << 14)
    print(f"case graph avalanche dt={dt}: instances of {[len(t.r) for t in etiles]} electrons", flush=True)
    g = dev.graph(3, max_generations=8)
    for launch in range(2):
        for _ in range(3):
            host.step()
        g.launch(3 * launch)
        torch.cuda.synchronize()
        g.check_error()
        host.check_error()
        host.sync_counts()
        dev.sync_counts()
        assert torch.equal(host.counts, dev.counts), (launch, host.counts, dev.counts)
        assert torch.equal(host.next_id, dev.next_id)
        for a, b in zip(host.stores, dev.stores):
            n = a.n
            for k in KEYS + (("prev_r", "prev_z", "prev_theta") if a.photon else ()):
                assert torch.equal(getattr(a, k)[:n], getattr(b, k)[:n]), (launch, a.species, k)
            assert torch.equal(a.inst[:n], b.inst[:n]), (launch, a.species)
        for k in ("jr", "jz", "rho"):
            x, y = getattr(host.grids, k), getattr(dev.grids, k)
            assert torch.allclose(x, y, rtol=0.0, atol=1e-12 * float(x.abs().max()) + 1e-300), (launch, k)
    grown = [int(c) for c in host.counts.cpu()]
    assert grown[1] + grown[2] > 0 and grown[0] > sum(len(t.r) for t in etiles), grown
    print(f"graph avalanche dt={dt}: stores grew to {grown} (e-, e+, gamma) over 6 steps -- identical", flush=True)
