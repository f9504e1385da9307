"""Fail-closed steps: a transport error stops the commit of the step's
survivors, as the reference aborts before AddNParticles
(warpx_overlay/Source/Rrea/RreaSecondaryCommit.cpp:240-265).

- A non-finite gathered field (RreaPathFieldWork.H:113-121): Ez is NaN on
  the z-faces 14-16 that some positrons reach within the 1 ns step; it reports
  "invalid" (4) at a positron, the oracle's positron step stops with the same
  code, and nothing is committed although the drain of the other particles'
  secondaries made survivors.
- The strict transport guard (RreaTransportStep.H:186-193): one allowed
  subcycle where a 1 ns electron step needs more; "strict guard" (3), the
  oracle the same, nothing committed.  With the guard off the same step
  commits.
- The staged device drain: a FIFO smaller than the step's staged records
  reports a capacity error at index -1, and a generation limit reached with
  children pending reports index -2; neither hands out a survivor."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import binding as ob  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402
from test_gpu_simulation import mixed_start  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def state(sim):
    torch.cuda.synchronize()
    return sim.counts.cpu().numpy().copy(), int(sim.next_id.item()), [s.host()["id"].copy() for s in sim.stores]


def assert_nothing_committed(sim, before):
    counts, next_id, ids = state(sim)
    assert np.array_equal(counts, before[0]), (counts, before[0])
    assert next_id == before[1]
    assert [s.n for s in sim.stores] == before[0].tolist()
    for a, b in zip(ids, before[2]):
        assert len(a) == len(b)
    assert (sim.created == 0).all()


def nan_rows(fields, rows):
    ez = fields.ez.copy().reshape(fields.nz + 1, fields.nr)
    ez[rows, :] = np.nan
    return dataclasses.replace(fields, ez=ez.reshape(-1))


def test_nonfinite_field_commits_nothing(tables, otables):
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.simulation import Simulation
    fields, e, g, p = mixed_start(1000, 11)
    cfg = StepConfig(rng_seed=5, dt_s=1e-9)
    bad = nan_rows(fields, slice(14, 17))
    print("case fail-closed: NaN Ez on z-faces 14-16, 1000 e- + 200 gamma + 100 e+, one 1 ns step", flush=True)
    ok = Simulation(tables, fields, cfg, e, g, p, next_id=1 << 30)
    ok.step()
    assert ok.created.sum() > 0
    sim = Simulation(tables, bad, cfg, e, g, p, next_id=1 << 30)
    before = state(sim)
    with pytest.raises(_lib.RreaHipError, match="invalid"):
        sim.step()
    code, index = sim.out[0].error.cpu().tolist()
    assert code == 4 and 0 <= index < len(p.r), (code, index)
    zi = int((p.z[index] - bad.z_lo) // bad.dz)
    assert 13 <= zi <= 17, (index, p.z[index])  # its chord (at most 0.3 m, downward) gathers a NaN face
    assert int(sim.drain.survivor_count.item()) == 0
    assert int(sim.drain.work_count.item()) > 0  # the drain ran; its survivors were withheld
    assert_nothing_committed(sim, before)
    assert ob.positron_step(otables, bad, cfg, p.as_dict(), nthreads=8)["code"] == 4
    assert ob.electron_step(otables, bad, cfg, e.as_dict(), nthreads=8)["code"] == 0


@pytest.mark.parametrize("strict", [True, False])
def test_strict_guard_commits_nothing(tables, otables, strict):
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.simulation import Simulation
    fields, e, g, p = mixed_start(1000, 11)
    cfg = StepConfig(rng_seed=5, dt_s=1e-9, max_subcycles=1, strict_transport_guards=strict)
    print(f"case fail-closed: max_subcycles=1, strict={strict}, one 1 ns step", flush=True)
    sim = Simulation(tables, fields, cfg, e, g, p, next_id=1 << 30)
    before = state(sim)
    o = ob.electron_step(otables, fields, cfg, e.as_dict(), nthreads=8)
    if not strict:
        sim.step()
        assert o["code"] == 0 and sim.created.sum() > 0
        return
    assert o["code"] == 3
    with pytest.raises(_lib.RreaHipError, match="strict guard"):
        sim.step()
    code, index = sim.out[0].error.cpu().tolist()
    assert code == 3 and 0 <= index < len(e.r), (code, index)
    assert int(sim.drain.survivor_count.item()) == 0
    assert_nothing_committed(sim, before)


def staged_drain(tables, coupling, cfg, outs, capacity, max_generations):
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DrainOutputs, _step_config
    outs[0].error.zero_()
    d = DrainOutputs(capacity, outs[0])
    d.survivor_count.fill_(-7)
    recs = (C.c_void_p * 3)(*[o.secondaries.data_ptr() for o in outs])
    cnts = (C.c_void_p * 3)(*[o.secondary_count.data_ptr() for o in outs])
    caps = (C.c_int64 * 3)(*[o.secondary_capacity for o in outs])
    io = d.struct(0)
    c = _step_config(cfg)
    _lib.check(_lib.lib().rrea_newborn_drain_staged(tables.h, coupling.h, C.byref(c), recs, cnts, caps, 3,
                                                    C.byref(io), max_generations,
                                                    torch.cuda.current_stream().cuda_stream),
               "rrea_newborn_drain_staged")
    torch.cuda.synchronize()
    return d, outs[0].error.cpu().tolist()


def test_staged_drain_capacity_and_generation_errors(tables):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DeviceTile, DrainOutputs, PhotonTile,
                                                                RreaParticleInteraction, StepOutputs)
    fields = wl.c2_fields()
    n = 1 << 14
    e = wl.c2_electrons(n, fields, seed=2)
    g = wl.c2_photons(n // 2, fields, dt_s=1e-8, seed=4, first_id=1 << 41)
    p = wl.c2_electrons(n // 4, fields, seed=3, first_id=1 << 40, charge_sign=1.0)
    print(f"case staged drain errors: {n} e-, {n // 2} gamma, {n // 4} e+, one 10 ns step", flush=True)
    coupling = Coupling(fields, tables)
    oe = StepOutputs(n, fields, secondary_capacity=1 << 15)
    og = StepOutputs(n // 2, fields, secondary_capacity=1 << 15, share_grid=oe)
    op = StepOutputs(n // 4, fields, secondary_capacity=1 << 15, share_grid=oe)
    outs = (oe, og, op)
    cfg = StepConfig(rng_seed=81, step=0, dt_s=1e-8)  # long enough for newborns to have children
    apply = RreaParticleInteraction
    apply.apply_electrons(tables, coupling, DeviceTile.from_host(e), cfg, oe)
    apply.apply_photons(tables, coupling, PhotonTile.from_host(g), cfg, og)
    apply.apply_positrons(tables, coupling, DeviceTile.from_host(p), cfg, op)
    torch.cuda.synchronize()
    staged = sum(int(o.secondary_count.item()) for o in outs)
    # the host-driven drain of the same records: how many generations, how many survivors
    ref = DrainOutputs(1 << 17, StepOutputs(1, fields, secondary_capacity=1, per_particle=False))
    apply.drain_secondaries(tables, coupling, cfg, list(outs), ref)
    torch.cuda.synchronize()
    work = ref.work_host()
    gens = int(work["generation"].max())
    assert staged > 16 and gens >= 1 and int(ref.survivor_count.item()) > 0, (staged, gens)
    print(f"staged drain: {staged} staged records, {len(work)} FIFO entries, {gens} generations", flush=True)
    # the full drain on the device agrees with it
    d, err = staged_drain(tables, coupling, cfg, outs, 1 << 17, 16)
    assert err == [0, 0]
    assert (d.survivors_host() == ref.survivors_host()).all()
    # FIFO smaller than the staged records: truncated, capacity error at -1, no survivors
    d, err = staged_drain(tables, coupling, cfg, outs, staged // 2, 16)
    assert err == [5, -1], err
    assert int(d.survivor_count.item()) == 0
    # one generation allowed, children pending: capacity error at -2, no survivors
    d, err = staged_drain(tables, coupling, cfg, outs, 1 << 17, 1)
    assert err == [5, -2], err
    assert int(d.survivor_count.item()) == 0
    assert int(d.work_count.item()) > staged  # generation 1 was staged, never advanced
