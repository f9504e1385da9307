"""Benchmark of the RREA transport step on one MI355X.

Workload: the at-scale configuration -- 2^24 macroparticles on a 256 x 1024 RZ
mesh over 2.56 km x 10 km, rho = 1.225 exp(-h / 8.4 km) with h = 4 km + z,
|E_z| piecewise-linear 0 -> 500 kV/m, dt = 1 ps, FP64, net-chord current
deposition and material-source accumulation on, secondaries staged and
ordered each step.  The charged shares
are transported: 70 % electrons (11,744,051) and 5 % positrons (838,861),
K ~ Exp(7.3 MeV) clipped to [10 keV, 100 MeV] in a 30-degree cone about the
force; 25 % photons (4,194,304), isotropic, same energy law.

Every --sort-every steps (default 25) the three tiles are reordered by cell
and energy band (rrea_tile_sort, the particle sort WarpX applies to its
tiles) inside the timed region.

One "step" is one RreaParticleInteraction::Apply (RreaParticleInteraction.cpp:57-60):
the electron kernels, the photon kernel, the positron kernels, each followed
by the deterministic secondary ordering -- the three passes run on three
streams, the photon and positron passes starting when the electron fast-path
kernel ends, under the electron hand-over (ConcurrentApply; --sequential runs
them one after another; the results are the same) -- then the newborn
residual advance of
every secondary the step made (the transient FIFO of
DrainAndCommitSecondaries), all charged chords depositing into one J grid.  Charged tiles step in place (the fused pre-push snapshot); the photon
tile is pushed on the device after its step (prev <- current, current += c dt
u/|u|, WarpX's photon push), so each step reads a freshly pushed chord.  The
push is inside the timed region.  Particles a step invalidates stay in their
tile with the invalid id and are skipped by later steps, as WarpX keeps them
until its redistribute.

Prints one JSON line (rank 0, N=1).  `roofline` is for the dominant kernel,
the electron kernel: algorithmic bytes per launch (112 B per live electron)
over its mean launch time from HIP events on its own stream over the timed
region, and `roofline.isolated` the same for the kernel timed alone on a
few steps after the timed region; `traffic` is that kernel's measured HBM
bytes per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE record in
profiles/electron_kernel_traffic.json when present.  The batched (plain per
instance, batched, graph-captured), drift-adjoint, field-side, population
and feedback-column legs and the CPU baseline follow (see their functions).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

B_ALG = 112.0  # bytes per particle-step: read r,z,theta,ux,uy,uz,w,id (64), write r,z,theta,ux,uy,uz (48)
HBM_PEAK = 8.0e12
TRAFFIC_RECORD = os.path.join(ROOT, "profiles", "electron_kernel_traffic.json")
INVALID = np.iinfo(np.int64).min


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n_particles: int, steps: int, threads: int):
    """The CPU oracle (a restatement of the reference step, OpenMP over
    64-particle chunks) timed on this host on a bounded sample of the same
    workload: electrons, photons and positrons in the 70:25:5 ratio, then
    the newborn drain of the step's secondaries, material sources
    accumulated.  Only the oracle's calls are timed (not the numpy photon
    push between steps)."""
    from oracle import binding as ob
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.transport import load_bundle

    ob.build()
    tables = ob.Tables(load_bundle())
    fields = wl.c2_fields()
    n_e, n_g, n_p = wl.c2_species_counts(n_particles)
    g = wl.c2_photons(n_g, fields, dt_s=1e-12, seed=4, first_id=1 << 41)
    photons = g.as_dict()
    photons.update(g.extra)
    tiles = {"electron": wl.c2_electrons(n_e, fields, seed=2).as_dict(), "photon": photons,
             "positron": wl.c2_electrons(n_p, fields, seed=3, first_id=1 << 40, charge_sign=1.0).as_dict()}
    done = 0
    dt = 0.0
    src = ob.new_sources(fields)
    for s in range(steps):
        cfg = StepConfig(rng_seed=20240623, step=s, dt_s=1e-12)
        src[:] = 0.0
        staged = []
        for species, p in tiles.items():
            t0 = time.perf_counter()
            if species == "photon":
                o = ob.photon_step(tables, fields, cfg, p, nthreads=threads, sec_cap=1 << 16, sources=src)
            else:
                o = ob.charged_step(species, tables, fields, cfg, p, nthreads=threads, sec_cap=1 << 16, sources=src)
            dt += time.perf_counter() - t0
            if o["code"] != 0:
                raise RuntimeError(f"oracle error {o['code']}")
            staged.append(o["secondaries"])
            done += int((p["id"] != INVALID).sum())
            p = dict(p)
            for k in ("r", "z", "theta", "ux", "uy", "uz"):
                p[k] = o[k]
            p["id"] = np.where(np.isin(o["status"], (1, 2, 3, 6, 7)), INVALID, p["id"])
            if species == "photon":
                p["r_cur"], p["theta_cur"], p["z_cur"] = wl.push_photons(p["r"], p["theta"], p["z"], p["ux"],
                                                                         p["uy"], p["uz"], cfg.dt_s)
            tiles[species] = p
        # the newborn residual advance of the step's secondaries
        t0 = time.perf_counter()
        d = ob.newborn_drain(tables, fields, cfg, np.concatenate(staged), cap=1 << 16, sources=src)
        dt += time.perf_counter() - t0
        if d["code"] != 0:
            raise RuntimeError(f"oracle drain error {d['code']}")
    return done / dt, dt, done


def batched_leg(args, tables):
    """The batched configuration on the same device: 256 independent
    avalanche instances (|E| ~ U[250, 1000] kV/m, rho ~ U[0.3, 1.225] kg/m^3,
    N0 ~ log-U[1e3, 1e6] electrons at K ~ U[1, 10] MeV, 64 x 256 RZ grids of
    1 m cells), dt = 1 ps, each step the electron step of every instance then
    the newborn drain of its secondaries, J deposited per instance.  Timed
    three times on the same inputs: the batched path
    (rrea_transport_step_batched / rrea_newborn_drain_batched: one launch set
    for all instances), the plain per-instance path (the drop-in Apply run
    instance after instance, one launch set and drain each), and the batched
    steps captured in one HIP graph (rrea_batch_graph_*).  Returns the report
    object."""
    import ctypes as C

    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Batch, BatchGraph, Coupling, DeviceTile,
                                                                DrainOutputs, InstanceGrids, RreaParticleInteraction,
                                                                StepOutputs)

    keys = ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")
    n_inst = args.batched_instances
    fields, tiles, seeds = wl.c3_batch(n_inst, seed=3, scale=args.batched_scale)
    sizes = np.array([len(t.r) for t in tiles], dtype=np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    n = int(offs[-1])
    log(f"batched: {n_inst} instances, {n} electrons (smallest {sizes.min()}, largest {sizes.max()})")
    dev_b = DeviceTile.from_host(type(tiles[0])(*(np.concatenate([getattr(t, k) for t in tiles]) for k in keys)))
    del tiles
    dev_p = dev_b.clone()
    dev_g = dev_b.clone()
    inst = torch.as_tensor(np.repeat(np.arange(n_inst, dtype=np.int32), sizes), device="cuda")
    couplings = [Coupling(f, tables) for f in fields]
    grids_b, grids_p = InstanceGrids(fields, sources=True), InstanceGrids(fields, sources=True)
    batch = Batch(couplings, grids_b.grids, seeds)
    out_b = StepOutputs(n, fields[0], secondary_capacity=max(1 << 14, n >> 6), per_particle=False)
    drain_b = DrainOutputs(max(1 << 15, n >> 5), out_b)
    # the plain path: instance i's tile is a view of its rows, its J grid its own
    plain_tiles = [DeviceTile(*(getattr(dev_p, k)[offs[i]:offs[i + 1]] for k in keys)) for i in range(n_inst)]
    counts = torch.zeros(n_inst, dtype=torch.int64, device="cuda")
    plain_outs, plain_drains = [], []
    for i, f in enumerate(fields):
        g = grids_p.grids[i]
        o = StepOutputs(int(sizes[i]), f, secondary_capacity=max(1 << 10, int(sizes[i]) >> 6), per_particle=False,
                        grid_buffers=(g.jr, g.jz, g.rho, g.sources))
        o.secondary_count = counts[i:i + 1]
        plain_outs.append(o)
        plain_drains.append(DrainOutputs(max(1 << 11, int(sizes[i]) >> 5), o))
    apply = RreaParticleInteraction
    dt = 1e-12

    def step_batched(s):
        out_b.secondary_count.zero_()
        grids_b.zero_()
        cfg = StepConfig(rng_seed=0, step=s, dt_s=dt)
        batch.step(tables, "electron", dev_b, inst, cfg, out_b, check=False)
        batch.drain(tables, cfg, [out_b], drain_b, check=False)

    def step_plain(s):
        counts.zero_()
        grids_p.zero_()
        for i in range(n_inst):
            cfg = StepConfig(rng_seed=seeds[i], step=s, dt_s=dt)
            apply.apply_electrons(tables, couplings[i], plain_tiles[i], cfg, plain_outs[i], order_secondaries=True,
                                  check=False)
            apply.drain_secondaries(tables, couplings[i], cfg, [plain_outs[i]], plain_drains[i], check=False)

    def live(t):
        return int((t.id != _lib.INVALID_ID).sum().item())

    def timed(run, tile):
        for s in range(args.batched_warmup):
            run(s)
        torch.cuda.synchronize()
        l0 = live(tile)
        t0 = time.perf_counter()
        for s in range(args.batched_steps):
            run(args.batched_warmup + s)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        l1 = live(tile)
        return 0.5 * (l0 + l1) * args.batched_steps / el, el, l0, l1

    L = _lib.lib()
    L.rrea_kernel_timing(batch.workspace, 1)
    rate_b, el_b, b0, b1 = timed(step_batched, dev_b)
    ms = C.c_double(0.0)
    cnt = C.c_int64(0)
    _lib.check(L.rrea_kernel_timing_read(batch.workspace, 0, C.byref(ms), C.byref(cnt)), "timing read")
    L.rrea_kernel_timing(batch.workspace, 0)
    out_b.check_error()
    mean_launch_s = ms.value * 1e-3 / max(cnt.value, 1)
    achieved = 0.5 * (b0 + b1) * B_ALG / mean_launch_s
    log(f"batched: {args.batched_steps} steps in {el_b:.3f} s, {rate_b:.4e} particle-steps/s; electron kernel "
        f"{mean_launch_s * 1e3:.3f} ms/launch")
    rate_p, el_p, p0, p1 = timed(step_plain, dev_p)
    for o in plain_outs:
        o.check_error()
    log(f"plain per instance: {args.batched_steps} steps in {el_p:.3f} s, {rate_p:.4e} particle-steps/s")
    # the graph-captured batched steps on a third copy: the warm-up steps and
    # the timed steps each one graph (rrea_batch_graph_*), no host round trip
    # inside a step
    grids_g = InstanceGrids(fields, sources=True)
    batch_g = Batch(couplings, grids_g.grids, seeds)
    clear = (grids_g.jr, grids_g.jz, grids_g.rho, grids_g.sources)
    gcfg = StepConfig(rng_seed=0, dt_s=dt)
    t0 = time.perf_counter()
    # a graph holds at least one step: with no warm-up steps only the timed
    # graph is captured
    graphs = [BatchGraph(tables, batch_g, {"electron": (dev_g, inst)}, gcfg, record_capacity=drain_b.capacity,
                         clear=clear, steps=k, max_generations=args.graph_generations)
              for k in (args.batched_warmup, args.batched_steps) if k > 0]
    capture_s = time.perf_counter() - t0
    if args.batched_warmup > 0:
        graphs[0].launch(0)
    torch.cuda.synchronize()
    g0 = live(dev_g)
    t0 = time.perf_counter()
    graphs[-1].launch(args.batched_warmup)
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t0
    g1 = live(dev_g)
    for g in graphs:
        g.check_error()
    rate_g = 0.5 * (g0 + g1) * args.batched_steps / el_g
    log(f"graph-captured batched: {args.batched_steps} steps in {el_g:.3f} s, {rate_g:.4e} particle-steps/s "
        f"(capture + instantiate {capture_s:.2f} s)")
    # the three paths must have advanced the same particles to the same state
    same = all(torch.equal(getattr(dev_b, k), getattr(dev_p, k)) for k in keys)
    same_g = all(torch.equal(getattr(dev_b, k), getattr(dev_g, k)) for k in keys)
    # the graph's loop continued to the configuration's step count, relaunched
    # back to back
    launches = max(1, args.graph_loop_steps // args.batched_steps)
    t0 = time.perf_counter()
    for i in range(1, launches):
        graphs[-1].launch(args.batched_warmup + i * args.batched_steps)
    torch.cuda.synchronize()
    el_loop = el_g + time.perf_counter() - t0
    g2 = live(dev_g)
    graphs[-1].check_error()
    loop_steps = launches * args.batched_steps
    rate_loop = 0.5 * (g0 + g2) * loop_steps / el_loop
    log(f"graph-captured loop: {loop_steps} steps in {el_loop:.3f} s, {rate_loop:.4e} particle-steps/s")
    for g in graphs:
        g.close()
    batch_g.close()
    batch.close()
    return {
        "workload": f"{n_inst} independent instances, {n} electrons (N0 log-uniform 1e3..1e6 x {args.batched_scale}),"
                    " |E| U[250,1000] kV/m, rho U[0.3,1.225] kg/m^3, 64x256 RZ grids, dt=1 ps, electron step + "
                    "newborn drain per step, J and material sources per instance",
        "instances": n_inst, "electrons": n, "steps": args.batched_steps, "warmup": args.batched_warmup,
        "batched": {"value": rate_b, "ms_per_step": el_b / args.batched_steps * 1e3,
                    "electron_kernel_ms": mean_launch_s * 1e3, "hbm_frac": achieved / HBM_PEAK,
                    "achieved_GBps": achieved / 1e9},
        "plain_per_instance": {"value": rate_p, "ms_per_step": el_p / args.batched_steps * 1e3},
        "graph_captured": {"value": rate_g, "ms_per_step": el_g / args.batched_steps * 1e3,
                           "steps_per_graph": args.batched_steps, "max_generations": args.graph_generations,
                           "capture_and_instantiate_s": capture_s,
                           "loop": {"steps": loop_steps, "launches": launches, "value": rate_loop,
                                    "ms_per_step": el_loop / loop_steps * 1e3}},
        "ratio": rate_b / rate_p,
        "ratio_graph": rate_g / rate_p,
        "unit": "particle-steps/s",
        "states_identical": bool(same),
        "graph_states_identical": bool(same_g),
    }


def drift_leg(args, tables):
    """The differentiation configuration: 2^20 electrons (K ~ U[1, 20] MeV),
    64 drift steps of 100 ps (push + CSDA loss + deposition, events frozen)
    on a 128 x 512 grid, checkpoints every 8; L = sum(J_z W), W ~ N(0, 1).
    Times the forward (with checkpoints) and the adjoint (reverse sweep with
    recompute) separately, HIP events around each on the current stream."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import DriftConfig, DriftProblem

    n = args.drift_particles
    fields, parts = wl.c4_drift(n)
    prob = DriftProblem(tables, fields, parts, DriftConfig(steps=64, dt_s=1e-10, checkpoint_every=8))
    g = torch.Generator(device="cpu").manual_seed(11)
    gjr = torch.zeros(prob.jr_size, dtype=torch.float64).cuda()
    gjz = torch.randn(prob.jz_size, generator=g, dtype=torch.float64).cuda()
    fw, bw = [], []
    for it in range(4):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record()
        out = prob.forward(prob.er0, prob.ez0, gjr, gjz)
        b.record()
        prob.adjoint(prob.er0, prob.ez0, out["ck"], out["ck_alive"], gjr, gjz)
        c.record()
        torch.cuda.synchronize()
        if it > 0:
            fw.append(a.elapsed_time(b))
            bw.append(b.elapsed_time(c))
    steps = n * 64
    log(f"drift: forward {np.mean(fw):.2f} ms, adjoint {np.mean(bw):.2f} ms for {n} electrons x 64 steps")
    return {"workload": f"{n} electrons x 64 drift steps of 100 ps, 128x512 grid, checkpoints every 8, "
                        "L = sum(J_z W)",
            "forward_ms": float(np.mean(fw)), "adjoint_ms": float(np.mean(bw)),
            "forward_particle_steps_per_s": steps / (np.mean(fw) * 1e-3),
            "adjoint_particle_steps_per_s": steps / (np.mean(bw) * 1e-3),
            "checkpoint_bytes": int(out["ck"].numel() * 8 + out["ck_alive"].numel())}


def field_leg(args, tables):
    """The field side of the feedback loop at the plain configuration's grid
    (256 x 1024 RZ, dr = 10 m, dz = 9.77 m): per step the net-chord deposit
    (a synthetic padded index-space grid, half its faces non-zero) converted
    to physical face currents, one TM Maxwell leapfrog step driven by them,
    and the total field handed back to the transport's gather
    (rrea_current_to_physical, rrea_maxwell_step, rrea_fields_update).  HIP
    events around the timed steps on the current stream."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.maxwell import MaxwellField
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling

    fields = wl.c2_fields()
    nrp, nzp = fields.padded_shape()
    g = torch.Generator(device="cpu").manual_seed(5)
    jr_i = torch.randn((nrp + 1) * nzp, generator=g, dtype=torch.float64)
    jz_i = torch.randn(nrp * (nzp + 1), generator=g, dtype=torch.float64)
    jr_i[torch.rand(jr_i.numel(), generator=g) < 0.5] = 0.0
    jz_i[torch.rand(jz_i.numel(), generator=g) < 0.5] = 0.0
    jr_i, jz_i = jr_i.cuda(), jz_i.cuda()
    cp = Coupling(fields, tables)
    m = MaxwellField(fields)
    for _ in range(5):
        m.feedback_step(jr_i, jz_i, 1e-12, coupling=cp)
    steps = args.field_steps
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        m.feedback_step(jr_i, jz_i, 1e-12, coupling=cp)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    nr, nz = fields.nr, fields.nz
    n_er, n_ez, n_bt, cells = (nr + 1) * nz, nr * (nz + 1), (nr + 1) * (nz + 1), nr * nz
    # algorithmic bytes: read both index grids, write jr/jz; Faraday + absorber
    # read Er, Ez and read/write Btheta; Ampere reads Btheta, jr, jz and
    # reads/writes Er, Ez; the gather's cell totals read Er, Ez, both ambients,
    # write two cell arrays, and copy Btheta
    alg = 8 * ((nrp + 1) * nzp + nrp * (nzp + 1) + n_er + n_ez
               + n_er + n_ez + 2 * n_bt
               + n_bt + n_er + n_ez + 2 * (n_er + n_ez)
               + 2 * (n_er + n_ez) + 2 * cells + 2 * n_bt)
    log(f"field side: {ms * 1e3:.1f} us per feedback step on {nr}x{nz} ({alg / (ms * 1e-3) / 1e9:.0f} GB/s algorithmic)")
    return {"workload": f"{nr}x{nz} RZ grid: deposit -> physical current -> TM Maxwell step -> gather field totals",
            "us_per_step": ms * 1e3, "alg_bytes_per_step": alg, "achieved_GBps": alg / (ms * 1e-3) / 1e9,
            "steps": steps}


def population_leg(args, tables):
    """Population control at the plain configuration's scale: the
    charge-exact thinning plan (rrea_population_thin) over the electrons of
    the 256 x 1024 grid, a target of half the count.  Times the whole call
    (record formation, the radix passes, capacities, the per-support plans
    and the write-back); checks every support's CIC charges are unchanged."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.population import thin_tile
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile

    fields = wl.c2_fields()
    n = args.population_particles
    host = wl.c2_electrons(n, fields, seed=31)
    times = []
    res = None
    for it in range(3):
        tile = DeviceTile.from_host(host)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = thin_tile(tile, fields, target=n // 2, rng_seed=20240623, step=it)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    ms = float(np.median(times)) * 1e3
    log(f"population control: {n} electrons -> target {n // 2}: removed {res.removed} "
        f"(capacity {res.capacity}) in {ms:.1f} ms")
    return {"workload": f"{n} C2 electrons on the 256x1024 grid, target {n // 2}", "ms": ms,
            "removed": res.removed, "capacity": res.capacity, "requested": res.requested,
            "records_per_s": n / (ms * 1e-3)}


def column_leg(args, tables):
    """The feedback column (one instance of the feedback configuration):
    every step injects electrons at the column top (power law E^-1.5 over
    1-100 MeV, r < 500 m, 1e5 electrons/ns), runs the photon push, Apply,
    the drain, the commit and compaction, converts the deposit to current,
    advances the TM Maxwell field and hands it back to the gather; every 10
    steps population control thins electrons and positrons above 1.25 x
    their targets.  dt = 1 ns on the 256 x 1024 column grid.  Times the whole
    loop; particle-steps counted are the live particles at each step."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd.column import FeedbackColumn

    steps = args.column_steps
    col = FeedbackColumn(tables, dt_s=1e-9, inject_macros=args.column_inject, targets=(1 << 18, 1 << 15),
                         control_every=10, capacity=1 << 22)
    col.run(10)  # warm-up: the first injections and control pass
    torch.cuda.synchronize()
    live = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        col.step()
        live += sum(s.n for s in col.sim.stores)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    obs = col.observe()
    log(f"feedback column: {steps} steps in {el:.2f} s ({el / steps * 1e3:.1f} ms/step), "
        f"{live / el:.3e} particle-steps/s; {obs}")
    single = {"workload": "feedback column, 256x1024 grid, dt = 1 ns, injection + transport + drain + commit + "
                          "Maxwell feedback + population control", "steps": steps, "ms_per_step": el / steps * 1e3,
              "particle_steps_per_s": live / el, "final": obs}
    # the feedback configuration's batch: 16 columns, E_z peaks 300-600 kV/m
    from rrea_feedback_warpx_amrex_geant4_amd.column import BatchedColumn
    peaks = np.linspace(3.0e5, 6.0e5, 16)
    bc = BatchedColumn(tables, peaks, inject_macros=args.column_inject // 8, targets=(1 << 16, 1 << 13),
                       control_every=10, capacity=1 << 23)
    bc.run(10)
    torch.cuda.synchronize()
    live = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        bc.step()
        live += sum(s.n for s in bc.stores)
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    obs_b = bc.observe()
    bc.close()
    log(f"batched columns: 16 instances, {steps} steps in {el_b:.2f} s ({el_b / steps * 1e3:.1f} ms/step), "
        f"{live / el_b:.3e} particle-steps/s; electrons per instance {obs_b['electrons']}")
    single["batched_16"] = {"steps": steps, "ms_per_step": el_b / steps * 1e3, "particle_steps_per_s": live / el_b,
                            "electrons_per_instance": obs_b["electrons"],
                            "removed_electrons": obs_b["removed_electrons"]}
    return single


def traffic_record():
    """Measured HBM bytes per launch of the electron kernel, or None."""
    try:
        with open(TRAFFIC_RECORD) as f:
            rec = json.load(f)
        return float(rec["bytes_per_launch"]), rec
    except (OSError, KeyError, ValueError):
        return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=None, help="total macroparticles (default 2^24)")
    ap.add_argument("--sort-every", type=int, default=25,
                    help="reorder the tiles by cell every this many steps (0: never), inside the timed region")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick the sort interval and energy-band bits by timing candidates before the warm-up")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sequential", action="store_true",
                    help="run the three species passes one after another on one stream (default: concurrently)")
    ap.add_argument("--cpu-particles", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=150)
    ap.add_argument("--batched-instances", type=int, default=256, help="instances of the batched leg (0: skip it)")
    ap.add_argument("--batched-scale", type=float, default=1.0, help="scale of the batched leg's populations")
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batched-warmup", type=int, default=2)
    ap.add_argument("--field-steps", type=int, default=200, help="timed feedback steps of the field leg (0: skip)")
    ap.add_argument("--device-drain", type=int, default=0,
                    help="concurrent step: drain with the counts on the device, up to this many generations "
                         "(0: the host-driven drain)")
    ap.add_argument("--push-on-photon-stream", action=argparse.BooleanOptionalAction, default=False,
                    help="concurrent step: push the photons on their own stream right after their pass "
                         "(default: on the main stream after the drain, measured 1-3 %% faster)")
    ap.add_argument("--column-steps", type=int, default=100, help="timed steps of the feedback-column leg (0: skip)")
    ap.add_argument("--column-inject", type=int, default=2048, help="macroparticles injected per column step")
    ap.add_argument("--population-particles", type=int, default=1 << 22,
                    help="electrons of the population-control leg (0: skip)")
    ap.add_argument("--graph-loop-steps", type=int, default=500,
                    help="steps of the batched leg's graph-captured loop (the timed graph relaunched)")
    ap.add_argument("--graph-generations", type=int, default=4,
                    help="newborn FIFO generations captured per graph step of the batched leg")
    ap.add_argument("--drift-particles", type=int, default=1 << 20,
                    help="electrons of the drift-adjoint leg (0: skip it)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run only a device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 8, "copy_calibration_bytes_written": n * 8}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU runs are out of scope "
              "(the J-grid allreduce design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C

    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (ConcurrentApply, Coupling, DeviceTile, DrainOutputs,
                                                                InteractionTables, PhotonTile, RreaParticleInteraction,
                                                                StepOutputs)

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    total = args.particles if args.particles else 1 << 24
    n_e, n_g, n_p = wl.c2_species_counts(total)
    log(f"setup: {n_e} e- + {n_g} gamma + {n_p} e+ = {total}, building tables")
    tables = InteractionTables()
    fields = wl.c2_fields()
    coupling = Coupling(fields, tables)
    dev_e = DeviceTile.from_host(wl.c2_electrons(n_e, fields, seed=2))
    dev_p = DeviceTile.from_host(wl.c2_electrons(n_p, fields, seed=3, first_id=1 << 40, charge_sign=1.0))
    dev_g = PhotonTile.from_host(wl.c2_photons(n_g, fields, dt_s=1e-12, seed=4, first_id=1 << 41))
    out_e = StepOutputs(n_e, fields, secondary_capacity=1 << 20, per_particle=False, sources=True)
    out_p = StepOutputs(n_p, fields, secondary_capacity=1 << 18, per_particle=False, share_grid=out_e)
    out_g = StepOutputs(n_g, fields, secondary_capacity=1 << 18, per_particle=False, share_grid=out_e)
    drain = DrainOutputs(1 << 21, out_e)
    L = _lib.lib()
    apply = RreaParticleInteraction
    conc = None if args.sequential else ConcurrentApply(tables, fields)
    # the handle each species' kernels are timed on
    h_e, h_g, h_p = (coupling.h,) * 3 if conc is None else tuple(c.h for c in conc.couplings)

    knobs = {"sort_every": args.sort_every, "energy_bits": -1, "first_step": 0}

    def one_step(s, sort_now, energy_bits):
        if sort_now:
            # WarpX's periodic particle sort: cell-ordered tiles
            dev_e.sort_by_cell(coupling, energy_bits=energy_bits)
            dev_p.sort_by_cell(coupling, energy_bits=energy_bits)
            dev_g.sort_by_cell(coupling, energy_bits=energy_bits)
        for o in (out_e, out_g, out_p):
            o.secondary_count.zero_()
        out_e.jr.zero_()
        out_e.jz.zero_()
        out_e.rho.zero_()
        out_e.sources.zero_()
        cfg = StepConfig(rng_seed=20240623, step=s, dt_s=1e-12)
        if conc is not None:
            conc.run(dev_e, dev_g, dev_p, cfg, out_e, out_g, out_p, drain, push_photons=args.push_on_photon_stream,
                     device_drain=args.device_drain)
            if not args.push_on_photon_stream:
                dev_g.push(cfg.dt_s)
        else:
            apply.apply_electrons(tables, coupling, dev_e, cfg, out_e, order_secondaries=True, check=False)
            apply.apply_photons(tables, coupling, dev_g, cfg, out_g, order_secondaries=True, check=False)
            apply.apply_positrons(tables, coupling, dev_p, cfg, out_p, order_secondaries=True, check=False)
            apply.drain_secondaries(tables, coupling, cfg, [out_e, out_g, out_p], drain, check=False)
            dev_g.push(cfg.dt_s)

    def run_step(s):
        # s counts from 0 (the first sort); the RNG step index continues past the tuning steps
        every = knobs["sort_every"]
        one_step(knobs["first_step"] + s, every > 0 and s % every == 0, knobs["energy_bits"])

    def live():
        return tuple(int((t.id != _lib.INVALID_ID).sum().item()) for t in (dev_e, dev_g.cur, dev_p))

    tuned = None
    if args.autotune:
        from rrea_feedback_warpx_amrex_geant4_amd.autotune import tune
        log("autotune: tile-sort interval x energy-band bits")
        choice = tune(one_step, first_step=0)
        knobs["sort_every"], knobs["energy_bits"] = choice.sort_every, choice.energy_bits
        knobs["first_step"] = choice.steps_run
        tuned = choice.as_dict()
        log(f"autotune: sort every {choice.sort_every} steps, {choice.energy_bits} band bits "
            f"({choice.ms_per_step:.3f} ms/step)")
    log(f"warmup {args.warmup}")
    for s in range(args.warmup):
        run_step(s)
    torch.cuda.synchronize()
    out_e.check_error()
    e0, g0, p0 = live()

    for h in {h_e.value, h_g.value, h_p.value}:
        L.rrea_kernel_timing(h, 1)
    log(f"timed {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        run_step(args.warmup + s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = {}
    for name, kid, h in (("electron", 0, h_e), ("electron_handover", 1, h_e), ("positron", 2, h_p),
                         ("positron_handover", 3, h_p), ("photon", 4, h_g), ("photon_handover", 5, h_g),
                         ("newborn", 6, h_e)):
        ms = C.c_double(0.0)
        cnt = C.c_int64(0)
        _lib.check(L.rrea_kernel_timing_read(h, kid, C.byref(ms), C.byref(cnt)), "timing read")
        kms[name] = (ms.value, cnt.value)
    for h in {h_e.value, h_g.value, h_p.value}:
        L.rrea_kernel_timing(h, 0)
    out_e.check_error()
    e1, g1, p1 = live()
    # live counts decline monotonically and by a tiny fraction per step; the
    # mean of the ends bounds the processed count
    processed = 0.5 * ((e0 + e1) + (g0 + g1) + (p0 + p1)) * args.steps
    value = processed / elapsed
    e_ms, e_launches = kms["electron"]
    mean_launch_s = e_ms * 1e-3 / max(e_launches, 1)
    alg_bytes = 0.5 * (e0 + e1) * B_ALG
    achieved = alg_bytes / mean_launch_s
    log(f"live e- {e0} -> {e1}, gamma {g0} -> {g1}, e+ {p0} -> {p1}; wall {elapsed:.3f}s; kernels " +
        ", ".join(f"{k} {v[0]:.2f} ms / {v[1]}" for k, v in kms.items()))
    traffic, rec = traffic_record()

    # the electron kernel alone: a few sequential Applies after the timed
    # region (the timed region overlaps the species passes, which stretches
    # the kernel's own launch time while it shares the CUs)
    iso = None
    if conc is not None:
        L.rrea_kernel_timing(coupling.h, 1)
        e_before = live()[0]
        for s in range(3):
            cfg = StepConfig(rng_seed=20240623, step=args.warmup + args.steps + s, dt_s=1e-12)
            for o in (out_e, out_g, out_p):
                o.secondary_count.zero_()
            apply.apply_electrons(tables, coupling, dev_e, cfg, out_e, order_secondaries=True, check=False)
        ms = C.c_double(0.0)
        cnt = C.c_int64(0)
        _lib.check(L.rrea_kernel_timing_read(coupling.h, 0, C.byref(ms), C.byref(cnt)), "timing read")
        L.rrea_kernel_timing(coupling.h, 0)
        iso_s = ms.value * 1e-3 / max(cnt.value, 1)
        iso_ach = e_before * B_ALG / iso_s
        iso = {"mean_launch_ms": iso_s * 1e3, "achieved": iso_ach / 1e9, "frac": iso_ach / HBM_PEAK,
               "note": "the electron kernel timed alone (3 electron steps after the timed region)"}
        log(f"electron kernel alone: {iso_s * 1e3:.3f} ms/launch, {iso_ach / 1e9:.0f} GB/s")

    batched = batched_leg(args, tables) if args.batched_instances > 0 else None
    drift = drift_leg(args, tables) if args.drift_particles > 0 else None
    field = field_leg(args, tables) if args.field_steps > 0 else None
    population = population_leg(args, tables) if args.population_particles > 0 else None
    column = column_leg(args, tables) if args.column_steps > 0 else None

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle, {threads} threads, {args.cpu_particles} particles x {args.cpu_steps} steps")
        rate, secs, done = cpu_baseline(args.cpu_particles, args.cpu_steps, threads)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_particles} C2 particles (e-:gamma:e+ = 70:25:5) x {args.cpu_steps} steps "
                         f"({done} particle-steps, {secs:.1f} s), oracle restatement, OpenMP dynamic over "
                         f"64-particle chunks"}

    line = {
        "metric": "particle-steps/s on 1 MI355X, plain vs batched, at 1e-12 step match; % HBM peak",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"C2 plain at scale, {total} macroparticles: {n_e} e- + {n_g} gamma + {n_p} e+, "
                               "dt=1 ps, RZ 256x1024, deposit on, material sources accumulated, secondaries ordered and their newborn residual "
                               "advance drained, photon push on device",
                   "electrons": n_e, "photons": n_g, "positrons": n_p, "path": "plain",
                   "species_passes": "sequential" if args.sequential else "concurrent (three streams)",
                   "tile_sort_every": knobs["sort_every"], "tile_sort_energy_bits": knobs["energy_bits"],
                   "autotune": tuned,
                   "kernel_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in kms.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "electron_fast_kernel", "alg_bytes_per_launch": alg_bytes,
                     "mean_launch_ms": mean_launch_s * 1e3, "isolated": iso},
        "cpu_baseline": cpu,
        "batched_vs_plain": batched,
        "field_side": field,
        "population_control": population,
        "feedback_column": column,
        "drift_adjoint": drift,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
