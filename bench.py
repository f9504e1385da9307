"""Benchmark of the RREA transport step on one MI355X.

Workload: the at-scale configuration -- 2^24 macroparticles on a 256 x 1024 RZ
mesh over 2.56 km x 10 km, rho = 1.225 exp(-h / 8.4 km) with h = 4 km + z,
|E_z| piecewise-linear 0 -> 500 kV/m, dt = 1 ps, FP64, net-chord current
deposition and material-source accumulation on, secondaries staged and
ordered each step.  The charged shares
are transported: 70 % electrons (11,744,051) and 5 % positrons (838,861),
K ~ Exp(7.3 MeV) clipped to [10 keV, 100 MeV] in a 30-degree cone about the
force; 25 % photons (4,194,304), isotropic, same energy law.

The three tiles are reordered by cell and energy band (rrea_tile_sort, the
particle sort WarpX applies to its tiles) at the first timed step and every
--sort-every steps after it (default 25; the autotuner picks the interval
and the band width before the warm-up), inside the timed region, so it holds
at least its amortised share of sorts.

One "step" is one RreaParticleInteraction::Apply (RreaParticleInteraction.cpp:57-60)
and the WarpX calls around it: the electron kernels, the photon kernel, the
positron kernels, each followed by the deterministic secondary ordering --
the three passes run on three streams, the positron pass starting with
the electron pass (its latency-bound hand-over runs under the electron fast
kernel) and the photon pass when the electron fast-path kernel ends, under
the electron hand-over (ConcurrentApply; --early chooses the passes that
start with the electron pass; --sequential runs them one after another; the
results are the same) -- then the newborn residual advance of every
secondary the step made (the transient FIFO of DrainAndCommitSecondaries,
with its generation counts on the device: no host round trip in the step),
all charged chords depositing into one J grid; then the commit of the
drain's survivors into the species' stores (AddNParticles,
RreaSecondaryCommit.cpp:594-629), so newborns are transported from the next
step on; then WarpX's photon push on the device (prev <- current, current
+= c dt u/|u|), so each step reads a freshly pushed chord.  Charged tiles
step in place (the fused pre-push snapshot).  Everything is inside the timed
region.  Particles a step invalidates stay in their store with the invalid
id and are skipped by later steps, as WarpX keeps them until its
redistribute.

Prints one JSON line (rank 0, N=1).  `roofline` is for the dominant kernel,
the electron kernel: algorithmic bytes per launch (112 B per live electron)
over its mean launch time from HIP events on its own stream over the timed
region, and `roofline.isolated` the same for the kernel timed alone on a
few steps after the timed region; `traffic` is that kernel's measured HBM
bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
record in profiles/electron_kernel_traffic.json when present (labelled as
such), and `valu` its VALU issue from profiles/electron_kernel_valu.json.  The batched (plain per
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
VALU_RECORD = os.path.join(ROOT, "profiles", "electron_kernel_valu.json")
INVALID = np.iinfo(np.int64).min


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_c2_sample(tables, n_particles: int, steps: int, threads: int):
    """The CPU oracle (a restatement of the reference step, OpenMP over
    64-particle chunks) on a bounded sample of the at-scale workload:
    electrons, photons and positrons in the 70:25:5 ratio, each step followed
    by the newborn drain of its secondaries, material sources accumulated.
    Only the oracle's calls are timed (not the numpy photon push between
    steps).  Returns (particle-steps/s, seconds, particle-steps)."""
    from oracle import binding as ob
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig

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


def cpu_c1_run(tables, steps: int, threads: int):
    """The plain CPU configuration on the oracle's multi-step loop
    (oracle/driver.py: photon push, the three species steps, the drain, the
    commit of its survivors): 1,000 e- at 7.2 MeV along the force in
    E_z = -400 kV/m, air at 1.225 kg/m^3, 10 keV cutoff, dt = 1 ps, seed
    20240623.  The whole loop is timed (it is the CPU path's own cost).
    Returns (particle-steps/s, seconds, particle-steps, final live counts)."""
    from oracle import driver as od
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig

    fields, tile = wl.c1_plain(1000)
    sim = od.OracleSimulation(tables, fields, StepConfig(rng_seed=20240623, dt_s=1e-12), tile, nthreads=threads)
    done = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        done += sum(int((sim.sp[k]["id"] != INVALID).sum()) for k in sim.sp)
        sim.step()
    el = time.perf_counter() - t0
    live = {name: int((sim.sp[k]["id"] != INVALID).sum()) for name, k in
            (("electrons", od.ELECTRON), ("photons", od.PHOTON), ("positrons", od.POSITRON))}
    return done / el, el, done, live


def cpu_baseline(args):
    """The CPU baselines on this host's cores: the at-scale sample (the
    headline, same per-particle work as the GPU metric) with every host
    thread this process may use and with one thread, and the plain CPU
    configuration -- in full (20,000 steps) with every thread, a bounded
    2,000 steps with one thread."""
    from oracle import binding as ob
    from rrea_feedback_warpx_amrex_geant4_amd.transport import load_bundle

    ob.build()
    tables = ob.Tables(load_bundle())
    nproc = os.cpu_count() or 1
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    threads = max(1, min(16, allowed))
    rate, secs, done = cpu_c2_sample(tables, args.cpu_particles, args.cpu_steps, threads)
    log(f"cpu baseline: {rate:.4e} particle-steps/s on {threads} threads ({secs:.1f} s)")
    rate1, secs1, done1 = cpu_c2_sample(tables, args.cpu_particles_1t, args.cpu_steps_1t, 1)
    log(f"cpu baseline, 1 thread: {rate1:.4e} particle-steps/s ({secs1:.1f} s)")
    c1 = c1_1 = None
    if args.cpu_c1_steps > 0:
        r, el, dn, live = cpu_c1_run(tables, args.cpu_c1_steps, threads)
        c1 = {"value": r, "seconds": el, "particle_steps": dn, "steps": args.cpu_c1_steps, "threads": threads,
              "live_at_end": live}
        log(f"cpu plain configuration: {args.cpu_c1_steps} steps, {r:.4e} particle-steps/s on {threads} threads "
            f"({el:.1f} s)")
        r1, el1, dn1, _ = cpu_c1_run(tables, min(args.cpu_c1_steps, 2000), 1)
        c1_1 = {"value": r1, "seconds": el1, "particle_steps": dn1, "steps": min(args.cpu_c1_steps, 2000),
                "threads": 1}
        log(f"cpu plain configuration, 1 thread: {r1:.4e} particle-steps/s ({el1:.1f} s)")
    return {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_particles} at-scale particles (e-:gamma:e+ = 70:25:5) x {args.cpu_steps} steps "
                      f"({done} particle-steps, {secs:.1f} s), oracle restatement, OpenMP dynamic over 64-particle "
                      f"chunks, each step's newborn drain and material sources included",
            "nproc": nproc, "threads_allowed": allowed,
            "single_thread": {"value": rate1, "particles": args.cpu_particles_1t, "steps": args.cpu_steps_1t,
                              "seconds": secs1, "particle_steps": done1},
            "plain_cpu_configuration": c1, "plain_cpu_configuration_single_thread": c1_1}


def batched_leg(args, tables):
    """The batched configuration on the same device: 256 independent
    avalanche instances (|E| ~ U[250, 1000] kV/m, rho ~ U[0.3, 1.225] kg/m^3,
    N0 ~ log-U[1e3, 1e6] electrons at K ~ U[1, 10] MeV, 64 x 256 RZ grids of
    1 m cells), dt = 1 ps.  A step is the whole Apply of every instance --
    the photon push, the electron, photon and positron steps, the newborn
    drain -- then the commit of the drain's survivors into the stores
    (AddNParticles), so each instance avalanches: its newborn electrons,
    positrons and photons are transported from the next step on.  Timed on
    the same inputs three ways: the batched entry points driven from the
    host (simulation.BatchedAvalanche.step: one launch set per species and
    one drain for all instances), the same steps captured into one HIP graph
    (rrea_batch_graph_* with the stores' counts: commit inside the graph, no
    host round trip in a step), relaunched to the configuration's 500 steps,
    and the plain per-instance path (simulation.Simulation: the drop-in
    Apply + commit run instance after instance).  Particle-steps are the
    stores' live particles per step (the mean of the live counts before and
    after the timed steps, times the steps).  Returns the report object."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.simulation import BatchedAvalanche, Simulation

    n_inst = args.batched_instances
    fields, tiles, seeds = wl.c3_batch(n_inst, seed=3, scale=args.batched_scale)
    sizes = np.array([len(t.r) for t in tiles], dtype=np.int64)
    n = int(sizes.sum())
    log(f"batched: {n_inst} instances, {n} electrons (smallest {sizes.min()}, largest {sizes.max()})")
    dt = 1e-12
    cfg = StepConfig(rng_seed=0, dt_s=dt)
    steps_total = args.batched_warmup + args.batched_steps + args.graph_loop_steps
    # room for every newborn of the run: well above the ~1e-4 secondaries per
    # particle-step at 1 ps
    caps = (n + max(1 << 20, n // 8), max(1 << 18, n // 64), max(1 << 20, n // 16))
    sec_cap = max(1 << 15, n >> 6)

    def live(av):
        return sum(av.live())

    # host-driven batched steps; the batch's particle sort (instance, cell,
    # energy band) tuned on this workload first: its interval and band width
    tuned = None
    every, bits = args.batched_sort_every, -1
    block = 256
    if args.batched_autotune:
        # on a copy of the batch, so that the timed legs below all start from
        # the seeded state
        from rrea_feedback_warpx_amrex_geant4_amd.autotune import tune
        tav = BatchedAvalanche(tables, fields, tiles, seeds, cfg, capacity=caps, secondary_capacity=sec_cap)

        def tune_step(s, sort_now, energy_bits):
            if sort_now:
                tav.sort(energy_bits)
            tav.step()
        # intervals up to a fifth of the graph loop; the candidates are timed
        # one after another on the evolving avalanche, the never-sort one
        # first, after 100 steps of warm-up
        choice = tune(tune_step, sort_every=(0, 50, 100), energy_bits=(0, 6), unsorted_steps=100, warmup=100)
        every, bits = choice.sort_every, choice.energy_bits
        tuned = choice.as_dict()
        from rrea_feedback_warpx_amrex_geant4_amd.autotune import tune_block
        bchoice = tune_block(lambda s: tav.step(), tav.batch.set_block, steps=10)
        block = bchoice.block
        tuned["block"] = bchoice.as_dict()
        del tav
        torch.cuda.empty_cache()
        log(f"batched autotune: sort every {every} steps, {bits} band bits ({choice.ms_per_step:.3f} ms/step), "
            f"fast-kernel workgroups of {block} lanes ({', '.join(f'{b}: {t:.3f}' for b, t in bchoice.trials)})")
    bav = BatchedAvalanche(tables, fields, tiles, seeds, cfg, capacity=caps, secondary_capacity=sec_cap)
    bav.batch.set_block(block)
    for _ in range(args.batched_warmup):
        bav.step()
    torch.cuda.synchronize()
    b0 = live(bav)
    t0 = time.perf_counter()
    for k in range(args.batched_steps):
        if every > 0 and k % every == 0:
            bav.sort(bits)
        bav.step()
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    b1 = live(bav)
    rate_b = 0.5 * (b0 + b1) * args.batched_steps / el_b
    log(f"batched: {args.batched_steps} steps in {el_b:.3f} s, {rate_b:.4e} particle-steps/s "
        f"(live {b0} -> {b1}, stores {[s.n for s in bav.stores]})")

    # the same steps captured in one graph: its first launch (graph upload
    # included) is the warm-up, untimed
    gav = BatchedAvalanche(tables, fields, tiles, seeds, cfg, capacity=caps, secondary_capacity=sec_cap)
    gav.batch.set_block(block)
    del tiles
    t0 = time.perf_counter()
    graph = gav.graph(args.batched_steps, max_generations=args.graph_generations)
    capture_s = time.perf_counter() - t0
    graph.launch(0)
    torch.cuda.synchronize()
    graph.check_error()
    gav.sync_counts()
    g0 = live(gav)
    # the sort (outside the graph: hipcub clears its scratch with memsets)
    # before the launches its interval falls on, counted in the time
    sort_launches = max(1, -(-every // args.batched_steps)) if every > 0 else 0
    t0 = time.perf_counter()
    if sort_launches:
        gav.sort(bits)
    graph.launch(args.batched_steps)
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t0
    graph.check_error()
    gav.sync_counts()
    g1 = live(gav)
    rate_g = 0.5 * (g0 + g1) * args.batched_steps / el_g
    log(f"graph-captured batched: {args.batched_steps} steps in {el_g:.3f} s, {rate_g:.4e} particle-steps/s "
        f"(capture + instantiate {capture_s:.2f} s)")
    # the graph's loop continued to the configuration's 500 steps, relaunched back to back
    launches = max(1, args.graph_loop_steps // args.batched_steps)
    t0 = time.perf_counter()
    for i in range(1, launches):
        if sort_launches and i % sort_launches == 0:
            gav.sort(bits)
        graph.launch((i + 1) * args.batched_steps)
    torch.cuda.synchronize()
    el_loop = el_g + time.perf_counter() - t0
    graph.check_error()
    gav.sync_counts()
    g2 = live(gav)
    loop_steps = launches * args.batched_steps
    rate_loop = 0.5 * (g0 + g2) * loop_steps / el_loop
    log(f"graph-captured loop: {loop_steps} steps in {el_loop:.3f} s, {rate_loop:.4e} particle-steps/s "
        f"(live {g0} -> {g2}, stores {[s.n for s in gav.stores]})")
    grown = [s.n for s in gav.stores]
    graph.close()
    del gav
    torch.cuda.empty_cache()

    # the plain per-instance path: one Simulation per instance, stepped one after another
    if not args.batched_plain:
        return {"batched": {"value": rate_b, "ms_per_step": el_b / args.batched_steps * 1e3, "live": [b0, b1]},
                "graph_captured": {"value": rate_g, "ms_per_step": el_g / args.batched_steps * 1e3},
                "plain_per_instance": None}
    _, tiles, _ = wl.c3_batch(n_inst, seed=3, scale=args.batched_scale)
    sims = [Simulation(tables, f, StepConfig(rng_seed=sd, dt_s=dt), t, capacity=int(len(t.r)) + 4096,
                       secondary_capacity=max(1 << 10, len(t.r) >> 6), compact_every=0)
            for f, t, sd in zip(fields, tiles, seeds)]
    del tiles

    def live_plain():
        return sum(sum(s.live() for s in sim.stores) for sim in sims)

    for _ in range(args.batched_warmup):
        for sim in sims:
            sim.step()
    torch.cuda.synchronize()
    p0 = live_plain()
    t0 = time.perf_counter()
    for _ in range(args.batched_steps):
        for sim in sims:
            sim.step()
    torch.cuda.synchronize()
    el_p = time.perf_counter() - t0
    p1 = live_plain()
    rate_p = 0.5 * (p0 + p1) * args.batched_steps / el_p
    log(f"plain per instance: {args.batched_steps} steps in {el_p:.3f} s, {rate_p:.4e} particle-steps/s")
    # the host-driven batch and the plain runs agree on the seed electrons' first-step
    # outcomes only up to the ids of committed newborns (one counter per batch,
    # one per run), so the report compares live counts
    del sims
    torch.cuda.empty_cache()
    return {
        "workload": f"{n_inst} independent avalanche instances, {n} seed electrons (N0 log-uniform 1e3..1e6 x "
                    f"{args.batched_scale}), |E| U[250,1000] kV/m, rho U[0.3,1.225] kg/m^3, 64x256 RZ grids, "
                    "dt=1 ps; per step the photon push, the electron, photon and positron steps, the newborn "
                    "drain and the commit of its survivors (the avalanche), J per instance",
        "instances": n_inst, "seed_electrons": n, "steps": args.batched_steps, "warmup": args.batched_warmup,
        "autotune": tuned, "sort_every": every, "sort_energy_bits": bits, "block": block,
        "batched": {"value": rate_b, "ms_per_step": el_b / args.batched_steps * 1e3, "live": [b0, b1]},
        "plain_per_instance": {"value": rate_p, "ms_per_step": el_p / args.batched_steps * 1e3, "live": [p0, p1]},
        "graph_captured": {"value": rate_g, "ms_per_step": el_g / args.batched_steps * 1e3,
                           "steps_per_graph": args.batched_steps, "max_generations": args.graph_generations,
                           "capture_and_instantiate_s": capture_s,
                           "loop": {"steps": loop_steps, "launches": launches, "value": rate_loop,
                                    "ms_per_step": el_loop / loop_steps * 1e3, "live": [g0, g2],
                                    "store_rows_end": grown}},
        "ratio": rate_b / rate_p,
        "ratio_graph": rate_g / rate_p,
        "unit": "particle-steps/s",
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
    loop; particle-steps counted are the live particles at each step.  Then
    the configuration itself (``configuration_16``): 16 columns of E_z peaks
    300-600 kV/m holding 5e7 live macroparticles, stepped as a batch."""
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
    # the feedback configuration's batch at its size: 16 columns (E_z peaks
    # 300-600 kV/m) seeded with 5e7 live electrons in all, then injection
    # (--column-inject per instance per step), the batched species steps,
    # drain, commit, compaction, each instance's Maxwell feedback, and
    # population control every 10 steps at 1.25 x 3,125,000 electrons per
    # instance (the seeded level)
    from rrea_feedback_warpx_amrex_geant4_amd.column import BatchedColumn
    peaks = np.linspace(3.0e5, 6.0e5, 16)
    per = args.column_seed // 16
    bc = BatchedColumn(tables, peaks, inject_macros=args.column_inject, targets=(per, 1 << 18), control_every=10,
                       capacity=16 * per + (1 << 23))
    bc.seed_population(per)
    bc.run(2)
    torch.cuda.synchronize()
    live0 = bc.live()
    live = 0
    n_b = args.column_batched_steps
    t0 = time.perf_counter()
    for _ in range(n_b):
        live += bc.live()
        bc.step()
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t0
    live1 = bc.live()
    obs_b = bc.observe()
    bc.out[0].check_error()
    bc.close()
    del bc
    torch.cuda.empty_cache()
    log(f"batched columns: 16 instances, live {live0} -> {live1}, {n_b} steps in {el_b:.2f} s "
        f"({el_b / n_b * 1e3:.1f} ms/step), {live / el_b:.3e} particle-steps/s; electrons per instance "
        f"{obs_b['electrons']}")
    single["configuration_16"] = {
        "workload": f"16 columns (E_z peaks 300-600 kV/m, 256x1024, 4-14 km), {16 * per} seeded electrons, "
                    f"{args.column_inject} injected per instance per step, dt = 1 ns, batched e-/e+/gamma steps + "
                    "drain + commit + compaction + per-instance Maxwell feedback, control every 10 steps",
        "steps": n_b, "ms_per_step": el_b / n_b * 1e3, "particle_steps_per_s": live / el_b,
        "live": [live0, live1], "electrons_per_instance": obs_b["electrons"],
        "removed_electrons": obs_b["removed_electrons"]}
    return single


def valu_record():
    """The electron kernel's VALU issue from the committed rocprofv3 SQ
    counter record (profiles/electron_kernel_valu.json), or None: VALU
    instructions per electron-step, the share of the launch's SIMD cycles
    they occupy, and how a wave's cycles split between waiting on memory or
    LDS, waiting for issue and issuing."""
    try:
        with open(VALU_RECORD) as f:
            rec = json.load(f)
        return {k: rec[k] for k in ("valu_insts_per_particle_step", "valu_issue_frac", "waves_per_simd",
                                    "vgprs_compiler", "wave_cycle_split", "source") if k in rec}
    except (OSError, KeyError, ValueError):
        return None


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
    ap.add_argument("--block", type=int, default=0, choices=[0, 128, 256, 512],
                    help="fast kernels' workgroup size without the autotuner (0: 256, or the autotuner's choice)")
    ap.add_argument("--sort-every", type=int, default=25,
                    help="reorder the tiles by cell every this many steps (0: never), inside the timed region")
    ap.add_argument("--batched-autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="tune the batched leg's particle-sort interval and band bits first")
    ap.add_argument("--batched-plain", action=argparse.BooleanOptionalAction, default=True,
                    help="time the plain per-instance path on the batched inputs (the batched/plain ratio)")
    ap.add_argument("--batched-sort-every", type=int, default=0,
                    help="the batched leg's particle-sort interval without the autotuner (0: never)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick the sort interval and energy-band bits by timing candidates before the warm-up")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sequential", action="store_true",
                    help="run the three species passes one after another on one stream (default: concurrently)")
    ap.add_argument("--cpu-particles", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--cpu-particles-1t", type=int, default=1 << 16)
    ap.add_argument("--cpu-steps-1t", type=int, default=20)
    ap.add_argument("--cpu-c1-steps", type=int, default=20000,
                    help="steps of the plain CPU configuration timed on the CPU (0: skip)")
    ap.add_argument("--batched-instances", type=int, default=256, help="instances of the batched leg (0: skip it)")
    ap.add_argument("--batched-scale", type=float, default=1.0, help="scale of the batched leg's populations")
    ap.add_argument("--batched-steps", type=int, default=10, help="timed steps of the batched leg (>= 1)")
    ap.add_argument("--batched-warmup", type=int, default=2)
    ap.add_argument("--field-steps", type=int, default=200, help="timed feedback steps of the field leg (0: skip)")
    ap.add_argument("--early", nargs="*", default=["positron"], choices=["photon", "positron"],
                    help="species passes that start with the electron pass instead of after its fast kernel")
    ap.add_argument("--device-drain", type=int, default=4,
                    help="concurrent step: drain with the counts on the device, up to this many generations "
                         "(0: the host-driven drain); a FIFO still holding a generation after them is an error")
    ap.add_argument("--column-steps", type=int, default=100, help="timed steps of the feedback-column leg (0: skip)")
    ap.add_argument("--column-inject", type=int, default=2048, help="macroparticles injected per column step")
    ap.add_argument("--column-seed", type=int, default=50_000_000,
                    help="live electrons seeded over the 16 columns of the feedback configuration")
    ap.add_argument("--column-batched-steps", type=int, default=10,
                    help="timed steps of the 16-column feedback configuration")
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
    if args.batched_steps < 1 or args.batched_warmup < 0:
        ap.error("--batched-steps must be at least 1 and --batched-warmup not negative")

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
    # the three species in growing stores (room for the newborns committed
    # every step); the tiles a step transports are views of their rows
    from rrea_feedback_warpx_amrex_geant4_amd.simulation import ELECTRON, PHOTON, POSITRON, ParticleStore
    stores = [ParticleStore(ELECTRON, n_e + (1 << 20)), ParticleStore(POSITRON, n_p + (1 << 18)),
              ParticleStore(PHOTON, n_g + (1 << 20))]
    stores[ELECTRON].load(wl.c2_electrons(n_e, fields, seed=2))
    stores[POSITRON].load(wl.c2_electrons(n_p, fields, seed=3, first_id=1 << 40, charge_sign=1.0))
    g_host = wl.c2_photons(n_g, fields, dt_s=1e-12, seed=4, first_id=1 << 41)
    stores[PHOTON].load(g_host)
    if args.sequential:
        # the first step's chords run to the pushed points
        for k in ("r", "z", "theta"):
            getattr(stores[PHOTON], k)[:n_g].copy_(torch.as_tensor(g_host.extra[k + "_cur"]))
    # (otherwise the step's own push, before its photon pass, moves the
    # photons there)
    del g_host
    counts = torch.tensor([s.n for s in stores], dtype=torch.int64, device="cuda")
    next_id = torch.tensor([(1 << 42)], dtype=torch.int64, device="cuda")
    store_structs = (_lib.ParticleStore * 3)(*[s.struct() for s in stores])
    # every step transports the stores' whole capacity: rows past a store's
    # length hold the invalid id and are skipped (8 B read each), so the
    # commit grows the stores on the device with no host round trip
    dev_e, dev_g, dev_p = stores[ELECTRON].full_tile(), stores[PHOTON].full_photon_tile(), stores[POSITRON].full_tile()
    out_e = StepOutputs(n_e, fields, secondary_capacity=1 << 20, per_particle=False, sources=True)
    out_p = StepOutputs(n_p, fields, secondary_capacity=1 << 18, per_particle=False, share_grid=out_e)
    out_g = StepOutputs(n_g, fields, secondary_capacity=1 << 18, per_particle=False, share_grid=out_e)
    drain = DrainOutputs(1 << 21, out_e)
    L = _lib.lib()
    apply = RreaParticleInteraction
    conc = None if args.sequential else ConcurrentApply(tables, fields, early=tuple(args.early))
    # one list per dtype: torch's multi-tensor zero then takes one launch each
    step_clear = ([o.secondary_count for o in (out_e, out_g, out_p)], [out_e.jr, out_e.jz, out_e.rho, out_e.sources])
    # the handle each species' kernels are timed on
    h_e, h_g, h_p = (coupling.h,) * 3 if conc is None else tuple(c.h for c in conc.couplings)

    knobs = {"sort_every": args.sort_every, "energy_bits": -1, "first_step": 0, "block": args.block or 256}

    def one_step(s, sort_now, energy_bits):
        if sort_now:
            # WarpX's periodic particle sort: cell-ordered tiles
            dev_e.sort_by_cell(coupling, energy_bits=energy_bits)
            dev_p.sort_by_cell(coupling, energy_bits=energy_bits)
            dev_g.sort_by_cell(coupling, energy_bits=energy_bits)
        # the step's counters and grids cleared (one launch per dtype)
        for group in step_clear:
            torch._foreach_zero_(group)
        cfg = StepConfig(rng_seed=20240623, step=s, dt_s=1e-12)
        if conc is not None:
            # WarpX's photon push of this step on the photon stream, before
            # its pass, under the electron hand-over
            conc.run(dev_e, dev_g, dev_p, cfg, out_e, out_g, out_p, drain, device_drain=args.device_drain,
                     pre_push=True)
        else:
            apply.apply_electrons(tables, coupling, dev_e, cfg, out_e, order_secondaries=True, check=False)
            apply.apply_photons(tables, coupling, dev_g, cfg, out_g, order_secondaries=True, check=False)
            apply.apply_positrons(tables, coupling, dev_p, cfg, out_p, order_secondaries=True, check=False)
            apply.drain_secondaries(tables, coupling, cfg, [out_e, out_g, out_p], drain, check=False)
        # AddNParticles of the drain's survivors (RreaSecondaryCommit.cpp:594-629);
        # the sequential step then pushes the photons for the next step,
        # newborn photons included (the concurrent one pushes them before its photon pass)
        _lib.check(L.rrea_commit_survivors(drain.survivors.data_ptr(), drain.survivor_count.data_ptr(),
                                           store_structs, counts.data_ptr(), next_id.data_ptr(), 0,
                                           out_e.error.data_ptr(), torch.cuda.current_stream().cuda_stream),
                   "rrea_commit_survivors")
        if conc is None:
            dev_g.push(cfg.dt_s)

    def run_step(s, sort_now):
        # the RNG step index continues past the tuning steps
        one_step(knobs["first_step"] + s, sort_now, knobs["energy_bits"])

    def live():
        return tuple(int((t.id != _lib.INVALID_ID).sum().item()) for t in (dev_e, dev_g.cur, dev_p))

    tuned = None
    if args.autotune:
        from rrea_feedback_warpx_amrex_geant4_amd.autotune import tune
        log("autotune: tile-sort interval x energy-band bits")
        choice = tune(one_step, first_step=0)
        knobs["sort_every"], knobs["energy_bits"] = choice.sort_every, choice.energy_bits
        tuned = choice.as_dict()
        log(f"autotune: sort every {choice.sort_every} steps, {choice.energy_bits} band bits "
            f"({choice.ms_per_step:.3f} ms/step)")
        # then the fast kernels' workgroup size, on tiles freshly sorted with
        # the chosen band width
        from rrea_feedback_warpx_amrex_geant4_amd.autotune import tune_block
        s0 = choice.steps_run
        one_step(s0, True, choice.energy_bits)
        bchoice = tune_block(lambda s: one_step(s, False, choice.energy_bits),
                             conc.set_block if conc is not None else coupling.set_block, first_step=s0 + 1)
        knobs["first_step"] = s0 + 1 + bchoice.steps_run
        knobs["block"] = bchoice.block
        tuned["block"] = bchoice.as_dict()
        log(f"autotune: fast-kernel workgroups of {bchoice.block} lanes "
            f"({', '.join(f'{b}: {t:.3f}' for b, t in bchoice.trials)} ms/step)")
    # the workgroup size on every handle the step and the isolated timing use
    coupling.set_block(knobs["block"])
    if conc is not None:
        conc.set_block(knobs["block"])
    log(f"warmup {args.warmup}")
    every = knobs["sort_every"]
    for s in range(args.warmup):
        run_step(s, every > 0 and s == 0)
    torch.cuda.synchronize()
    out_e.check_error()
    e0, g0, p0 = live()

    for h in {h_e.value, h_g.value, h_p.value}:
        L.rrea_kernel_timing(h, 1)
    log(f"timed {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # the timed region opens with a tile sort and sorts every `every` steps
    # after it, so it holds ceil(steps / every) whole sorts: at least its
    # amortised share of them
    sorts_timed = 0
    for s in range(args.steps):
        sort_now = every > 0 and s % every == 0
        sorts_timed += int(sort_now)
        run_step(args.warmup + s, sort_now)
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
        cpu = cpu_baseline(args)

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
                   "species_passes": "sequential" if args.sequential else
                   ("concurrent (three streams): " + (", ".join(args.early) or "no pass") + " starting with the "
                    "electron pass, the others after the electron fast kernel"),
                   "drain": (f"device-driven, {args.device_drain} generations" if args.device_drain > 0 and not args.sequential
                             else "host-driven"),
                   "tile_sort_every": knobs["sort_every"], "tile_sort_energy_bits": knobs["energy_bits"],
                   "fast_kernel_block": knobs["block"],
                   "tile_sorts_in_timed_region": sorts_timed,
                   "autotune": tuned,
                   "kernel_ms_per_step": {k: v[0] / max(v[1], 1) for k, v in kms.items()}},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": ("committed rocprofv3 FETCH_SIZE / WRITE_SIZE record of this kernel on this "
                                        "workload (profiles/electron_kernel_traffic.json), not measured in this run"
                                        if traffic is not None else None),
                     "valu": valu_record(),
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
