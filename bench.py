"""Benchmark of the RREA transport step on one MI355X.

Workload: the charged electron share of the at-scale configuration -- 2^24
macroparticles, of which 70 % (11,744,051) are electrons with K ~ Exp(7.3 MeV)
clipped to [10 keV, 100 MeV] in a 30-degree cone about the force, on a
256 x 1024 RZ mesh over 2.56 km x 10 km with rho = 1.225 exp(-h / 8.4 km),
h = 4 km + z, and |E_z| piecewise-linear 0 -> 500 kV/m; dt = 1 ps; FP64;
net-chord current deposition on; secondaries staged and ordered each step.

One "step" is one transport pass of the HIP electron kernel over the whole
tile, in place (the fused pre-push snapshot), followed by the deterministic
secondary ordering.  Particles the step invalidates stay in the tile with the
invalid id and are skipped by later steps, as WarpX keeps them until its
redistribute.

Prints one JSON line (rank 0, N=1):
  metric / value (particle-steps/s of live electrons) / roofline / cpu_baseline ...
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(n_particles: int, steps: int, threads: int):
    """The CPU oracle (a restatement of the reference step, OpenMP over
    64-particle chunks) timed on this host on a bounded sample."""
    from oracle import binding as ob
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.transport import load_bundle

    ob.build()
    tables = ob.Tables(load_bundle())
    fields = wl.c2_fields()
    tile = wl.c2_electrons(n_particles, fields, seed=2)
    p = tile.as_dict()
    done = 0
    t0 = time.perf_counter()
    for s in range(steps):
        cfg = StepConfig(rng_seed=20240623, step=s, dt_s=1e-12)
        o = ob.electron_step(tables, fields, cfg, p, nthreads=threads, sec_cap=1 << 20)
        if o["code"] != 0:
            raise RuntimeError(f"oracle error {o['code']}")
        live = p["id"] != np.iinfo(np.int64).min
        done += int(live.sum())
        p = dict(p)
        for k in ("r", "z", "theta", "ux", "uy", "uz"):
            p[k] = o[k]
        p["id"] = np.where(np.isin(o["status"], (1, 2, 3)), np.iinfo(np.int64).min, p["id"])
    dt = time.perf_counter() - t0
    return done / dt, dt, done


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--particles", type=int, default=None, help="electrons (default: 70%% of 2^24)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=12)
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

    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DeviceTile, InteractionTables,
                                                                RreaParticleInteraction, StepOutputs)

    assert torch.cuda.is_available(), "bench needs a GPU"
    torch.cuda.set_device(0)
    total = 1 << 24
    n_e, n_g, n_p = wl.c2_species_counts(total)
    n = args.particles if args.particles else n_e
    log(f"setup: {n} electrons (C2 electron share), building tables")
    tables = InteractionTables()
    fields = wl.c2_fields()
    coupling = Coupling(fields, tables)
    tile = wl.c2_electrons(n, fields, seed=2)
    dev = DeviceTile.from_host(tile)
    del tile
    out = StepOutputs(n, fields, secondary_capacity=1 << 20, per_particle=False)
    L = _lib.lib()

    def run_step(s):
        out.secondary_count.zero_()
        out.jr.zero_()
        out.jz.zero_()
        out.rho.zero_()
        cfg = StepConfig(rng_seed=20240623, step=s, dt_s=1e-12)
        RreaParticleInteraction.apply_electrons(tables, coupling, dev, cfg, out, order_secondaries=True, check=False)

    log(f"warmup {args.warmup}")
    for s in range(args.warmup):
        run_step(s)
    torch.cuda.synchronize()
    out.check_error()
    live_before = int((dev.id != _lib.INVALID_ID).sum().item())

    import ctypes as C
    L.rrea_kernel_timing(coupling.h, 1)
    log(f"timed {args.steps} steps")
    live_counts = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        live_counts.append(None)
        run_step(args.warmup + s)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    kms = C.c_double(0.0)
    launches = C.c_int64(0)
    _lib.check(L.rrea_kernel_timing_read(coupling.h, C.byref(kms), C.byref(launches)), "timing read")
    L.rrea_kernel_timing(coupling.h, 0)
    out.check_error()
    live_after = int((dev.id != _lib.INVALID_ID).sum().item())
    # live particles decline monotonically; the mean of the ends bounds the
    # processed count within the (tiny) per-step loss
    particle_steps = 0.5 * (live_before + live_after) * args.steps
    value = particle_steps / elapsed
    kernel_s = kms.value * 1e-3
    achieved = particle_steps * B_ALG / kernel_s
    log(f"live {live_before} -> {live_after}; wall {elapsed:.3f}s; kernel {kernel_s:.3f}s over {launches.value}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: oracle, {threads} threads, {args.cpu_particles} e- x {args.cpu_steps} steps")
        rate, secs, done = cpu_baseline(args.cpu_particles, args.cpu_steps, threads)
        cpu = {"value": rate, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_particles} C2 electrons x {args.cpu_steps} steps ({done} particle-steps, "
                         f"{secs:.1f} s), oracle restatement, OpenMP dynamic over 64-particle chunks"}

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
        "config": {"workload": "C2 plain at scale, electron share: 11,744,051 e- of 2^24, dt=1 ps, RZ 256x1024, "
                               "deposit on, secondaries ordered", "particles": n, "path": "plain"},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
