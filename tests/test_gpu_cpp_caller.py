"""The drop-in layout: tiles whose pre-push snapshot (rrea_prev_*) and live
arrays are different buffers, as CaptureRreaPositionsBeforePush leaves them
(warpx_overlay/Source/Rrea/RreaWarpXCoupling.cpp:151-208), driven from Python
and from a compiled C++ caller of include/rrea_hip.h.

1. Through the Python host path, a step over separate prev_* and live arrays
   equals the fused in-place step on every written-back particle (survivors
   and escapes) bit for bit; it leaves the live arrays of every other particle
   and every prev_* array untouched, and invalidates the same ids.
2. The C++ caller (tests/cpp/rrea_caller.cpp, built with the library) runs the
   integration sequence: tables from the bundle's columns, fields and the
   density profile, the electron, photon and positron steps in Apply's order,
   then the staged newborn drain.  On the same inputs it equals the Python
   path: states, ids, statuses, substeps and ion pairs bit for bit;
   secondaries, the drain's FIFO and its survivors record for record; the
   tallies and the J grid to FP64 summation order."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CALLER = os.path.join(ROOT, "tests", "cpp", "rrea_caller")
KEYS = ("r", "z", "theta", "ux", "uy", "uz")
N_E, N_G, N_P = 4096, 2048, 1024
DT = 1e-8
CAP = 1 << 15
FIFO = 1 << 17
GENERATIONS = 16


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def live_of(tile, seed):
    """Live arrays that differ from the snapshot (as WarpX's push leaves them)."""
    rng = np.random.default_rng(seed)
    return {"r": tile.r * (1.0 + 1e-3 * rng.random(len(tile.r))), "z": tile.z + rng.random(len(tile.z)),
            "theta": tile.theta * 0.5, "ux": tile.ux * 1.01, "uy": tile.uy * 0.99, "uz": tile.uz * 1.02}


def inputs():
    fields = wl.c2_fields()
    e = wl.c2_electrons(N_E, fields, seed=2)
    g = wl.c2_photons(N_G, fields, dt_s=DT, seed=4, first_id=1 << 41)
    p = wl.c2_electrons(N_P, fields, seed=3, first_id=1 << 40, charge_sign=1.0)
    cfg = StepConfig(rng_seed=77, dt_s=DT, step=3)
    return fields, e, g, p, cfg


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def split_tiles(e, live):
    """(snapshot tile, live tile) of one charged species; w, id and cpu are the snapshot's."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile
    prev = DeviceTile.from_host(e)
    cur = prev.clone()
    for k in KEYS:
        setattr(cur, k, dev(live[k]))
    return prev, cur


def python_run(tables, fields, e, g, p, cfg):
    """The integration sequence through the Python host path, separate arrays."""
    import ctypes as C

    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (Coupling, DrainOutputs, PhotonTile,
                                                                RreaParticleInteraction, StepOutputs, _step_config)
    coupling = Coupling(fields, tables)
    pe, ce = split_tiles(e, live_of(e, 5))
    pp, cp = split_tiles(p, live_of(p, 6))
    tg = PhotonTile.from_host(g)
    oe = StepOutputs(N_E, fields, secondary_capacity=CAP)
    og = StepOutputs(N_G, fields, secondary_capacity=CAP, share_grid=oe)
    op = StepOutputs(N_P, fields, secondary_capacity=CAP, share_grid=oe)
    apply = RreaParticleInteraction
    apply.apply_electrons(tables, coupling, pe, cfg, oe, current=ce, order_secondaries=2)
    apply.apply_photons(tables, coupling, tg, cfg, og, 2)
    apply.apply_positrons(tables, coupling, pp, cfg, op, current=cp, order_secondaries=2)
    d = DrainOutputs(FIFO, oe)
    outs = (oe, og, op)
    recs = (C.c_void_p * 3)(*[o.secondaries.data_ptr() for o in outs])
    cnts = (C.c_void_p * 3)(*[o.secondary_count.data_ptr() for o in outs])
    caps = (C.c_int64 * 3)(*[o.secondary_capacity for o in outs])
    io = d.struct(0)
    c = _step_config(cfg)
    _lib.check(_lib.lib().rrea_newborn_drain_staged(tables.h, coupling.h, C.byref(c), recs, cnts, caps, 3,
                                                    C.byref(io), GENERATIONS, torch.cuda.current_stream().cuda_stream),
               "rrea_newborn_drain_staged")
    torch.cuda.synchronize()
    oe.check_error()
    return {"electron": (pe, ce, oe), "photon": (tg, tg.cur, og), "positron": (pp, cp, op), "drain": d,
            "grid": oe}


def test_separate_arrays_equal_fused_step(tables):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling, DeviceTile, RreaParticleInteraction, StepOutputs
    fields, e, _, p, cfg = inputs()
    print(f"case separate prev/live arrays: {N_E} e-, {N_P} e+, one 10 ns step", flush=True)
    coupling = Coupling(fields, tables)
    for species, tile, apply, seed in (("electron", e, RreaParticleInteraction.apply_electrons, 5),
                                       ("positron", p, RreaParticleInteraction.apply_positrons, 6)):
        live = live_of(tile, seed)
        fused = DeviceTile.from_host(tile)
        of = StepOutputs(len(tile.r), fields, secondary_capacity=CAP)
        apply(tables, coupling, fused, cfg, of, order_secondaries=2)
        prev, cur = split_tiles(tile, live)
        os_ = StepOutputs(len(tile.r), fields, secondary_capacity=CAP)
        apply(tables, coupling, prev, cfg, os_, current=cur, order_secondaries=2)
        torch.cuda.synchronize()
        st = of.status.cpu().numpy()
        assert np.array_equal(st, os_.status.cpu().numpy())
        assert np.array_equal(fused.id.cpu().numpy(), prev.id.cpu().numpy())
        written = np.isin(st, (0, 1))
        assert written.sum() > 0.9 * len(st) and (~written).sum() > 0, (written.sum(), len(st))
        for k in KEYS:
            a, b = getattr(fused, k).cpu().numpy(), getattr(cur, k).cpu().numpy()
            assert np.array_equal(a[written], b[written]), (species, k)
            assert np.array_equal(b[~written], live[k][~written]), (species, k)  # live arrays left alone
            assert np.array_equal(getattr(prev, k).cpu().numpy(), getattr(tile, k)), (species, k)  # snapshot read only
        sa, sb = of.secondaries_host(), os_.secondaries_host()
        assert len(sa) == len(sb) and (sa == sb).all()
        assert np.array_equal(of.ion_pairs.cpu().numpy(), os_.ion_pairs.cpu().numpy())
        print(f"{species}: {int(written.sum())} written back, {int((~written).sum())} not, "
              f"{len(sa)} secondaries -- equal to the fused step", flush=True)


def write(d, name, a, dtype):
    np.ascontiguousarray(a, dtype=dtype).tofile(os.path.join(d, name + {np.float64: ".f8", np.int64: ".i8",
                                                                         np.int32: ".i4"}[dtype]))


def test_cpp_caller_equals_python_path(tables, bundle):
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.transport import RATE_KEYS, SECONDARY_DTYPE
    if not os.access(CALLER, os.X_OK):
        pytest.fail(f"{CALLER} is not built (python -m rrea_feedback_warpx_amrex_geant4_amd.build)")
    fields, e, g, p, cfg = inputs()
    print(f"case C++ caller: {N_E} e-, {N_G} gamma, {N_P} e+, one 10 ns step and its drain", flush=True)
    ref = python_run(tables, fields, e, g, p, cfg)
    with tempfile.TemporaryDirectory() as tmp:
        din, dout = os.path.join(tmp, "in"), os.path.join(tmp, "out")
        os.makedirs(din)
        os.makedirs(dout)
        # the bundle's columns
        for k in RATE_KEYS:
            write(din, k + "_x", bundle[k + "_x"], np.float64)
            write(din, k + "_y", bundle[k + "_y"], np.float64)
            write(din, k + "_mode", [int(bundle[k + "_mode"])], np.int64)
        for k in ("moller", "bhabha", "ebrems", "pbrems"):
            for s, t in (("_primary", np.float64), ("_start", np.int64), ("_u", np.float64), ("_v", np.float64)):
                write(din, k + s, bundle[k + s], t)
        for k in ("gs_e", "gs_p", "wz_e", "wz_p"):
            for s in ("_e", "_n", "_u", "_omc"):
                write(din, k + s, bundle[k + s], np.float64)
        for k in ("el_e", "el_p"):
            for s in ("_e", "_mfp", "_eta"):
                write(din, k + s, bundle[k + s], np.float64)
        write(din, "scalars", bundle["scalars"], np.float64)
        for k, t in (("pe_primary", np.float64), ("pe_start", np.int64), ("pe_u", np.float64),
                     ("pe_binding", np.float64), ("pair_block_channel", np.int64), ("pair_block_energy", np.float64),
                     ("pair_block_start", np.int64), ("pair_frac", np.float64), ("pair_dir", np.float64)):
            write(din, k, bundle[k], t)
        # fields, density profile, configuration
        for k in ("er", "ez", "amb_er", "amb_ez", "btheta"):
            write(din, k, getattr(fields, k), np.float64)
        write(din, "profile_alt", fields.profile_alt_m, np.float64)
        write(din, "profile_rho", fields.profile_rho, np.float64)
        meta = {"nr": fields.nr, "nz": fields.nz, "dr": fields.dr, "dz": fields.dz, "z_lo": fields.z_lo,
                "altitude_at_z0_m": fields.altitude_at_z0_m, "reference_density": tables.reference_density,
                "rng_seed": cfg.rng_seed, "step": cfg.step, "dt_s": cfg.dt_s,
                "low_energy_cutoff_eV": cfg.low_energy_cutoff_eV, "photon_cutoff_eV": cfg.photon_cutoff_eV,
                "max_optical_depth_per_substep": cfg.max_optical_depth_per_substep,
                "max_subcycles": cfg.max_subcycles, "strict_transport_guards": int(cfg.strict_transport_guards),
                "max_gs_collisions_per_substep": cfg.transport_max_gs_collisions_per_substep,
                "gs_total_collision_cap": cfg.transport_gs_total_collision_cap,
                "max_field_impulse_fraction": cfg.transport_max_field_impulse_fraction,
                "field_impulse_floor_eV": cfg.transport_field_impulse_floor_eV,
                "secondary_capacity": CAP, "fifo_capacity": FIFO, "max_generations": GENERATIONS}
        with open(os.path.join(din, "meta.txt"), "w") as f:
            for k, v in meta.items():
                f.write(f"{k} {float(v)!r}\n")
        # the tiles: snapshot and live arrays apart
        for name, tile, live in (("electron", e, live_of(e, 5)), ("positron", p, live_of(p, 6))):
            for k in KEYS:
                write(din, f"{name}_prev_{k}", getattr(tile, k), np.float64)
                write(din, f"{name}_{k}", live[k], np.float64)
            write(din, f"{name}_w", tile.w, np.float64)
            write(din, f"{name}_id", tile.id, np.int64)
            write(din, f"{name}_cpu", tile.cpu, np.int32)
        for k in ("r", "z", "theta"):
            write(din, f"photon_prev_{k}", getattr(g, k), np.float64)
            write(din, f"photon_{k}", g.extra[k + "_cur"], np.float64)
        for k in ("ux", "uy", "uz", "w"):
            write(din, f"photon_{k}", getattr(g, k), np.float64)
        write(din, "photon_id", g.id, np.int64)
        write(din, "photon_cpu", g.cpu, np.int32)
        run = subprocess.run([CALLER, din, dout], capture_output=True, text=True, timeout=120)
        print(run.stdout.strip(), run.stderr.strip()[-2000:], flush=True)
        assert run.returncode == 0, run.stderr

        def rd(name, dtype):
            return np.fromfile(os.path.join(dout, name), dtype=dtype)

        for name in ("electron", "photon", "positron"):
            prev, cur, out = ref[name]
            for k in KEYS:
                assert np.array_equal(rd(f"{name}_{k}.f8", np.float64), getattr(cur, k).cpu().numpy()), (name, k)
            for k in ("r", "z", "theta"):
                pk = getattr(prev, f"prev_{k}") if name == "photon" else getattr(prev, k)
                assert np.array_equal(rd(f"{name}_prev_{k}.f8", np.float64), pk.cpu().numpy()), (name, k)
            ids = (prev.cur.id if name == "photon" else prev.id).cpu().numpy()
            assert np.array_equal(rd(f"{name}_id.i8", np.int64), ids), name
            assert np.array_equal(rd(f"{name}_status.i4", np.int32), out.status.cpu().numpy().astype(np.int32))
            assert np.array_equal(rd(f"{name}_substeps.i4", np.int32), out.substeps.cpu().numpy())
            assert np.array_equal(rd(f"{name}_ion_pairs.f8", np.float64), out.ion_pairs.cpu().numpy())
            sc, sp = rd(f"{name}_secondaries.bin", SECONDARY_DTYPE), out.secondaries_host()
            assert len(sc) == len(sp) and (sc == sp).all(), name
        d = ref["drain"]
        wc, wp = rd("work.bin", SECONDARY_DTYPE), d.work_host()
        sc, sp = rd("survivors.bin", SECONDARY_DTYPE), d.survivors_host()
        assert len(wc) == len(wp) and (wc == wp).all()
        assert len(sc) == len(sp) > 0 and (sc == sp).all()
        assert (wp["generation"] > 0).any()
        grid = ref["grid"]
        assert rd("error.i4", np.int32).tolist() == [0, 0]
        tc, tp = rd("tallies.f8", np.float64), grid.tallies.cpu().numpy()
        assert len(tc) == _lib.TALLY_COUNT
        assert np.allclose(tc, tp, rtol=1e-12, atol=0.0), np.flatnonzero(~np.isclose(tc, tp, rtol=1e-12, atol=0.0))
        for k in ("jr", "jz", "rho"):
            a, b = rd(f"{k}.f8", np.float64), getattr(grid, k).cpu().numpy()
            assert np.allclose(a, b, rtol=0.0, atol=1e-12 * float(np.abs(b).max())), k
        print(f"C++ caller: {len(wp)} FIFO entries, {len(sp)} survivors -- equal to the Python path", flush=True)
