"""Shared helpers of the GPU parity tests: run one step on the HIP path and on
the CPU oracle from identical inputs and compare them."""
from __future__ import annotations

import numpy as np

from oracle import binding as ob

# Per-particle continuous outputs must agree to 1e-12 relative (FP64); the
# absolute floor covers values that cancel to ~0 (r on the axis, u components
# orthogonal to the motion), scaled by the quantity's natural size.
REL = 1e-12


def close(a, b, scale):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tol = REL * np.maximum(np.maximum(np.abs(a), np.abs(b)), scale)
    return np.abs(a - b) <= tol


def gpu_step(tables, coupling, tile_dev, cfg, out):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import RreaParticleInteraction
    import torch

    out.reset()
    RreaParticleInteraction.apply_electrons(tables, coupling, tile_dev, cfg, out)  # fused in-place mode
    torch.cuda.synchronize()


def compare_step(fields, tile_host_prev, oracle_out, tile_dev, out, label=""):
    """Assert the HIP outputs of one step equal the oracle's."""
    status_o = oracle_out["status"]
    status_g = out.status.cpu().numpy().astype(np.int32)
    bad = np.flatnonzero(status_o != status_g)
    assert bad.size == 0, f"{label} status differs at {bad[:10]}: oracle {status_o[bad[:10]]} gpu {status_g[bad[:10]]}"
    c = 299792458.0
    scales = {"r": fields.dr, "z": fields.dz, "theta": 1.0, "ux": c, "uy": c, "uz": c}
    for k, sc in scales.items():
        g = getattr(tile_dev, k).cpu().numpy()
        o = oracle_out[k]
        ok = close(g, o, sc * 1e-3)
        idx = np.flatnonzero(~ok)
        assert idx.size == 0, (f"{label} {k} differs at {idx[:5]}: oracle {o[idx[:5]]} gpu {g[idx[:5]]} "
                               f"status {status_o[idx[:5]]}")
    ids = tile_dev.id.cpu().numpy()
    invalidated = np.isin(status_o, (1, 2, 3))
    assert np.all(ids[invalidated] == np.iinfo(np.int64).min), f"{label} invalidated ids not marked"
    assert np.all(ids[~invalidated] == tile_host_prev.id[~invalidated]), f"{label} live ids changed"
    ion_g = out.ion_pairs.cpu().numpy()
    ok = close(ion_g, oracle_out["ion_pairs"], 1e-3)
    assert ok.all(), f"{label} ion pairs differ at {np.flatnonzero(~ok)[:5]}"
    sub_g = out.substeps.cpu().numpy()
    assert np.array_equal(sub_g, oracle_out["substeps"].astype(np.int32)), f"{label} substeps differ"
    # secondaries: same count, same (parent, lineage) order, same payload
    sg = out.secondaries_host()
    so = oracle_out["secondaries"]
    assert len(sg) == len(so), f"{label} secondary count {len(sg)} vs {len(so)}"
    for f in ("parent_index", "species", "process", "root_key", "lineage"):
        assert np.array_equal(sg[f], so[f]), f"{label} secondary {f} differs"
    for f, sc in (("x", 1.0), ("y", 1.0), ("z", 1.0), ("dir_x", 1.0), ("dir_y", 1.0), ("dir_z", 1.0),
                  ("weight", 1.0), ("energy_eV", 1.0), ("birth_fraction", 1.0)):
        assert close(sg[f], so[f], sc).all(), f"{label} secondary {f} differs"
    # the net-chord grid: FP64 atomics sum in a different order
    for k in ("jr", "jz", "rho"):
        g = getattr(out, k).cpu().numpy()
        o = oracle_out[k]
        scale = max(np.abs(o).max(), 1e-300)
        assert np.abs(g - o).max() <= 1e-11 * scale, f"{label} {k} differs: {np.abs(g - o).max()} of {scale}"
    t_g = out.tallies.cpu().numpy()[:10]
    t_o = oracle_out["tallies"][:10]
    assert np.allclose(t_g, t_o, rtol=1e-10, atol=1e-300), f"{label} tallies {t_g} vs {t_o}"


def host_tile_from_dev(tile_dev):
    from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile

    return ParticleTile(*(getattr(tile_dev, k).cpu().numpy().copy() for k in
                          ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))


def run_and_compare(otables, tables, coupling, fields, tile_host, cfg, steps=1, label="", nthreads=8):
    """Advance `steps` steps on both paths from the same start, comparing every step."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile, StepOutputs

    tile_dev = DeviceTile.from_host(tile_host)
    out = StepOutputs(len(tile_host), fields, secondary_capacity=max(4096, len(tile_host)))
    host = tile_host
    for s in range(steps):
        cfg.step = s
        o = ob.electron_step(otables, fields, cfg, host.as_dict(), nthreads=nthreads)
        assert o["code"] == 0, f"oracle error {o['code']}"
        gpu_step(tables, coupling, tile_dev, cfg, out)
        compare_step(fields, host, o, tile_dev, out, label=f"{label} step {s}")
        # continue both from the oracle's state, keeping survivors only
        keep = o["status"] == 0
        host = host.take(np.flatnonzero(keep))
        for k in ("r", "z", "theta", "ux", "uy", "uz"):
            getattr(host, k)[:] = o[k][keep]
        if len(host) == 0:
            break
        tile_dev = DeviceTile.from_host(host)
        out = StepOutputs(len(host), fields, secondary_capacity=max(4096, len(host)))
    return host
