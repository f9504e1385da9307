"""Shared helpers of the GPU parity tests: run one step on the HIP path and on
the CPU oracle from identical inputs and compare them."""
from __future__ import annotations

import numpy as np

from oracle import binding as ob

# Per-particle continuous outputs must agree to 1e-12 relative (FP64); the
# absolute floor covers values that cancel to ~0 (r on the axis, u components
# orthogonal to the motion), scaled by the quantity's natural size.
REL = 1e-12
EPS = np.finfo(np.float64).eps
ME_C2 = 510998.95
C_LIGHT = 299792458.0


def conditioning(tile_prev, dt, substeps=None, fields=None, status=None):
    """Amplification of a last-place libm difference by the reference's
    Cartesian -> (r, theta) -> Cartesian round trip of every chord
    (RreaParticleInteraction.H:144-155 -> RreaTransportPath.cpp:109-114): the
    chord start and end carry ~ulp(r) each, so the chord's length and direction
    carry ~eps r / l relative, l = speed * dt."""
    u2 = tile_prev.ux ** 2 + tile_prev.uy ** 2 + tile_prev.uz ** 2
    gamma = np.sqrt(1.0 + u2 / C_LIGHT ** 2)
    speed = C_LIGHT * np.sqrt(np.maximum(0.0, 1.0 - 1.0 / (gamma * gamma)))
    ell = np.maximum(speed * dt, 1e-300)
    if substeps is not None:
        ell = ell / np.maximum(substeps, 1)  # each microchord is dt / substeps long
    kappa = 1.0 + np.maximum(tile_prev.r, np.abs(tile_prev.z)) / ell
    if fields is not None and status is not None:
        # an exiting chord's in-domain fraction is a root of r^2 - R^2 (or a
        # ratio (z_face - z) / dz): it carries R / (R - r) (or |z| / (z_face - z))
        esc = status == 1
        gap_r = np.maximum(fields.r_hi - tile_prev.r, 1e-300)
        gap_z = np.maximum(np.minimum(tile_prev.z - fields.z_lo, fields.z_hi - tile_prev.z), 1e-300)
        zscale = max(abs(fields.z_lo), abs(fields.z_hi))
        kappa = kappa + np.where(esc, fields.r_hi / gap_r + zscale / gap_z, 0.0)
    return kappa


def close(a, b, scale):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    tol = REL * np.maximum(np.maximum(np.abs(a), np.abs(b)), scale)
    return np.abs(a - b) <= tol


def gpu_step(tables, coupling, tile_dev, cfg, out, species="electron"):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import RreaParticleInteraction
    import torch

    out.reset()
    apply = RreaParticleInteraction.apply_positrons if species == "positron" else RreaParticleInteraction.apply_electrons
    apply(tables, coupling, tile_dev, cfg, out)  # fused in-place mode
    torch.cuda.synchronize()


def compare_step(fields, tile_host_prev, oracle_out, tile_dev, out, label="", dt=None, stats=None):
    """Assert the HIP outputs of one step equal the oracle's."""
    status_o = oracle_out["status"]
    status_g = out.status.cpu().numpy().astype(np.int32)
    bad = np.flatnonzero(status_o != status_g)
    assert bad.size == 0, f"{label} status differs at {bad[:10]}: oracle {status_o[bad[:10]]} gpu {status_g[bad[:10]]}"
    # The state is compared as vectors: the proper velocity to 1e-12 of |u|, the
    # Cartesian position (r cos theta, r sin theta, z) to 1e-12 of its
    # coordinates (a coordinate carries FP64 precision relative to its own size;
    # theta alone is ill-conditioned near the axis).
    g = {k: getattr(tile_dev, k).cpu().numpy() for k in ("r", "z", "theta", "ux", "uy", "uz")}
    o = oracle_out
    du = np.sqrt((g["ux"] - o["ux"]) ** 2 + (g["uy"] - o["uy"]) ** 2 + (g["uz"] - o["uz"]) ** 2)
    umag = np.sqrt(o["ux"] ** 2 + o["uy"] ** 2 + o["uz"] ** 2)
    rel_u = du / np.maximum(umag, 1.0)
    # tolerance: 1e-12, plus the round-trip conditioning times the step's own
    # relative change of u (the part of the state the chord length feeds)
    kappa = conditioning(tile_host_prev, dt, oracle_out["substeps"], fields, status_o)
    du_step = np.sqrt((o["ux"] - tile_host_prev.ux) ** 2 + (o["uy"] - tile_host_prev.uy) ** 2
                      + (o["uz"] - tile_host_prev.uz) ** 2) / np.maximum(umag, 1.0)
    tol_u = REL + 16.0 * EPS * kappa * du_step
    bad = np.flatnonzero(rel_u > tol_u)
    assert bad.size == 0, (f"{label} u differs at {bad[:5]}: |du|/|u| = {rel_u[bad[:5]]} tol {tol_u[bad[:5]]} "
                           f"status {status_o[bad[:5]]}")
    if stats is not None:
        live = status_o == 0
        stats.setdefault("n", 0)
        stats["n"] += int(live.sum())
        stats["within_1e-12"] = stats.get("within_1e-12", 0) + int((rel_u[live] <= REL).sum())
        stats["max_rel_u"] = max(stats.get("max_rel_u", 0.0), float(rel_u[live].max(initial=0.0)))
    xg, yg = g["r"] * np.cos(g["theta"]), g["r"] * np.sin(g["theta"])
    xo, yo = o["r"] * np.cos(o["theta"]), o["r"] * np.sin(o["theta"])
    dxy = np.hypot(xg - xo, yg - yo)
    bad = np.flatnonzero(dxy > REL * np.maximum(o["r"], fields.dr * 1e-3))
    assert bad.size == 0, f"{label} transverse position differs at {bad[:5]}: {dxy[bad[:5]]} at r {o['r'][bad[:5]]}"
    for name, a, b, sc in (("z", g["z"], o["z"], fields.dz), ("r", g["r"], o["r"], fields.dr)):
        ok = close(a, b, sc * 1e-3)
        idx = np.flatnonzero(~ok)
        assert idx.size == 0, (f"{label} {name} differs at {idx[:5]}: oracle {b[idx[:5]]} gpu {a[idx[:5]]} "
                               f"status {status_o[idx[:5]]}")
    ids = tile_dev.id.cpu().numpy()
    invalidated = np.isin(status_o, (1, 2, 3, 6))
    assert np.all(ids[invalidated] == np.iinfo(np.int64).min), f"{label} invalidated ids not marked"
    assert np.all(ids[~invalidated] == tile_host_prev.id[~invalidated]), f"{label} live ids changed"
    ion_g = out.ion_pairs.cpu().numpy()
    # ion pairs are proportional to the chord's density path: they carry the
    # round-trip conditioning at full weight
    ion_o = oracle_out["ion_pairs"]
    ok = np.abs(ion_g - ion_o) <= (REL + 16.0 * EPS * kappa) * np.maximum(np.abs(ion_o), 1e-3)
    assert ok.all(), f"{label} ion pairs differ at {np.flatnonzero(~ok)[:5]}"
    sub_g = out.substeps.cpu().numpy()
    assert np.array_equal(sub_g, oracle_out["substeps"].astype(np.int32)), f"{label} substeps differ"
    # secondaries: same count, same (parent, lineage) order, same payload
    sg = out.secondaries_host()
    so = oracle_out["secondaries"]
    assert len(sg) == len(so), f"{label} secondary count {len(sg)} vs {len(so)}"
    for f in ("parent_index", "species", "process", "root_key", "lineage", "root_species", "generation"):
        assert np.array_equal(sg[f], so[f]), f"{label} secondary {f} differs"
    assert np.array_equal(sg["weight"], so["weight"]), f"{label} secondary weight differs"
    # a secondary is born on its parent's chord: its payload carries the
    # parent's round-trip conditioning
    kp = kappa[so["parent_index"]] if len(so) else np.zeros(0)
    pos_scale = np.maximum(np.abs(so["z"]), np.hypot(so["x"], so["y"])) if len(so) else np.zeros(0)
    for f in ("x", "y", "z"):
        assert np.all(np.abs(sg[f] - so[f]) <= REL * np.maximum(pos_scale, 1e-3)), f"{label} secondary {f} differs"
    for f, sc in (("dir_x", 1.0), ("dir_y", 1.0), ("dir_z", 1.0), ("energy_eV", 1.0), ("birth_fraction", 1.0)):
        tol = (REL + 16.0 * EPS * kp) * np.maximum(np.maximum(np.abs(sg[f]), np.abs(so[f])), sc)
        bad = np.flatnonzero(np.abs(sg[f] - so[f]) > tol)
        assert bad.size == 0, f"{label} secondary {f} differs at {bad[:5]}: {sg[f][bad[:5]]} vs {so[f][bad[:5]]}"
    # The net-chord grid: FP64 atomics sum in another order, and a face flux
    # q/dt * (x1 - x0)/dx carries the endpoints' absolute precision ~ulp(coordinate)
    # divided by the cell size.
    wmax = float(np.max(np.abs(tile_host_prev.w), initial=0.0))
    coord = fields.r_hi / fields.dr + max(abs(fields.z_lo), abs(fields.z_hi)) / fields.dz
    for k in ("jr", "jz", "rho"):
        g = getattr(out, k).cpu().numpy()
        o = oracle_out[k]
        scale = max(np.abs(o).max(), 1e-300)
        tol = 1e-11 * scale + (64.0 * EPS * coord * wmax / dt if k != "rho" else 64.0 * EPS * coord * wmax)
        assert np.abs(g - o).max() <= tol, f"{label} {k} differs: {np.abs(g - o).max()} of {scale} (tol {tol})"
    t_g = out.tallies.cpu().numpy()
    t_o = oracle_out["tallies"]
    assert np.allclose(t_g[:17], t_o[:17], rtol=1e-10, atol=1e-300), f"{label} tallies {t_g[:17]} vs {t_o[:17]}"
    compare_ledger(t_g, t_o, label, kappa=float(np.max(kappa, initial=1.0)))


def compare_ledger(t_g, t_o, label, rtol=1e-10, kappa=1.0):
    """The event ledger (every RreaEventLedger field): counts exactly; sums
    and running maxima to rtol (the device adds in another order) plus the
    worst particle's round-trip conditioning kappa, as they are sums and
    maxima of per-particle values (path lengths, taus, losses) that agree to
    the state tolerance, 1e-12 + 16 eps kappa."""
    from oracle import binding as ob
    bad = []
    tol = rtol + 16.0 * EPS * kappa
    for i, n in enumerate(ob.LEDGER_NAMES):
        g, o = float(t_g[ob.LEDGER_OFFSET + i]), float(t_o[ob.LEDGER_OFFSET + i])
        ok = g == o if n.endswith("_count") else abs(g - o) <= tol * max(abs(g), abs(o)) + 1e-300
        if not ok:
            bad.append((n, g, o))
    assert not bad, f"{label} event ledger differs: {bad}"


def host_tile_from_dev(tile_dev):
    from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile

    return ParticleTile(*(getattr(tile_dev, k).cpu().numpy().copy() for k in
                          ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))


def run_and_compare(otables, tables, coupling, fields, tile_host, cfg, steps=1, label="", nthreads=8, stats=None,
                    species="electron"):
    """Advance `steps` steps on both paths from the same start, comparing every step."""
    stats = {} if stats is None else stats
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile, StepOutputs

    tile_dev = DeviceTile.from_host(tile_host)
    out = StepOutputs(len(tile_host), fields, secondary_capacity=max(4096, len(tile_host)))
    host = tile_host
    for s in range(steps):
        cfg.step = s
        step = ob.positron_step if species == "positron" else ob.electron_step
        o = step(otables, fields, cfg, host.as_dict(), nthreads=nthreads)
        assert o["code"] == 0, f"oracle error {o['code']}"
        gpu_step(tables, coupling, tile_dev, cfg, out, species)
        compare_step(fields, host, o, tile_dev, out, label=f"{label} step {s}", dt=cfg.dt_s, stats=stats)
        # continue both from the oracle's state, keeping survivors only
        keep = o["status"] == 0
        host = host.take(np.flatnonzero(keep))
        for k in ("r", "z", "theta", "ux", "uy", "uz"):
            getattr(host, k)[:] = o[k][keep]
        if len(host) == 0:
            break
        tile_dev = DeviceTile.from_host(host)
        out = StepOutputs(len(host), fields, secondary_capacity=max(4096, len(host)))
    print(f"{label}: {stats.get('within_1e-12', 0)} of {stats.get('n', 0)} live particle-steps within 1e-12 of |u|; "
          f"max |du|/|u| = {stats.get('max_rel_u', 0.0):.3e}", flush=True)
    return host


def compare_photon_step(fields, p_in, oracle_out, tile_dev, out, label="", dt=None):
    """Assert the HIP photon step's outputs equal the oracle's."""
    status_o = oracle_out["status"]
    status_g = out.status.cpu().numpy().astype(np.int32)
    bad = np.flatnonzero(status_o != status_g)
    assert bad.size == 0, f"{label} status differs at {bad[:10]}: oracle {status_o[bad[:10]]} gpu {status_g[bad[:10]]}"
    assert np.array_equal(out.substeps.cpu().numpy(), oracle_out["substeps"].astype(np.int32)), f"{label} events"
    c = tile_dev.cur
    g = {k: getattr(c, k).cpu().numpy() for k in ("r", "z", "theta", "ux", "uy", "uz")}
    o = oracle_out
    # the chord is the difference of two Cartesian points rebuilt from (r, theta):
    # lengths along it carry eps * r / l relative, l = c dt
    ell = C_LIGHT * dt
    kappa = 1.0 + np.maximum(p_in["r"], np.abs(p_in["z"])) / ell
    xg, yg = g["r"] * np.cos(g["theta"]), g["r"] * np.sin(g["theta"])
    xo, yo = o["r"] * np.cos(o["theta"]), o["r"] * np.sin(o["theta"])
    scale = np.maximum(np.maximum(o["r"], np.abs(o["z"])), fields.dr * 1e-3)
    d = np.sqrt((xg - xo) ** 2 + (yg - yo) ** 2 + (g["z"] - o["z"]) ** 2)
    bad = np.flatnonzero(d > REL * scale)
    assert bad.size == 0, f"{label} position differs at {bad[:5]}: {d[bad[:5]]} of {scale[bad[:5]]}"
    du = np.sqrt((g["ux"] - o["ux"]) ** 2 + (g["uy"] - o["uy"]) ** 2 + (g["uz"] - o["uz"]) ** 2)
    umag = np.sqrt(o["ux"] ** 2 + o["uy"] ** 2 + o["uz"] ** 2)
    bad = np.flatnonzero(du > REL * np.maximum(umag, 1.0))
    assert bad.size == 0, f"{label} u differs at {bad[:5]}: {du[bad[:5]] / umag[bad[:5]]}"
    ids = c.id.cpu().numpy()
    invalidated = np.isin(status_o, (1, 2, 3, 7))
    assert np.all(ids[invalidated] == np.iinfo(np.int64).min), f"{label} invalidated ids not marked"
    assert np.all(ids[~invalidated] == p_in["id"][~invalidated]), f"{label} live ids changed"
    ion_g = out.ion_pairs.cpu().numpy()
    ion_o = o["ion_pairs"]
    assert np.all(np.abs(ion_g - ion_o) <= REL * np.maximum(np.abs(ion_o), 1e-3)), f"{label} ion pairs differ"
    sg = out.secondaries_host()
    so = o["secondaries"]
    assert len(sg) == len(so), f"{label} secondary count {len(sg)} vs {len(so)}"
    for f in ("parent_index", "species", "process", "root_key", "lineage", "root_species", "generation"):
        assert np.array_equal(sg[f], so[f]), f"{label} secondary {f} differs"
    assert np.array_equal(sg["weight"], so["weight"]), f"{label} secondary weight differs"
    kp = kappa[so["parent_index"]] if len(so) else np.zeros(0)
    pos_scale = np.maximum(np.abs(so["z"]), np.hypot(so["x"], so["y"])) if len(so) else np.zeros(0)
    for f in ("x", "y", "z"):
        assert np.all(np.abs(sg[f] - so[f]) <= REL * np.maximum(pos_scale, 1e-3)), f"{label} secondary {f} differs"
    for f in ("dir_x", "dir_y", "dir_z", "energy_eV", "birth_fraction"):
        tol = (REL + 16.0 * EPS * kp) * np.maximum(np.maximum(np.abs(sg[f]), np.abs(so[f])), 1.0)
        bad = np.flatnonzero(np.abs(sg[f] - so[f]) > tol)
        assert bad.size == 0, f"{label} secondary {f} differs at {bad[:5]}: {sg[f][bad[:5]]} vs {so[f][bad[:5]]}"
    t_g = out.tallies.cpu().numpy()
    t_o = o["tallies"]
    assert np.allclose(t_g[:17], t_o[:17], rtol=1e-10, atol=1e-300), f"{label} tallies {t_g[:17]} vs {t_o[:17]}"
    compare_ledger(t_g, t_o, label, kappa=float(np.max(kappa, initial=1.0)))


def run_photons_and_compare(otables, tables, coupling, fields, tile_host, cfg, steps=1, label="", nthreads=8):
    """Advance photons `steps` steps on both paths from the same start (the
    host push of workloads.push_photons between steps), comparing every step."""
    import torch

    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.transport import PhotonTile, RreaParticleInteraction, StepOutputs

    p = tile_host.as_dict()
    p.update(tile_host.extra)
    p = {k: np.array(v) for k, v in p.items()}
    for s in range(steps):
        cfg.step = s
        o = ob.photon_step(otables, fields, cfg, p, nthreads=nthreads)
        assert o["code"] == 0, f"oracle error {o['code']}"
        host = tile_host.take(np.arange(len(p["r"])))
        for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu"):
            getattr(host, k)[:] = p[k]
        host.extra = {"r_cur": p["r_cur"], "z_cur": p["z_cur"], "theta_cur": p["theta_cur"]}
        dev = PhotonTile.from_host(host)
        out = StepOutputs(len(host), fields, secondary_capacity=max(4096, 4 * len(host)))
        RreaParticleInteraction.apply_photons(tables, coupling, dev, cfg, out)
        torch.cuda.synchronize()
        compare_photon_step(fields, p, o, dev, out, label=f"{label} step {s}", dt=cfg.dt_s)
        keep = np.flatnonzero(o["status"] == 0)
        if keep.size == 0:
            break
        nxt = {k: np.array(p[k][keep]) for k in ("w", "id", "cpu")}
        for k in ("ux", "uy", "uz"):
            nxt[k] = o[k][keep]
        nxt["r"], nxt["z"], nxt["theta"] = o["r"][keep], o["z"][keep], o["theta"][keep]
        nxt["r_cur"], nxt["theta_cur"], nxt["z_cur"] = wl.push_photons(nxt["r"], nxt["theta"], nxt["z"], nxt["ux"],
                                                                       nxt["uy"], nxt["uz"], cfg.dt_s)
        p = nxt
        tile_host = tile_host.take(keep)
    print(f"{label}: {steps} photon steps compared", flush=True)


def drain_conditioning(work, dt, substeps):
    """Round-trip conditioning of each FIFO entry's residual chords,
    kappa = (1 + max(r, |z|) / l) n with l the microchord length (v (1 - f)
    dt over the entry's n substeps, c (1 - f) dt for a photon, n = 1), carried
    down to the entry's children.  The factor n: each substep's multiple
    scattering feeds the previous substeps' last-place differences into the
    next chord's direction and field work, so they add up over the substeps
    rather than staying at one chord's worth."""
    k = work["energy_eV"]
    gamma = 1.0 + k / ME_C2
    speed = np.where(work["species"] == 2, C_LIGHT,
                     C_LIGHT * np.sqrt(np.maximum(0.0, 1.0 - 1.0 / (gamma * gamma))))
    nsub = np.where(work["species"] == 2, 1, np.maximum(substeps, 1))
    ell = np.maximum(speed * np.maximum(1.0 - work["birth_fraction"], 1e-6) * dt / nsub, 1e-300)
    kappa = (1.0 + np.maximum(np.hypot(work["x"], work["y"]), np.abs(work["z"])) / ell) * nsub
    gen = work["generation"]
    n0 = int(np.sum(gen == 0))
    # a child's payload inherits its parent's (and so its ancestors') conditioning
    for i in range(n0, len(work)):
        kappa[i] = max(kappa[i], kappa[work["parent_index"][i]])
    return kappa


def compare_drain(fields, cfg, oracle_out, out, label="", stats=None):
    """Assert the HIP newborn drain equals the oracle's FIFO: the same FIFO
    (order, lineage, generation, parent, species, process exact), the same
    statuses, payloads and survivors to 1e-12 plus the residual chords'
    conditioning, tallies, and the J grid."""
    wg = out.work_host()
    wo = oracle_out["work"]
    assert len(wg) == len(wo), f"{label} FIFO length {len(wg)} vs {len(wo)}"
    n = len(wo)
    for f in ("parent_index", "species", "process", "root_key", "lineage", "root_species", "generation"):
        bad = np.flatnonzero(wg[f] != wo[f])
        assert bad.size == 0, f"{label} FIFO {f} differs at {bad[:5]}: {wg[f][bad[:5]]} vs {wo[f][bad[:5]]}"
    assert np.array_equal(wg["weight"], wo["weight"]), f"{label} FIFO weight differs"
    st_g = out.status[:n].cpu().numpy().astype(np.int32)
    bad = np.flatnonzero(st_g != oracle_out["status"])
    assert bad.size == 0, f"{label} status differs at {bad[:10]}: {st_g[bad[:10]]} vs {oracle_out['status'][bad[:10]]}"
    sub_g = out.substeps[:n].cpu().numpy().astype(np.int64)
    bad = np.flatnonzero(sub_g != oracle_out["substeps"])
    assert bad.size == 0, f"{label} substeps differ at {bad[:10]}: {sub_g[bad[:10]]} vs {oracle_out['substeps'][bad[:10]]}"
    kappa = drain_conditioning(wo, cfg.dt_s, oracle_out["substeps"])
    kp = np.where(wo["generation"] > 0, kappa[np.minimum(wo["parent_index"], n - 1)], 1.0)
    scale = np.maximum(np.maximum(np.hypot(wo["x"], wo["y"]), np.abs(wo["z"])), 1e-3)
    for f in ("x", "y", "z"):
        d = np.abs(wg[f] - wo[f])
        bad = np.flatnonzero(d > (REL + 64.0 * EPS) * scale)
        assert bad.size == 0, f"{label} FIFO {f} differs at {bad[:5]}: {d[bad[:5]]}"
    for f in ("dir_x", "dir_y", "dir_z", "energy_eV", "birth_fraction"):
        tol = (REL + 16.0 * EPS * kp) * np.maximum(np.maximum(np.abs(wg[f]), np.abs(wo[f])), 1.0)
        bad = np.flatnonzero(np.abs(wg[f] - wo[f]) > tol)
        assert bad.size == 0, f"{label} FIFO {f} differs at {bad[:5]}: {wg[f][bad[:5]]} vs {wo[f][bad[:5]]}"
    ion_g = out.ion_pairs[:n].cpu().numpy()
    ion_o = oracle_out["ion_pairs"]
    bad = np.flatnonzero(np.abs(ion_g - ion_o) > (REL + 16.0 * EPS * kappa) * np.maximum(np.abs(ion_o), 1e-3))
    assert bad.size == 0, f"{label} ion pairs differ at {bad[:5]}: {ion_g[bad[:5]]} vs {ion_o[bad[:5]]}"
    sg = out.survivors_host()
    so = oracle_out["survivors"]
    assert len(sg) == len(so), f"{label} survivor count {len(sg)} vs {len(so)}"
    alive = np.flatnonzero(oracle_out["status"] == 0)
    ks = kappa[alive]
    for f in ("species", "root_key", "lineage", "generation"):
        assert np.array_equal(sg[f], so[f]), f"{label} survivor {f} differs"
    sscale = np.maximum(np.maximum(np.hypot(so["x"], so["y"]), np.abs(so["z"])), 1e-3)
    for f in ("x", "y", "z"):
        d = np.abs(sg[f] - so[f])
        bad = np.flatnonzero(d > (REL + 64.0 * EPS) * sscale)
        assert bad.size == 0, f"{label} survivor {f} differs at {bad[:5]}: {d[bad[:5]]}"
    rel_e = np.abs(sg["energy_eV"] - so["energy_eV"]) / np.maximum(so["energy_eV"], 1.0)
    bad = np.flatnonzero(rel_e > REL + 16.0 * EPS * ks)
    if bad.size:
        w_bad = wo[alive[bad[:5]]]
        detail = [(int(a), int(w["species"]), int(w["generation"]), float(w["birth_fraction"]), float(w["energy_eV"]),
                   float(so["energy_eV"][b]), int(oracle_out["substeps"][a]), float(ks[b]),
                   float(np.hypot(w["x"], w["y"])), float(w["z"]))
                  for a, b, w in zip(alive[bad[:5]], bad[:5], w_bad)]
        raise AssertionError(f"{label} survivor energy differs at {bad[:5]}: {rel_e[bad[:5]]}; "
                             f"(entry, species, generation, birth, E0, E1, substeps, kappa, r, z) = {detail}")
    dd = np.sqrt(sum((sg[f] - so[f]) ** 2 for f in ("dir_x", "dir_y", "dir_z")))
    bad = np.flatnonzero(dd > REL + 16.0 * EPS * ks)
    assert bad.size == 0, f"{label} survivor direction differs at {bad[:5]}: {dd[bad[:5]]}"
    if stats is not None:
        stats["entries"] = stats.get("entries", 0) + n
        stats["survivors"] = stats.get("survivors", 0) + len(so)
        stats["within_1e-12"] = stats.get("within_1e-12", 0) + int((rel_e <= REL).sum())
        stats["max_rel_e"] = max(stats.get("max_rel_e", 0.0), float(rel_e.max(initial=0.0)))
    t_g = out.grid.tallies.cpu().numpy()
    t_o = oracle_out["tallies"]
    assert np.array_equal(t_g[17:21], t_o[17:21]), f"{label} newborn counts {t_g[17:21]} vs {t_o[17:21]}"
    assert np.allclose(t_g[:17], t_o[:17], rtol=1e-9, atol=1e-300), f"{label} tallies {t_g[:17]} vs {t_o[:17]}"
    compare_ledger(t_g, t_o, label, rtol=1e-9, kappa=float(np.max(kappa, initial=1.0)))
    # a newborn's chord ends where its transport ended, which the comparison
    # above allows to differ by (1e-12 + 64 eps) of the coordinates; a face
    # flux q/dt (x1 - x0)/dx inherits that in index units
    wmax = float(np.max(wo["weight"], initial=0.0))
    coord = fields.r_hi / fields.dr + max(abs(fields.z_lo), abs(fields.z_hi)) / fields.dz
    for k in ("jr", "jz", "rho"):
        g = getattr(out.grid, k).cpu().numpy()
        o = oracle_out[k]
        sc = max(np.abs(o).max(), 1e-300)
        edge = (REL + 64.0 * EPS) * coord * wmax
        tol = 1e-11 * sc + (edge / cfg.dt_s if k != "rho" else edge)
        assert np.abs(g - o).max() <= tol, f"{label} {k} differs: {np.abs(g - o).max()} of {sc} (tol {tol})"
