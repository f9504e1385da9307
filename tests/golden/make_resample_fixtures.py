"""Writes tests/golden/resample.npz: inputs and outputs of the overlay's
adaptive_resample_v1 thinning plan on one rank, computed by the reference's
own RreaCicConservingResample.H and RreaRng (compiled from the checkout into
oracle/_ref by `make -C oracle ref`; the single-rank flow of
RreaCouplingPopulation.cpp:321-520 is ref_population_thin in
oracle/ref_driver.cpp).

Each case holds particles (r, z, theta, ux, uy, uz, w, stable ids,
candidate flags), the mesh, the target and the draw keys; the outputs are
every record's new weight and the plan's counts (capacity, geometric floor,
nominal request, removed).  The support cell, fractions and phase bin of a
record are formed here as the coupling's record lambda forms them
(RreaCouplingPopulation.cpp:233-277); `records()` is that restatement.

    python tests/golden/make_resample_fixtures.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

C_LIGHT = 299792458.0


def records(r, z, theta, ux, uy, uz, z_lo, dr, dz):
    """Support cell (lower-left of the order-1 cell-centred CIC stencil),
    fractions and phase bin of each particle (non-periodic, domain small end 0)."""
    rc = (r - 0.0) / dr + 0.0 - 0.5
    zc = (z - z_lo) / dz + 0.0 - 0.5
    i0, j0 = np.floor(rc), np.floor(zc)
    c, s = np.cos(theta), np.sin(theta)
    ur = ux * c + uy * s
    ut = -ux * s + uy * c
    mag = np.sqrt(ux * ux + uy * uy + uz * uz)
    _, e = np.frexp(mag)
    ex = np.where((mag > 0) & np.isfinite(mag), e - 1, -1024)
    ex = np.clip(ex, -1024, 1023)
    phase = (ex + 1024) * 8 + (ur >= 0) + 2 * (ut >= 0) + 4 * (uz >= 0)
    return (i0.astype(np.int32), j0.astype(np.int32), phase.astype(np.int32), rc - i0, zc - j0)


def particles(rng, n, cells_r, cells_z, dr, dz, z_lo):
    r = dr * cells_r * rng.random(n)
    z = z_lo + dz * cells_z * rng.random(n)
    theta = rng.uniform(-np.pi, np.pi, n)
    k = rng.uniform(1e5, 2e7, n)
    g = 1.0 + k / 510998.95
    u = C_LIGHT * np.sqrt(g * g - 1.0)
    mu = rng.uniform(-1, 1, n)
    phi = rng.uniform(0, 2 * np.pi, n)
    st = np.sqrt(1 - mu * mu)
    return r, z, theta, u * st * np.cos(phi), u * st * np.sin(phi), u * mu


def cases():
    rng = np.random.default_rng(20240623)
    out = []
    dr, dz, z_lo = 1.0, 1.25, -2.0
    # dense electrons in a 6 x 6 patch, thinned to a quarter
    n = 4000
    p = particles(rng, n, 6, 6, dr, dz, z_lo)
    out.append(dict(name="dense_electrons", species=0, target=1000, p=p, w=rng.uniform(0.5, 2.0, n),
                    low=np.arange(1, n + 1, dtype=np.uint64), high=None, cand=None))
    # positrons over a 3 x 3 patch, log-normal weights
    n = 600
    p = particles(rng, n, 3, 3, dr, dz, z_lo)
    out.append(dict(name="positrons", species=1, target=400, p=p, w=np.exp(rng.normal(0, 1, n)),
                    low=(np.arange(1, n + 1, dtype=np.uint64) << np.uint64(3)) ^ np.uint64(5), high=None,
                    cand=None))
    # split clones: every particle twice, same position, momentum and weight
    m = 300
    q = particles(rng, m, 4, 4, dr, dz, z_lo)
    p = tuple(np.concatenate([a, a]) for a in q)
    w = rng.uniform(0.5, 2.0, m)
    out.append(dict(name="split_clones", species=0, target=200, p=p, w=np.concatenate([w, w]),
                    low=np.arange(1, 2 * m + 1, dtype=np.uint64), high=None, cand=None))
    # below the control band: nothing thinned
    n = 100
    p = particles(rng, n, 2, 2, dr, dz, z_lo)
    out.append(dict(name="below_band", species=0, target=90, p=p, w=np.ones(n),
                    low=np.arange(1, n + 1, dtype=np.uint64), high=None, cand=None))
    # live particles and staged candidates (a secondary transaction)
    n_live, n_cand = 500, 300
    p = particles(rng, n_live + n_cand, 3, 4, dr, dz, z_lo)
    cand = np.concatenate([np.zeros(n_live, np.uint8), np.ones(n_cand, np.uint8)])
    high = np.concatenate([np.zeros(n_live, np.uint64), rng.integers(1, 50, n_cand).astype(np.uint64)])
    low = np.concatenate([np.arange(1, n_live + 1, dtype=np.uint64),
                          (np.arange(n_cand, dtype=np.uint64) << np.uint64(2)) | np.uint64(0)])
    out.append(dict(name="live_and_candidates", species=0, target=200, p=p, w=rng.uniform(0.2, 3.0, n_live + n_cand),
                    low=low, high=high, cand=cand))
    # many supports (more than one 64-lane workgroup of support plans)
    n = 3000
    p = particles(rng, n, 20, 10, dr, dz, z_lo)
    out.append(dict(name="many_supports", species=0, target=1500, p=p, w=rng.uniform(0.5, 2.0, n),
                    low=np.arange(1, n + 1, dtype=np.uint64), high=None, cand=None))
    return out, (z_lo, dr, dz)


def main():
    from oracle import binding as ob
    R = ob.ref()
    if R is None:
        raise SystemExit("the reference checkout is needed to write the fixtures")
    P = lambda t: C.POINTER(t)  # noqa: E731
    R.ref_population_thin.restype = C.c_int
    R.ref_population_thin.argtypes = [C.c_int64, P(C.c_int32), P(C.c_int32), P(C.c_int32), P(C.c_double),
                                      P(C.c_double), P(C.c_double), P(C.c_uint8), P(C.c_uint64), P(C.c_uint64),
                                      C.c_int64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, P(C.c_double),
                                      P(C.c_int64)]
    seed, salt, step = 20240623, 0x5eed, 12
    all_cases, (z_lo, dr, dz) = cases()
    arrays = {"z_lo": z_lo, "dr": dr, "dz": dz, "seed": seed, "salt": salt, "step": step,
              "names": np.array([c["name"] for c in all_cases])}
    for k, c in enumerate(all_cases):
        r, z, theta, ux, uy, uz = c["p"]
        sr, sz, phase, fr, fz = records(r, z, theta, ux, uy, uz, z_lo, dr, dz)
        n = len(r)
        w = np.ascontiguousarray(c["w"], dtype=np.float64)
        out_w = np.zeros(n)
        counts = np.zeros(4, dtype=np.int64)
        ptr = lambda a, t: None if a is None else np.ascontiguousarray(a).ctypes.data_as(C.POINTER(t))  # noqa: E731
        cand = None if c["cand"] is None else np.ascontiguousarray(c["cand"], dtype=np.uint8)
        high = None if c["high"] is None else np.ascontiguousarray(c["high"], dtype=np.uint64)
        low = np.ascontiguousarray(c["low"], dtype=np.uint64)
        st = R.ref_population_thin(n, ptr(sr, C.c_int32), ptr(sz, C.c_int32), ptr(phase, C.c_int32),
                                   ptr(fr, C.c_double), ptr(fz, C.c_double), ptr(w, C.c_double),
                                   ptr(cand, C.c_uint8), ptr(high, C.c_uint64), ptr(low, C.c_uint64),
                                   c["target"], seed, salt, step, c["species"], ptr(out_w, C.c_double),
                                   ptr(counts, C.c_int64))
        assert st == 0, (c["name"], st)
        for key, val in (("r", r), ("z", z), ("theta", theta), ("ux", ux), ("uy", uy), ("uz", uz), ("w", w),
                         ("low", low), ("out_w", out_w), ("counts", counts)):
            arrays[f"{k}_{key}"] = val
        arrays[f"{k}_species"] = c["species"]
        arrays[f"{k}_target"] = c["target"]
        if high is not None:
            arrays[f"{k}_high"] = high
        if cand is not None:
            arrays[f"{k}_cand"] = cand
        print(f"{c['name']}: n={n} capacity={counts[0]} floor={counts[1]} requested={counts[2]} "
              f"removed={counts[3]}")
    np.savez_compressed(os.path.join(HERE, "resample.npz"), **arrays)


if __name__ == "__main__":
    main()
