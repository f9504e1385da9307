"""GPU parity of population control (rrea_population_thin) against the
reference's own thinning plan: tests/golden/resample.npz holds inputs and
outputs of RreaPlanGlobalCicThinning / RreaThinCicWeights compiled from the
reference checkout (tests/golden/make_resample_fixtures.py).  The plan's
counts and every removal decision must be identical; surviving weights
agree to a few ulp (the reference eliminates in x87 long double, the device
in double-double); every support's four deposited CIC charges are conserved
to round-off."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd.config import RzFields  # noqa: E402

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "resample.npz"))


def support_charges(fix_arrays, w, z_lo, dr, dz):
    from golden.make_resample_fixtures import records
    r, z, theta, ux, uy, uz = (fix_arrays[k] for k in ("r", "z", "theta", "ux", "uy", "uz"))
    sr, sz, _, fr, fz = records(r, z, theta, ux, uy, uz, z_lo, dr, dz)
    key = sr.astype(np.int64) * 1_000_003 + sz
    cells = np.stack([w * (1 - fr) * (1 - fz), w * fr * (1 - fz), w * (1 - fr) * fz, w * fr * fz], axis=1)
    out = {}
    for k in np.unique(key):
        out[int(k)] = cells[key == k].sum(axis=0)
    return out


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


@pytest.mark.parametrize("case", range(len(FIX["names"])))
def test_thinning_matches_reference_plan(gpu, case):
    from rrea_feedback_warpx_amrex_geant4_amd.population import thin
    name = str(FIX["names"][case])
    a = {k[len(f"{case}_"):]: FIX[k] for k in FIX.files if k.startswith(f"{case}_")}
    z_lo, dr, dz = float(FIX["z_lo"]), float(FIX["dr"]), float(FIX["dz"])
    n = len(a["w"])
    print(f"case population {name}: {n} records, target {int(a['target'])}", flush=True)
    fields = RzFields(64, 64, dr, dz, z_lo, np.zeros(65 * 64), np.zeros(64 * 65), np.zeros(65 * 64),
                      np.zeros(64 * 65), np.zeros(65 * 65))
    d = {k: torch.as_tensor(a[k], dtype=torch.float64, device="cuda") for k in ("r", "z", "theta", "ux", "uy", "uz")}
    w = torch.as_tensor(a["w"], dtype=torch.float64, device="cuda").clone()
    low = torch.as_tensor(a["low"].view(np.int64), device="cuda")
    high = torch.as_tensor(a["high"].view(np.int64), device="cuda") if "high" in a else None
    cand = torch.as_tensor(a["cand"], device="cuda") if "cand" in a else None
    ids = torch.arange(1, n + 1, dtype=torch.int64, device="cuda")
    res = thin(d["r"], d["z"], d["theta"], d["ux"], d["uy"], d["uz"], w, low, fields, int(a["target"]),
               int(FIX["seed"]), int(FIX["step"]), int(a["species"]), int(FIX["salt"]), stable_id_high=high,
               candidate=cand, id=ids)
    want = a["counts"]
    assert (res.capacity, res.geometric_floor, res.requested, res.removed) == tuple(int(x) for x in want), name
    got = w.cpu().numpy()
    ref = a["out_w"]
    # the same records removed, the same survivors
    assert np.array_equal(got == 0.0, ref == 0.0), name
    assert np.array_equal(ids.cpu().numpy() == np.iinfo(np.int64).min, ref == 0.0), name
    np.testing.assert_allclose(got, ref, rtol=4 * np.finfo(float).eps, atol=0, err_msg=name)
    # every support keeps its four deposited charges
    before = support_charges(a, a["w"], z_lo, dr, dz)
    after = support_charges(a, got, z_lo, dr, dz)
    for k in before:
        assert np.allclose(after[k], before[k], rtol=1e-12, atol=1e-12 * max(1.0, np.abs(before[k]).max())), (name, k)
    print(f"population {name}: capacity {res.capacity}, removed {res.removed} of {n}; decisions identical, "
          f"max weight difference {np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)):.2e} relative", flush=True)


def test_thinning_at_scale_conserves_and_meets_the_request(gpu):
    """2^20 electrons over the 256 x 1024 grid (~260,000 supports), thinned
    to half: the nominal request is met, the removed count is the zeroed
    count, and every support's CIC charges are unchanged (size-independent
    properties; the reference's own plan completes this input too)."""
    from golden.make_resample_fixtures import records
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.population import thin_tile
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile
    fields = wl.c2_fields()
    n = 1 << 20
    print(f"case population at scale: {n} electrons", flush=True)
    host = wl.c2_electrons(n, fields, seed=31)
    dev = DeviceTile.from_host(host)
    res = thin_tile(dev, fields, target=n // 2, rng_seed=20240623, step=0)
    w = dev.w.cpu().numpy()
    assert res.requested > 0 and res.removed >= res.requested
    assert int((w == 0.0).sum()) == res.removed
    sr, sz, _, fr, fz = records(host.r, host.z, host.theta, host.ux, host.uy, host.uz, fields.z_lo, fields.dr,
                                fields.dz)
    key = (sr.astype(np.int64) + 1) * (fields.nz + 2) + (sz + 1)
    for wt, name in ((host.w, "before"), (w, "after")):
        cells = np.stack([np.bincount(key, wt * (1 - fr) * (1 - fz)), np.bincount(key, wt * fr * (1 - fz)),
                          np.bincount(key, wt * (1 - fr) * fz), np.bincount(key, wt * fr * fz)])
        if name == "before":
            c0 = cells
    assert np.allclose(cells, c0, rtol=1e-12, atol=1e-12)
    print(f"population at scale: removed {res.removed} (requested {res.requested}, capacity {res.capacity})",
          flush=True)


def test_thin_tile_invalidates_and_conserves(gpu):
    """A live tile thinned in place: zero-weight particles invalidated, the
    survivors' CIC charges per support unchanged, invalid particles untouched."""
    from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
    from rrea_feedback_warpx_amrex_geant4_amd.population import thin_tile
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile
    fields, tiles, _ = wl.c3_batch(1, seed=8, scale=0.02)
    host = tiles[0]
    host.id[::11] = np.iinfo(np.int64).min
    n = len(host.r)
    print(f"case population tile: {n} electrons", flush=True)
    dev = DeviceTile.from_host(host)
    valid = host.id != np.iinfo(np.int64).min
    res = thin_tile(dev, fields[0], target=int(valid.sum()) // 3, rng_seed=7, step=3)
    w = dev.w.cpu().numpy()
    ids = dev.id.cpu().numpy()
    assert res.removed >= res.requested > 0
    assert np.array_equal(ids[~valid], host.id[~valid]) and np.array_equal(w[~valid], host.w[~valid])
    zero = valid & (w == 0.0)
    assert int(zero.sum()) == res.removed
    assert np.all(ids[zero] == np.iinfo(np.int64).min)
    arrays = {k: getattr(host, k)[valid] for k in ("r", "z", "theta", "ux", "uy", "uz")}
    before = support_charges(arrays, host.w[valid], fields[0].z_lo, fields[0].dr, fields[0].dz)
    after = support_charges(arrays, w[valid], fields[0].z_lo, fields[0].dr, fields[0].dz)
    for k in before:
        assert np.allclose(after[k], before[k], rtol=1e-12, atol=1e-12 * max(1.0, np.abs(before[k]).max()))
    print(f"population tile: removed {res.removed} of {int(valid.sum())} (capacity {res.capacity})", flush=True)
