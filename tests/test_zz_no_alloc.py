"""rrea_transport_step_noalloc -- the species step whose workspace is one
fixed device array -- against the allocating entry points
(rrea_transport_step_electrons / _positrons / _photons, ordering with the
count on the device) on the same inputs: particle states, statuses, ion
pairs and substeps bit for bit, the ordered secondaries record for record,
the J grid and material sources to FP64 summation order, tallies to 1e-12.
Two calls in a row, the second on a smaller tile, so the second call runs
over what the first left in the shared array."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import _lib  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu

ENTRY = {0: "rrea_transport_step_electrons", 1: "rrea_transport_step_positrons", 2: "rrea_transport_step_photons"}
NAME = {0: "electrons", 1: "positrons", 2: "photons"}


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def host_tile(species, n, fields, dt, seed):
    if species == 2:
        return wl.c2_photons(n, fields, dt_s=dt, seed=seed)
    return wl.c2_electrons(n, fields, seed=seed, charge_sign=1.0 if species == 1 else -1.0)


def device_tile(species, host):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile, PhotonTile
    return PhotonTile.from_host(host) if species == 2 else DeviceTile.from_host(host)


def arrays(species, t):
    cur = t.cur if species == 2 else t
    out = {k: getattr(cur, k) for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id")}
    if species == 2:
        out.update(prev_r=t.prev_r, prev_z=t.prev_z, prev_theta=t.prev_theta)
    return out


@pytest.mark.parametrize("species", [0, 1, 2])
def test_noalloc_step_matches_allocating_step(tables, species):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling, StepOutputs, _step_config
    L = _lib.lib()
    fields = wl.c2_fields()
    coupling = Coupling(fields, tables)
    stream = torch.cuda.current_stream().cuda_stream
    dt = 1e-9
    for call, n in enumerate((1 << 15, 1 << 12)):
        print(f"case noalloc {NAME[species]}: call {call}, {n} particles, dt={dt}", flush=True)
        host = host_tile(species, n, fields, dt, seed=60 + 7 * call + species)
        tiles = [device_tile(species, host) for _ in range(2)]
        outs = [StepOutputs(n, fields, secondary_capacity=1 << 15, sources=True) for _ in range(2)]
        cfg = _step_config(StepConfig(rng_seed=91 + call, step=call, dt_s=dt))
        for k, (t, o) in enumerate(zip(tiles, outs)):
            ts = t.struct()
            so = o.struct(2)
            if k == 0:
                st = getattr(L, ENTRY[species])(tables.h, coupling.h, C.byref(ts), C.byref(cfg), C.byref(so), stream)
            else:
                st = L.rrea_transport_step_noalloc(species, tables.h, coupling.h, C.byref(ts), C.byref(cfg),
                                                   C.byref(so), stream)
            assert st == 0, (k, st)
        torch.cuda.synchronize()
        ref, got = outs
        ref.check_error()
        got.check_error()
        a, b = arrays(species, tiles[0]), arrays(species, tiles[1])
        for key in a:
            assert torch.equal(a[key], b[key]), (call, key)
        for key in ("status", "ion_pairs", "substeps"):
            assert torch.equal(getattr(ref, key), getattr(got, key)), (call, key)
        sa, sb = ref.secondaries_host(), got.secondaries_host()
        assert len(sa) == len(sb) and (sa == sb).all(), call
        grids = ("sources",) if species == 2 else ("jr", "jz", "rho", "sources")
        for key in grids:
            x, y = getattr(ref, key), getattr(got, key)
            scale = max(float(x.abs().max()), 1e-300)
            assert torch.allclose(x, y, rtol=0.0, atol=1e-12 * scale), (call, key)
        assert torch.allclose(ref.tallies, got.tallies, rtol=1e-12, atol=0.0), call
        print(f"noalloc {NAME[species]} call {call}: {n} particles, {len(sa)} secondaries, "
              f"{int((ref.status != 0).sum())} not surviving -- identical", flush=True)


def test_noalloc_limits_are_refused(tables):
    """Beyond the fixed workspace: more than 2^20 particles, or an ordered
    secondary capacity above 2^15 -- refused before any launch."""
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling, StepOutputs, _step_config
    L = _lib.lib()
    fields = wl.c2_fields()
    coupling = Coupling(fields, tables)
    cfg = _step_config(StepConfig(rng_seed=1, dt_s=1e-12))
    big = _lib.Tile()
    big.n = (1 << 20) + 1
    out = StepOutputs(16, fields, secondary_capacity=1 << 15)
    so = out.struct(2)
    assert L.rrea_transport_step_noalloc(0, tables.h, coupling.h, C.byref(big), C.byref(cfg), C.byref(so), None) == 16
    host = host_tile(0, 64, fields, 1e-12, seed=5)
    t = device_tile(0, host)
    ts = t.struct()
    wide = StepOutputs(64, fields, secondary_capacity=(1 << 15) + 1)
    sw = wide.struct(2)
    assert L.rrea_transport_step_noalloc(0, tables.h, coupling.h, C.byref(ts), C.byref(cfg), C.byref(sw), None) == 16
    assert L.rrea_transport_step_noalloc(3, tables.h, coupling.h, C.byref(ts), C.byref(cfg), C.byref(sw), None) == 16
