"""Population control on the device: the charge-exact thinning plan of the
overlay's adaptive_resample_v1 controller (rrea_population_thin; reference
RreaWarpXCoupling::PreflightAndApplySecondaryPopulationControl,
warpx_overlay/Source/Rrea/RreaCouplingPopulation.cpp:175-520 on one rank) and
its split refill (RreaAdaptiveResampleSplitOutcome, RreaRng.H:163-174;
AdaptiveResampleOneSpecies, RreaCouplingPopulation.cpp:686-787).

``thin`` plans over explicit records (live particles and staged candidates
with their stable ids); ``thin_tile`` over a live species tile, the stable
id of a particle being its RNG key |id| ^ (cpu << 40).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from . import _lib
from .config import RzFields

ACTIONS = ("none", "thin", "split")


@dataclass
class ThinResult:
    capacity: int        # exact removal capacity over all supports
    geometric_floor: int # records no null direction can remove
    requested: int       # the nominal request: min(count - target, capacity) above 1.25 x target
    removed: int


def action(count: int, target: int, split_min_weight: float) -> str:
    """RreaAdaptiveResampleAction (RreaRng.H:139-150): thin strictly above
    1.25 x target, split strictly below 0.5 x target (with a positive split
    floor), else nothing."""
    if count > target + target // 4:
        return "thin"
    if 0 < count < target // 2 and split_min_weight > 0.0:
        return "split"
    return "none"


def _p(t):
    return None if t is None else t.data_ptr()


def thin(r, z, theta, ux, uy, uz, w, stable_id, fields: RzFields, target: int, rng_seed: int, step: int,
         species: int = 0, rng_salt: int = 0, stable_id_high=None, candidate=None, id=None,
         error: torch.Tensor | None = None, stream: int | None = None) -> ThinResult:
    """Thin the records in place (w; id invalidated where given).  Tensors are
    device float64 (positions, u, w), uint64-as-int64 stable ids, uint8
    candidate flags.  Raises on a device error (weights then unchanged)."""
    n = w.numel()
    rec = _lib.PopulationRecords(n, *[_p(t) for t in (r, z, theta, ux, uy, uz, w, stable_id, stable_id_high,
                                                        candidate, id)])
    cfg = _lib.PopulationConfig(fields.z_lo, fields.dr, fields.dz, int(target), int(rng_seed) & (2**64 - 1),
                                int(rng_salt) & (2**64 - 1), int(step), int(species), 0)
    counts = (C.c_int64 * 4)()
    own_error = error is None
    if own_error:
        error = torch.zeros(2, dtype=torch.int32, device=w.device)
    if stream is None:
        stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib().rrea_population_thin(C.byref(rec), C.byref(cfg), counts, error.data_ptr(), stream),
               "rrea_population_thin")
    if own_error:
        code, index = (int(v) for v in error.cpu().tolist())
        if code:
            raise _lib.RreaHipError(f"population control: {_lib.ERRORS.get(code, code)} at record {index}")
    return ThinResult(*[int(c) for c in counts])


def stable_ids(tile) -> torch.Tensor:
    """The stable id of each particle of a tile: |id| ^ (cpu << 40)."""
    return torch.abs(tile.id) ^ (tile.cpu.to(torch.int64) << 40)


def thin_tile(tile, fields: RzFields, target: int, rng_seed: int, step: int, species: int = 0,
              rng_salt: int = 0) -> ThinResult:
    """Thin a live charged tile (every valid particle a record); particles
    whose weight reaches zero are invalidated.  Invalid particles are left
    alone."""
    valid = tile.id != _lib.INVALID_ID
    idx = torch.nonzero(valid).flatten()
    sub = {k: getattr(tile, k)[idx].contiguous() for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")}
    sid = (torch.abs(sub["id"]) ^ (sub["cpu"].to(torch.int64) << 40)).contiguous()
    res = thin(sub["r"], sub["z"], sub["theta"], sub["ux"], sub["uy"], sub["uz"], sub["w"], sid, fields, target,
               rng_seed, step, species, rng_salt, id=sub["id"])
    tile.w[idx] = sub["w"]
    tile.id[idx] = sub["id"]
    return res


def split_outcome(weight: torch.Tensor, split_min_weight: float):
    """RreaAdaptiveResampleSplitOutcome per particle: (split mask, child
    weight); a parent at or above twice the floor halves and clones."""
    if not split_min_weight > 0.0:
        return torch.zeros_like(weight, dtype=torch.bool), weight
    split = weight >= 2.0 * split_min_weight
    return split, torch.where(split, 0.5 * weight, weight)
