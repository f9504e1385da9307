"""Autotuning of the transport loop's launch-side knobs on the caller's own
workload.

The transport results do not depend on these knobs (only on the physics
configuration), but the speed does:

- ``sort_every``: how often the tiles are reordered by cell and energy band
  (``rrea_tile_sort_banded``; 0 = never).  Cell order lets a workgroup's net
  chords and source records go through its LDS window and makes the field
  gathers cache-resident; the sort itself costs a radix sort and a gather of
  every array of the tile.
- ``energy_bits``: the width of the energy-band part of the sort key (0 =
  cell order only, up to 6 = quarter-octave bands; a width b merges
  2^(6 - b) quarter-octaves, so every width spans 16 octaves).  Bands
  keep a wavefront's lanes on similar branches and table brackets.

- ``block``: the fast kernels' workgroup size (128, 256 or 512 lanes;
  ``rrea_fields_set_block``): a workgroup keeps one LDS deposit window and
  one partial row of tallies, so the size trades window flushes and fold
  rows against scheduling granularity.

``tune`` times every sort candidate over one sort period of the real
workload, ``tune_block`` every workgroup size over a fixed number of steps
(the steps are real: the caller's simulation advances through them, as
through a warm-up); each keeps the fastest per step.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Iterable

import torch


@dataclass
class Choice:
    sort_every: int
    energy_bits: int
    ms_per_step: float
    trials: list = field(default_factory=list)  # (sort_every, energy_bits, ms_per_step) of every candidate
    steps_run: int = 0  # workload steps the tuning advanced through

    def as_dict(self) -> dict:
        return {"sort_every": self.sort_every, "energy_bits": self.energy_bits, "ms_per_step": self.ms_per_step,
                "trials": [list(t) for t in self.trials], "steps_run": self.steps_run}


def tune(step: Callable[[int, bool, int], None], sort_every: Iterable[int] = (0, 10, 25, 50),
         energy_bits: Iterable[int] = (0, 3, 6), unsorted_steps: int = 10, warmup: int = 2,
         first_step: int = 0) -> Choice:
    """``step(s, sort_now, energy_bits)`` runs step ``s`` of the workload,
    first reordering the tiles with ``energy_bits`` band bits when
    ``sort_now``.

    After ``warmup`` untimed steps, each candidate is timed over one whole
    period: a sort and the ``sort_every`` steps that follow it, so that the
    sort's cost is amortised over the interval and the order's decay over it
    is in the time.  The never-sort candidate (0) is timed first, over
    ``unsorted_steps`` steps, while the tiles are still in the caller's
    order.  Step numbers run on from ``first_step``; the fastest candidate
    per step wins, ties to the first timed."""
    every_list = sorted(set(int(e) for e in sort_every), key=lambda e: (e != 0, e))
    if any(e < 0 for e in every_list):
        raise ValueError("sort intervals must be >= 0")
    bits_list = [int(b) for b in energy_bits]
    s = first_step
    for _ in range(warmup):
        step(s, False, 0)
        s += 1
    trials = []
    for every in every_list:
        for bits in (bits_list if every > 0 else [0]):
            n = every if every > 0 else unsorted_steps
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for k in range(n):
                step(s, every > 0 and k == 0, bits)
                s += 1
            torch.cuda.synchronize()
            trials.append((every, bits, (time.perf_counter() - t0) / n * 1e3))
    best = min(trials, key=lambda t: t[2])
    return Choice(best[0], best[1], best[2], trials, s - first_step)


@dataclass
class BlockChoice:
    block: int
    ms_per_step: float
    trials: list = field(default_factory=list)  # (block, ms_per_step) of every candidate
    steps_run: int = 0

    def as_dict(self) -> dict:
        return {"block": self.block, "ms_per_step": self.ms_per_step, "trials": [list(t) for t in self.trials],
                "steps_run": self.steps_run}


def tune_block(step: Callable[[int], None], set_block: Callable[[int], None], blocks: Iterable[int] = (128, 256, 512),
               steps: int = 10, warmup: int = 1, first_step: int = 0, default: int = 256,
               margin: float = 0.02) -> BlockChoice:
    """The fast kernels' workgroup size (rrea_fields_set_block): each
    candidate is set, run for ``warmup`` untimed steps, then timed over
    ``steps`` steps of the workload (``step(s)`` runs step s, in whatever
    tile order the caller keeps).  The fastest per step wins, ties to the
    first timed, and is left set; the ``default`` size (when it is a
    candidate) is kept unless another is faster by more than ``margin`` of
    its time, since a few steps' timings differ by about a percent from run
    to run."""
    s = first_step
    trials = []
    for b in blocks:
        set_block(int(b))
        for _ in range(warmup):
            step(s)
            s += 1
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step(s)
            s += 1
        torch.cuda.synchronize()
        trials.append((int(b), (time.perf_counter() - t0) / steps * 1e3))
    best = min(trials, key=lambda t: t[1])
    base = [t for t in trials if t[0] == default]
    if base and best[1] >= base[0][1] * (1.0 - margin):
        best = base[0]
    set_block(best[0])
    return BlockChoice(best[0], best[1], trials, s - first_step)
