"""The autotuner's selection (autotune.tune) on a synthetic workload whose
per-step cost is known: every candidate is timed over one whole sort
period, the never-sort candidate first, and the cheapest per step wins."""
import time

import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import autotune  # noqa: E402


class FakeWorkload:
    """Steps cost 2 units unsorted; after a sort with b band bits a step costs
    (1 + 0.02 k - 0.05 b) units, k steps after the sort; a sort costs 5 units
    (a unit sleeps 2 ms)."""

    def __init__(self):
        self.since = None
        self.bits = 0
        self.calls = []

    def __call__(self, s, sort_now, bits):
        self.calls.append((s, sort_now, bits))
        cost = 0.0
        if sort_now:
            cost += 5.0
            self.since, self.bits = 0, bits
        if self.since is None:
            cost += 2.0
        else:
            cost += 1.0 + 0.02 * self.since - 0.05 * self.bits
            self.since += 1
        time.sleep(cost * 2e-3)  # doubled: the margins between candidates stay well above sleep jitter


@pytest.fixture(autouse=True)
def no_device(monkeypatch):
    monkeypatch.setattr(autotune.torch.cuda, "synchronize", lambda: None)


def test_tune_amortises_the_sort_and_its_decay():
    w = FakeWorkload()
    ch = autotune.tune(w, sort_every=(50, 0, 10, 25), energy_bits=(0, 3), unsorted_steps=4, warmup=2, first_step=7)
    # per step: never 2.0; every E with b bits: 5/E + 1 + 0.01 (E - 1) - 0.05 b
    # E=10: 1.59 - .05b; E=25: 1.44 - .05b; E=50: 1.59 - .05b
    assert (ch.sort_every, ch.energy_bits) == (25, 3)
    assert [t[:2] for t in ch.trials] == [(0, 0), (10, 0), (10, 3), (25, 0), (25, 3), (50, 0), (50, 3)]
    # the never-sort candidate is timed before any sort
    assert not any(c[1] for c in w.calls[:6])
    assert ch.steps_run == 2 + 4 + 2 * (10 + 25 + 50) == len(w.calls)
    assert [c[0] for c in w.calls] == list(range(7, 7 + ch.steps_run))
    # one sort per period, on its first step
    assert sum(c[1] for c in w.calls) == 6
    d = ch.as_dict()
    assert d["sort_every"] == 25 and len(d["trials"]) == 7


def test_tune_rejects_negative_interval():
    with pytest.raises(ValueError):
        autotune.tune(FakeWorkload(), sort_every=(-1,))


def test_tune_block_picks_the_cheapest_and_leaves_it_set():
    """tune_block times each workgroup size over the same number of steps
    after its warm-up and leaves the fastest set."""
    cost = {128: 3.0, 256: 1.0, 512: 2.0}
    state = {"block": None, "sets": []}

    def set_block(b):
        state["block"] = b
        state["sets"].append(b)

    def step(s):
        time.sleep(cost[state["block"]] * 2e-3)

    c = autotune.tune_block(step, set_block, blocks=(128, 256, 512), steps=4, warmup=1, first_step=7)
    assert c.block == 256 and state["block"] == 256
    assert [t[0] for t in c.trials] == [128, 256, 512]
    assert state["sets"] == [128, 256, 512, 256]
    assert c.steps_run == 15


def test_tune_block_keeps_the_default_within_the_margin():
    # 128 is timed faster than 256 by far less than the margin
    state = {"block": 0}

    def set_block(b):
        state["block"] = b

    times = iter([0.0, 1.0, 1.0, 2.0 + 1e-9, 2.0 + 1e-9, 4.0])
    orig = autotune.time.perf_counter
    autotune.time.perf_counter = lambda: next(times)
    try:
        c = autotune.tune_block(lambda s: None, set_block, blocks=(128, 256, 512), steps=1, warmup=0)
    finally:
        autotune.time.perf_counter = orig
    assert c.block == 256 and state["block"] == 256
    assert c.trials[0][1] < c.trials[1][1]
