"""The autotuner's search (host logic only, with a synthetic timer)."""
from msim import autotune


def test_descent_per_layout_finds_a_layout_that_needs_its_own_segment_size():
    # binning 3 is slow at the automatic segment size and fastest at 2^12;
    # a single descent that settles the layout first would keep binning 2
    def timer(o):
        if o["binning"] == 3:
            return 500.0 if o["segment_log2"] == 12 else 800.0
        if o["binning"] == 2:
            return 700.0 - (5.0 if o["bin_unroll"] == 8 else 0.0)
        return 900.0

    single, single_us, _ = autotune._descend(timer, autotune.DEFAULT, autotune.SPACE, lambda m: None)
    assert single["binning"] == 2 and single_us == 695.0
    best, us, trials = autotune._descend_per_layout(timer, autotune.DEFAULT, autotune.SPACE, lambda m: None)
    assert best["binning"] == 3 and best["segment_log2"] == 12 and us == 500.0
    assert trials[0][0] == autotune.DEFAULT  # the first trial is the library default
    assert len(trials) > len(autotune.SPACE["binning"]) * sum(len(v) - 1 for k, v in autotune.SPACE.items()
                                                             if k != "binning") // 2


def test_batch_space_covers_the_sorted_state():
    assert 3 in autotune.BATCH_SPACE["binning"] and 2 in autotune.BATCH_SPACE["bin_unroll"]
    assert 3 in autotune.SPACE["binning"] and 2 in autotune.SPACE["bin_unroll"]


class _Handle:
    """Records run_ticks / sync / time_ticks calls in order."""

    def __init__(self):
        self.calls = []

    def run_ticks(self, n):
        self.calls.append(("run", n))

    def sync(self):
        self.calls.append(("sync",))

    def time_ticks(self, n, kernels=False):
        self.calls.append(("time", n))
        return 1.0, {}


def test_timed_run_replays_graphs_captured_before_it():
    # the warm-up, then one untimed run of the timed length rounded up to an
    # even number of ticks (it ends on the buffers it started from), so the
    # graph segments the timed run replays are captured outside it
    for ticks, primed in ((100, 100), (101, 102), (200, 200)):
        h = _Handle()
        autotune._prime(h, 20, ticks)
        assert h.calls == [("run", 20), ("run", primed), ("sync",)]
