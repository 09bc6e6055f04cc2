"""Host logic of the launch-shape autotuner (no GPU): which candidates reach
the timing loop, and that the winner is applied and cached."""
import json

import numpy as np


class _Batch:
    def __init__(self, n):
        self.n = np.asarray(n, dtype=np.int32)
        self.count = len(self.n)
        self.threads = None

    def set_threads(self, t):
        self.threads = t

    def run(self, nsteps, use_graph=True):
        pass


class _Tuner:
    """Records the candidates and picks the largest."""

    def __init__(self):
        self.seen = None

    def tune(self, kernel, key, candidates, apply, run):
        self.seen = (kernel, key, list(candidates))
        best = max(candidates)
        apply(best)
        return best


def test_burgers_threads_skip_block_sizes_that_cannot_cover_the_instance(ec):
    """The persistent Burgers kernel is built with at most min(64, 32768 / t)
    points per thread for t threads, so 128 threads cover 8192 points: a
    batch holding 16384-point instances must never time or pick 128 (its
    launch would leave the upper half of those instances untouched)."""
    A = ec.autotune
    tn = _Tuner()
    b = _Batch([2048, 16384, 4096])
    assert A.tune_burgers_threads(tn, b) == 1024 and b.threads == 1024
    assert tn.seen[2] == [256, 512, 1024] and tn.seen[1] == "count3|n16384"
    b = _Batch([2048, 8192])
    A.tune_burgers_threads(tn, b)
    assert tn.seen[2] == [128, 256, 512, 1024]
    b = _Batch([32768])
    A.tune_burgers_threads(tn, b)
    assert tn.seen[2] == [512, 1024]


def test_autotuner_cache_round_trip(ec, tmp_path):
    """tune() times every candidate once, applies and caches the fastest;
    a second tuner on the same cache file applies the cached winner without
    timing anything."""
    A = ec.autotune
    path = str(tmp_path / "cache.json")
    t = A.Autotuner(path=path)
    cost = {64: 3.0, 128: 1.0, 256: 2.0}
    state = {"c": None, "runs": 0}
    t.time_ms = lambda run: cost[state["c"]]
    best = t.tune("k", "w", [64, 128, 256], lambda c: state.__setitem__("c", c), lambda: None)
    assert best == 128 and state["c"] == 128
    doc = json.load(open(path))
    assert doc[f"{t.device}|k|w"]["best"] == "128"
    t2 = A.Autotuner(path=path)
    t2.time_ms = lambda run: (_ for _ in ()).throw(AssertionError("cached shape was timed again"))
    state["c"] = None
    assert t2.tune("k", "w", [64, 128, 256], lambda c: state.__setitem__("c", c), lambda: None) == 128
    assert state["c"] == 128 and t2.log[-1]["cached"]
