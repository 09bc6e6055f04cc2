"""Host logic of the launch-shape autotuner (no GPU): which candidates reach
the timing loop, and that the winner is applied and cached."""
import json

import numpy as np


class _Batch:
    """Stands in for BurgersBatch: refuses the block sizes in `refuse` the way
    ec_burgers_batch_set_threads does (an ECError)."""

    def __init__(self, n, refuse=(), c=None, dtype="f32"):
        self.n = np.asarray(n, dtype=np.int32)
        self.c = np.asarray(c if c is not None else [4] * len(n), dtype=np.int32)
        self.count = len(self.n)
        self.dtype = dtype
        self.refuse = set(refuse)
        self.threads = None

    def set_threads(self, t):
        from exactclosure_jl_amd._abi import ECError

        if t in self.refuse:
            raise ECError(-1, f"threads={t} cannot cover the largest instance")
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


def test_burgers_threads_skip_block_sizes_the_library_refuses(ec):
    """A block size that ec_burgers_batch_set_threads refuses (its threads
    cannot cover the largest instance, e.g. 128 threads for 16384 points) is
    never timed or picked: the library's own check decides, not a copy of it."""
    A = ec.autotune
    tn = _Tuner()
    b = _Batch([2048, 16384, 4096], refuse={128})
    assert A.tune_burgers_threads(tn, b) == 1024 and b.threads == 1024
    assert tn.seen[2] == [256, 512, 1024]
    b = _Batch([2048, 8192])
    A.tune_burgers_threads(tn, b)
    assert tn.seen[2] == [128, 256, 512, 1024]
    b = _Batch([32768], refuse={128, 256})
    A.tune_burgers_threads(tn, b)
    assert tn.seen[2] == [512, 1024]


def test_burgers_threads_cache_key_separates_precision_and_compressions(ec):
    """The LDS per block, and so the occupancy and the winner, depends on the
    state precision and the compression mix: batches that differ only there
    get different cache keys."""
    A = ec.autotune
    keys = set()
    for dtype in ("f32", "f64"):
        for c in ([4, 4, 64], [4, 64, 64]):
            tn = _Tuner()
            A.tune_burgers_threads(tn, _Batch([2048, 16384, 4096], c=c, dtype=dtype))
            keys.add(tn.seen[1])
    assert len(keys) == 4
    assert all(k.startswith("count3|n16384|") for k in keys)


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
