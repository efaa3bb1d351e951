"""Autotuner selection and cache logic (eagleimp_amd.tune) with stand-in
contexts whose timings are known; no device."""
import numpy as np

from eagleimp_amd import tune
from eagleimp_amd.engine import BatchResult


class _Targets:
    def __init__(self, n, K):
        self.gt = [np.zeros(16, np.uint64) for _ in range(n)]
        self.s0 = np.zeros(n, np.uint64)
        self.W = np.full(n, 1000, np.uint64)
        self.K = np.full(n, K, np.uint32)


class _Stream:
    def __init__(self, eng, batch):
        self.eng, self.batch, self.n = eng, batch, 0

    def submit(self, gts, s0, W, K):
        self.n = len(gts)
        self.eng.submits += 1

    def wait(self):
        a = np.ones((self.n, 6), np.uint64)
        a[self.eng.cap:, 1] = 0  # Mp == 0: deferred, the pool is full
        return BatchResult(*(a[:, i] for i in range(6))), max(0, self.n - self.eng.cap), (0.0, 0.0)

    def kernel_ms(self):
        v = self.eng.variant
        # variant 3 is fastest per target, and larger batches amortise better;
        # preparation runs for every submitted target, the rest for those placed
        per = self.eng.per.get(v, 9.0)
        done = min(self.n, self.eng.cap)
        return {"distance": 0.5 * self.n, "select": 0.1, "monos": 0.0, "splits": 0.1,
                "cref": 0.2 * done, "pbwt": per * done * (1.0 if self.batch > 128 else 1.2),
                "copy": 5.0}

    def close(self):
        pass


class _Engine:
    Nad = Ncorr = 20000
    M = 200000

    def __init__(self, per=None, cap=10**9):
        self.variant = -1
        self.submits = 0
        self.cap = cap  # targets a pool holds
        # device ms per target of each PBWT variant
        self.per = per or {2: 2.0, 3: 1.0, 4: 1.5}

    def set_pbwt_variant(self, v):
        self.variant = v

    def mem_info(self):
        return 200 << 30, 288 << 30

    def stream(self, b, W, K, pool, timed=False, graph=True, digest=True):
        assert timed and not graph
        return _Stream(self, b)


def test_candidates_hold_k():
    for K in (100, 2000, 10000, 16384):
        for v in tune.candidates(K):
            assert v["max_k"] >= (K + 63) // 64 * 64
    # 512 x 20, 512 x 32, 1024 x 24, then 512 x 20 and 1024 x 24 with the DPP wave scan
    assert [v["variant"] for v in tune.candidates(10000)] == [2, 3, 4, 5, 6]


def test_tune_picks_fastest_and_caches(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    eng = _Engine()
    tg = _Targets(300, 10000)
    best = tune.tune(eng, tg, range(300), batches=(128, 256))
    assert (best["variant"], best["batch"]) == (3, 256) and not best["cached"]
    assert eng.variant == 3
    nv = len(tune.candidates(10000))
    assert len(best["table"]) == nv * 2 and eng.submits == nv * 2
    eng2 = _Engine()
    again = tune.tune(eng2, tg, range(300), batches=(128, 256))
    assert again["cached"] and again["variant"] == 3 and eng2.submits == 0 and eng2.variant == 3


def test_near_tie_goes_to_lowest_variant(tmp_path, monkeypatch):
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    eng = _Engine(per={2: 1.005, 3: 1.0, 4: 1.5})  # 2 is within 2% of 3
    best = tune.tune(eng, _Targets(300, 10000), range(300), batches=(128, 256), use_cache=False)
    assert (best["variant"], best["batch"]) == (2, 256)


def test_batches_capped_at_pool_capacity(tmp_path, monkeypatch):
    """When a pool holds fewer targets than a candidate batch, the tuner
    measures at the capacity and never picks a batch that defers targets."""
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    eng = _Engine(cap=150)
    best = tune.tune(eng, _Targets(300, 10000), range(300), batches=(256, 300), use_cache=False)
    assert best["batch"] == 150 and best["variant"] == 3
    assert all(r["batch"] <= 150 or r["deferred"] > 0 for r in best["table"])
    assert {r["batch"] for r in best["table"] if r["deferred"] == 0} == {150}


def test_capacity_estimate_tried_once(tmp_path, monkeypatch):
    """After the first deferral the tuner estimates the capacity of the
    smaller submission's pool from the targets' sizes; when that estimate
    still defers, what it placed is taken: two trials, not a descent."""
    monkeypatch.setattr(tune, "CACHE", str(tmp_path / "tune.json"))
    eng = _Engine(cap=150, per={2: 1.0})
    monkeypatch.setattr(tune, "candidates", lambda K: [dict(variant=2, threads=512,
                                                            chunks_per_wave=20, max_k=10240)])
    best = tune.tune(eng, _Targets(300, 10000), range(300), batches=(256,), use_cache=False)
    assert [r["batch"] for r in best["table"]] == [256, 255, 150]
    assert best["batch"] == 150 and eng.submits == 3


def test_pool_budget_sets_aside_contexts_and_margin():
    """pool_budget: free memory less `contexts` times one context's own
    buffers (measured as the drop in free memory when a context with a 1 MiB
    pool is made) and the margin."""
    GiB = 1 << 30

    class _MemEngine(_Engine):
        def __init__(self, ctx_bytes):
            super().__init__()
            self.free, self.ctx = 200 * GiB, ctx_bytes

        def mem_info(self):
            return self.free, 288 * GiB

        def stream(self, b, W, K, pool, timed=False, graph=True, digest=True):
            eng = self
            eng.free -= self.ctx + pool

            class _S(_Stream):
                def close(s):
                    eng.free += eng.ctx + pool
            return _S(self, b)

    eng = _MemEngine(3 * GiB)
    assert tune.context_bytes(eng, 124, 500000, 10000) == 3 * GiB
    assert eng.free == 200 * GiB  # the probe context is closed again
    assert tune.pool_budget(eng, 2, 124, 500000, 10000) == 200 * GiB - 2 * 3 * GiB - 4 * GiB
    assert tune.pool_budget(eng, 1, 124, 500000, 10000, margin=0) == 197 * GiB
