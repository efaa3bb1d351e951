"""Host logic of the serving loop (eagleimp_amd.serve) against stand-in
contexts, no device: every request completes, results do not depend on what
else is in flight, deferred targets are admitted again, and a target that can
never fit is reported instead of looping."""
import numpy as np
import pytest

from eagleimp_amd import serve
from eagleimp_amd.engine import BatchResult


class _Targets:
    def __init__(self, n):
        self.gt = [np.full(4, i, np.uint64) for i in range(n)]
        self.s0 = np.zeros(n, np.uint64)
        self.W = np.arange(1, n + 1, dtype=np.uint64)
        self.K = (np.arange(n, dtype=np.uint64) % 5) + 2


def _result_of(gt):
    i = int(gt[0])
    return (3 * i, 6 * i + 1, i, 0, (i * 2654435761) % (1 << 32), i ^ 0x55)


class _FakeStream:
    """Completes a submission after `lag` polls; fits `cap` targets per submission."""

    def __init__(self, cap, lag):
        self.cap, self.lag = cap, lag
        self.items = None
        self.polls = 0

    def submit(self, gts, s0, W, K):
        assert self.items is None
        self.items = [_result_of(g) for g in gts]
        self.polls = 0

    def done(self):
        self.polls += 1
        return self.polls > self.lag

    def wait(self):
        rows = []
        for j, r in enumerate(self.items):
            rows.append(r if j < self.cap else (r[0], 0, r[2], r[3], 0, 0))
        a = np.array(rows, np.uint64).reshape(-1, 6)
        self.items = None
        nd = sum(1 for j in range(len(rows)) if j >= self.cap)
        return BatchResult(S=a[:, 0], Mp=a[:, 1], ncalls=a[:, 2], K=a[:, 3], dig_pbwt=a[:, 4],
                           dig_count0=a[:, 5]), nd, (0.0, 0.0)

    def close(self):
        pass


class _FakeEngine:
    def __init__(self, cap=1000, lags=(1, 3, 2)):
        self.cap, self.lags, self.made = cap, list(lags), 0

    def mem_info(self):
        return 64 << 30, 64 << 30

    def stream(self, max_batch, max_window, max_K, pool_bytes):
        s = _FakeStream(self.cap, self.lags[self.made % len(self.lags)])
        self.made += 1
        return s


def _run(eng, pool, reqs, contexts=3, max_batch=8):
    srv = serve.Server(eng, pool, 100, 10, contexts=contexts, max_batch=max_batch)
    wall = serve.run_requests(srv, reqs, realtime=False)
    srv.close()
    return srv, wall


def test_all_requests_complete_with_their_own_results():
    pool = _Targets(40)
    t = np.cumsum(np.full(30, 0.001))
    size = (np.arange(30) % 7) + 1
    first = (np.arange(30) * 11) % 40
    reqs = serve.make_requests(t, size, first, 40)
    srv, wall = _run(_FakeEngine(), pool, reqs)
    for r in reqs:
        assert r.finish is not None and r.latency >= 0
        for p, i in enumerate(r.targets):
            assert r.results[p] == _result_of(pool.gt[i])
    s = serve.summarize(reqs, wall)
    assert s["targets"] == int(size.sum()) and s["p99_ms"] >= s["p50_ms"]


def test_same_results_alone_and_in_the_mix():
    pool = _Targets(25)
    reqs = serve.make_requests([0.0] * 6, [5, 1, 9, 3, 8, 2], [0, 7, 3, 20, 11, 5], 25)
    _run(_FakeEngine(), pool, reqs, max_batch=6)
    for r in reqs:
        alone = serve.make_requests([0.0], [len(r.targets)], [r.targets[0]], 25)
        _run(_FakeEngine(lags=(0,)), pool, alone, contexts=1, max_batch=32)
        assert alone[0].results == r.results


def test_deferred_targets_are_admitted_again():
    pool = _Targets(30)
    reqs = serve.make_requests([0.0, 0.0, 0.0], [10, 10, 10], [0, 10, 20], 30)
    srv, _ = _run(_FakeEngine(cap=3), pool, reqs, contexts=2, max_batch=8)
    assert srv.deferred > 0
    for r in reqs:
        assert len(r.results) == 10
        for p, i in enumerate(r.targets):
            assert r.results[p] == _result_of(pool.gt[i])


def test_target_that_never_fits_raises():
    pool = _Targets(4)
    reqs = serve.make_requests([0.0], [2], [0], 4)
    with pytest.raises(RuntimeError, match="exceed"):
        _run(_FakeEngine(cap=0), pool, reqs, contexts=1, max_batch=2)
