"""The benchmark workloads at their full panel sizes on the device, against the
CPU oracle, through the C ABI (stream contexts: the batched engine that
bench.py times).

  workload 1  2,000 x 10,000 panel, 20 targets, K = 1,000: every target in
              full (best haps, monos, split sites, count0, rows, digests),
              through a stream context and the synchronous engine
  workload 2  10,000 x 100,000 panel, K = 10,000 = N (whole-panel targets):
              8 targets by digest plus every small intermediate, one in full
              rows, under the PBWT variants the autotuner chooses among
  workload 3  20,000 x 200,000 panel, windows 20k-200k, K in {2k..10k}: the
              shortest and longest windows and every K among 12 targets, by
              digest plus every small intermediate, the shortest in full rows
  workload 4  200,000 x 500,000 panel, MAF ~ Beta(0.3, 0.3), K = 10,000:
              4 targets by digest plus every small intermediate, the oracle
              spread over host threads
  workload 5  the serving loop with the workload-3 panel resident: seeded
              Poisson request batches of 1-256 targets, every request equal
              alone and in the mix, and an oracle sample

The generator is prefix-stable (target i is the same for any target count),
so these are the benchmark's own first targets.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from eagleimp_amd import serve, synth
from eagleimp_amd.engine import Engine, pbwt_variants

from conftest import say
from parity import check_target

pytestmark = pytest.mark.gpu

GB = 1 << 30


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def _engine(panel):
    eng = Engine(0)
    eng.initialize(2, panel.M, panel.N, panel.N, panel.site)
    eng.set_panel_extras(panel.cM, hap=panel.hap)
    return eng


def _oracle(pv, tg, idx, rows_for=()):
    """Oracle results of targets idx on host threads (ctypes releases the GIL);
    cPBWT rows kept only for the targets in rows_for."""
    def one(t):
        return oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t],
                                     rows=t in rows_for)
    with ThreadPoolExecutor(max_workers=min(len(idx), _threads())) as ex:
        return list(ex.map(one, idx))


def _submit(st, tg, idx):
    st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx], [tg.W[t] for t in idx],
              [tg.K[t] for t in idx])
    res, nd, _ = st.wait()
    assert nd == 0, "targets deferred"
    return res


def _compare(src, res, refs, idx, tag, rows_for=(), layout="gpu"):
    """src: a Stream or the Engine (its last batch).  Summary, digests and every
    small intermediate of every target; rows for the targets in rows_for."""
    for j, t in enumerate(idx):
        r = refs[j]
        what = f"{tag} target {t} (K={r['K']}, W={r['W']}, Mp={r['Mp']})"
        assert int(res.S[j]) == r["S"], f"{what}: S"
        assert int(res.Mp[j]) == r["Mp"], f"{what}: Mp"
        assert int(res.ncalls[j]) == r["ncalls"], f"{what}: ncalls"
        assert int(res.dig_pbwt[j]) == r["dig_pbwt"], f"{what}: pbwt digest"
        assert int(res.dig_count0[j]) == r["dig_count0"], f"{what}: count0 digest"
        got = src.copy_target(j, rows=t in rows_for)
        check_target(r, got, int(r["K"]), layout, what, monos=src.copy_monos(j))


# ---------------------------------------------------------------------------
def test_workload1_cpu_panel_every_target(built):
    cfg, panel, tg = synth.make_workload(1)
    T = len(tg)
    say("workload", cfg["name"], "targets", T)
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    idx = list(range(T))
    refs = _oracle(pv, tg, idx, rows_for=set(idx))
    eng = _engine(panel)
    st = eng.stream(T, panel.M, cfg["K"], 2 * GB)
    res = _submit(st, tg, idx)
    _compare(st, res, refs, idx, "workload 1 stream", rows_for=set(idx))
    st.close()
    res = eng.run_batch(tg.gt, tg.s0, tg.W, tg.K, digest=True, compact=True)
    _compare(eng, res, refs, idx, "workload 1 batch", rows_for=set(idx), layout="compact")
    eng.close()


@pytest.fixture(scope="module")
def wl2():
    cfg, panel, tg = synth.make_workload(2, T=8)
    return cfg, panel, tg


def test_workload2_whole_panel_targets(built, wl2):
    cfg, panel, tg = wl2
    idx = list(range(len(tg)))
    say("workload", cfg["name"], "targets", len(idx))
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    refs = _oracle(pv, tg, idx, rows_for={0})
    eng = _engine(panel)
    Kpad = (cfg["K"] + 63) // 64 * 64
    # the variants the autotuner chooses among at this K (all that hold it)
    vs = [v["variant"] for v in pbwt_variants() if v["max_k"] >= Kpad]
    assert vs
    for v in vs:
        say("workload 2 PBWT variant", v)
        eng.set_pbwt_variant(v)
        st = eng.stream(len(idx), panel.M, cfg["K"], 8 * GB)
        res = _submit(st, tg, idx)
        _compare(st, res, refs, idx, f"workload 2 variant {v}", rows_for={0} if v == vs[0] else ())
        st.close()
    eng.close()


@pytest.fixture(scope="module")
def wl3():
    cfg, panel, tg = synth.make_workload(3, T=400)
    return cfg, panel, tg


def _wl3_pick(tg, n=12):
    """Targets covering the shortest and longest windows and every K."""
    W = tg.W.astype(np.int64)
    pick = [int(np.argmin(W)), int(np.argmax(W))]
    for k in sorted(set(int(k) for k in tg.K)):
        cand = [i for i in np.argsort(W) if int(tg.K[i]) == k and i not in pick]
        pick.append(int(cand[len(cand) // 2]))
    i = 0
    while len(pick) < n:
        if i not in pick:
            pick.append(i)
        i += 1
    return pick


def test_workload3_windows_and_mixed_K(built, wl3):
    cfg, panel, tg = wl3
    idx = _wl3_pick(tg)
    assert set(int(tg.K[t]) for t in idx) == set(cfg["kset"])
    say("workload", cfg["name"], "targets", idx, "W", [int(tg.W[t]) for t in idx],
        "K", [int(tg.K[t]) for t in idx])
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    shortest = idx[0]
    refs = _oracle(pv, tg, idx, rows_for={shortest})
    eng = _engine(panel)
    st = eng.stream(len(idx), int(tg.W.max()), int(tg.K.max()), 16 * GB)
    res = _submit(st, tg, idx)
    _compare(st, res, refs, idx, "workload 3", rows_for={shortest})
    st.close()
    eng.close()


def test_workload4_large_panel(built):
    T = 4
    cfg, panel, tg = synth.make_workload(4, T=T, nthreads=_threads())
    idx = list(range(T))
    say("workload", cfg["name"], "panel", panel.N, "x", panel.M, "targets", T)
    eng = _engine(panel)
    st = eng.stream(T, panel.M, cfg["K"], 24 * GB)
    res = _submit(st, tg, idx)
    say("workload 4 device done; oracle on", min(T, _threads()), "threads")
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    refs = _oracle(pv, tg, idx)
    _compare(st, res, refs, idx, "workload 4")
    st.close()
    eng.close()


def test_workload5_serving_on_workload3_panel(built, wl3):
    """Continuous batching: requests overlap in two contexts; each request's
    results equal the same request run alone; a sample equals the oracle."""
    cfg5 = synth.WORKLOADS[5]
    _, panel, _ = wl3  # the serving workload's panel is the workload-3 panel (seed 3)
    npool = 96
    pool = synth.make_targets(panel, 5, npool, kset=cfg5["kset"], wmin=cfg5["wmin"],
                              wmax=cfg5["wmax"], miss=cfg5["miss"], miss_hi=cfg5["miss_hi"])
    eng = _engine(panel)
    srv = serve.Server(eng, pool, int(pool.W.max()), int(pool.K.max()), contexts=2,
                       max_batch=128, pool_bytes=12 * GB)
    t, size, first = synth.make_arrivals(cfg5["panel_seed"] + 2, 10, 50.0, npool,
                                         cfg5["bmin"], cfg5["bmax"])
    mix = serve.make_requests(t, size, first, npool)
    say("workload 5 serving:", len(mix), "requests, sizes", [int(x) for x in size])
    wall = serve.run_requests(srv, mix)
    s = serve.summarize(mix, wall)
    say("mix", s)
    assert s["requests"] == len(mix)
    for r in mix:
        alone = serve.make_requests([0.0], [len(r.targets)], [r.targets[0]], npool)
        serve.run_requests(srv, alone, realtime=False)
        assert alone[0].results == r.results, f"request {r.rid}"
    srv.close()
    eng.close()
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    byt = {}
    for r in mix:
        for p, i in enumerate(r.targets):
            byt[i] = r.results[p]
    sample = sorted(byt)[::max(1, len(byt) // 10)]
    say("workload 5 oracle sample", sample)
    refs = _oracle(pv, pool, sample)
    for i, ref in zip(sample, refs):
        S, Mp, ncalls, K, dig, dig0 = byt[i]
        assert (S, Mp, ncalls, dig, dig0) == (ref["S"], ref["Mp"], ref["ncalls"], ref["dig_pbwt"],
                                              ref["dig_count0"]), f"target {i}"


def test_pbwt_variants_at_K10000(built):
    """The PBWT kernel shapes the autotuner picks at K = 10,000 (every variant
    that holds K; 20 or 24 slots per wave), on a short panel: rows, count0
    and digests of 4 whole-panel and 4 subset targets against the oracle."""
    panel = synth.make_panel(21, N=10048, M=400, F=64, T=8)
    tg = synth.make_targets(panel, 21, 8, kset=[10048, 10000, 9999, 9473], miss=0.01)
    idx = list(range(8))
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    refs = _oracle(pv, tg, idx, rows_for=set(idx))
    eng = _engine(panel)
    for v in [v["variant"] for v in pbwt_variants() if v["max_k"] >= 10048]:
        say("K=10,000 PBWT variant", v, "K", [int(k) for k in tg.K])
        eng.set_pbwt_variant(v)
        st = eng.stream(8, panel.M, panel.N, 2 * GB)
        res = _submit(st, tg, idx)
        _compare(st, res, refs, idx, f"variant {v}", rows_for=set(idx))
        st.close()
    eng.close()
