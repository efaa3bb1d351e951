"""Stream contexts (eimp_stream_*): the device-driven batch path -- planner,
queue-driven cref/PBWT kernels, captured hipGraph -- vs the CPU oracle, bit
for bit, through the C ABI.  Also: graph replay == direct launches, deferral
when the output pool is short, and results of a context alone == the same
context's results while other contexts run at the same time.
"""
import numpy as np
import pytest

import oracle
from eagleimp_amd import synth
from eagleimp_amd.engine import EimpError

from conftest import say
from parity import check_target, panel_engine

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.fixture(scope="module")
def small():
    return synth.make_panel(12, N=300, M=4000, F=20, T=24)


def _refs(pv, tg, idx):
    return [oracle.process_target(pv if not callable(pv) else pv(t), tg.gt[t], tg.s0[t], tg.W[t],
                                  tg.K[t]) for t in idx]


def _compare(st, res, refs, idx, layout, tag, digest=True):
    for j, t in enumerate(idx):
        r = refs[j]
        what = f"{tag} target {j}"
        assert int(res.S[j]) == r["S"], f"{what}: S"
        assert int(res.Mp[j]) == r["Mp"], f"{what}: Mp"
        assert int(res.ncalls[j]) == r["ncalls"], f"{what}: ncalls"
        got = st.copy_target(j, cref=True)
        check_target(r, got, int(r["K"]), layout, what, monos=st.copy_monos(j))
        if digest:
            assert int(res.dig_pbwt[j]) == r["dig_pbwt"], f"{what}: pbwt digest"
            assert int(res.dig_count0[j]) == r["dig_count0"], f"{what}: count0 digest"


@pytest.mark.parametrize("layout", ["gpu", "compact", "raw"])
def test_stream_matches_oracle(built, small, layout):
    """Two submissions in a row on one context (the second smaller), mixed
    windows and K including the whole-panel identity case."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 41, 12, kset=[300, 64, 150, 299, 2, 257], wmin=60, wmax=4000,
                            miss=0.0, miss_hi=0.05)
    st = eng.stream(16, 4000, 300, 256 * MB, compact=layout == "compact", raw=layout == "raw")
    for idx in (list(range(12)), [3, 7, 1]):
        say("stream layout", layout, "targets", idx, "K", [int(tg.K[t]) for t in idx])
        st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx], [tg.W[t] for t in idx],
                  [tg.K[t] for t in idx])
        res, nd, (kms, sms) = st.wait()
        assert nd == 0
        assert 0 <= kms <= sms
        _compare(st, res, _refs(pv, tg, idx), idx, layout, f"stream {idx}")
    st.close()
    eng.close()


@pytest.mark.parametrize("layout", ["gpu", "raw"])
def test_stream_without_digest_matches_oracle(built, small, layout):
    """A context without the device digest, launched directly with events
    between kernel classes (how the benchmark's timed passes run): its rows,
    count0 and condensed reference, copied out, equal the oracle's."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 47, 8, kset=[300, 150, 63, 299], wmin=100, wmax=4000,
                            miss=0.0, miss_hi=0.05)
    st = eng.stream(8, 4000, 300, 256 * MB, raw=layout == "raw", digest=False, timed=True,
                    graph=False)
    idx = list(range(8))
    say("stream without digest", layout, "K", [int(tg.K[t]) for t in idx])
    st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx], [tg.W[t] for t in idx],
              [tg.K[t] for t in idx])
    res, nd, _ = st.wait()
    assert nd == 0
    assert not any(int(x) for x in res.dig_pbwt)  # no digest computed
    refs = _refs(pv, tg, idx)
    _compare(st, res, refs, idx, layout, f"no digest {layout}", digest=False)
    if layout == "raw":
        with pytest.raises(EimpError):  # raw rows hold no zero counts
            st.digest_outputs()
    else:
        # the digests recomputed afterwards from the stored rows equal the oracle's
        dp, dc = st.digest_outputs()
        for j in range(len(idx)):
            assert int(dp[j]) == refs[j]["dig_pbwt"], f"target {j}: digest of stored rows"
            assert int(dc[j]) == refs[j]["dig_count0"], f"target {j}: digest of stored count0"
    st.close()
    eng.close()


def test_stream_haploid_and_ignored(built, small):
    panel = small
    Nad = panel.N
    hapl = np.zeros(Nad // 2, np.uint8)
    hapl[[2, 9, 40]] = 1
    Ncorr = Nad - int(hapl.sum()) - 2
    ignore = [4, 33, 0, 149]
    eng = panel_engine(panel, Ncorr=Ncorr, haploid=hapl)
    tg = synth.make_targets(panel, 43, 4, kset=[30, 180, Ncorr], wmin=400, wmax=4000)
    pvs = [oracle.PanelView(panel.hap, panel.site, panel.cM, haploid_ref=hapl, ignore_sample=ig,
                            N_corr=Ncorr) for ig in ignore]
    st = eng.stream(4, 4000, Ncorr, 64 * MB)
    say("stream haploid", list(tg.K))
    st.submit(tg.gt, tg.s0, tg.W, tg.K, ignore=ignore)
    res, nd, _ = st.wait()
    assert nd == 0
    _compare(st, res, _refs(lambda t: pvs[t], tg, range(4)), range(4), False, "haploid")
    st.close()
    eng.close()


def test_graph_replay_equals_direct_launches(built, small):
    panel = small
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 45, 10, kset=[100, 300, 200], wmin=500, wmax=4000)
    out = []
    for mode in (dict(graph=True), dict(graph=False), dict(graph=False, timed=True)):
        st = eng.stream(10, 4000, 300, 128 * MB, **mode)
        reps = []
        for rep in range(3):  # replays with host work between them
            say("mode", mode, "replay", rep)
            st.submit(tg.gt, tg.s0, tg.W, tg.K)
            res, nd, _ = st.wait()
            assert nd == 0
            reps.append((res.S.copy(), res.dig_pbwt.copy(), res.dig_count0.copy()))
            if mode.get("timed"):
                ms = st.kernel_ms()
                assert ms["pbwt"] > 0 and ms["cref"] > 0 and ms["select"] > 0
            else:
                with pytest.raises(EimpError):
                    st.kernel_ms()
        for r in reps[1:]:
            for a, b in zip(reps[0], r):
                assert np.array_equal(a, b)
        out.append(reps[0])
        st.close()
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)
    eng.close()


def test_short_pool_defers_a_suffix(built, small):
    """Targets whose outputs do not fit are deferred (Mp = 0) and come out
    right when submitted again."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 47, 8, K=300, wmin=3000, wmax=4000)
    refs = _refs(pv, tg, range(8))

    def region(r):
        mp, kpad, rs = r["Mp"], (300 + 63) // 64 * 64, 4 * 8 * ((300 + 511) // 512)
        up = lambda x: (x + 255) // 256 * 256  # noqa: E731
        return up((mp + 15) // 16 * kpad * 2) + up(mp * 4) + up(2 * mp * rs * 4)

    need = [region(r) for r in refs]
    pool = sum(need[:3]) + need[3] // 2  # first three fit, the fourth does not
    st = eng.stream(8, 4000, 300, pool)
    say("short pool", pool, need)
    st.submit(tg.gt, tg.s0, tg.W, tg.K)
    res, nd, _ = st.wait()
    assert nd == 5
    assert list(res.Mp[3:]) == [0] * 5
    assert [int(s) for s in res.S] == [r["S"] for r in refs]
    _compare(st, res, refs[:3], range(3), False, "fits")
    with pytest.raises(EimpError):
        st.copy_target(4)
    idx = list(range(3, 6))
    st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx], [tg.W[t] for t in idx],
              [tg.K[t] for t in idx])
    res, nd, _ = st.wait()
    assert nd == 0
    _compare(st, res, refs[3:6], idx, False, "resubmitted")
    st.close()
    eng.close()


def test_contexts_alone_and_in_the_mix(built, small):
    """Each context's results are the same alone and while three contexts
    overlap on the device."""
    panel = small
    eng = panel_engine(panel)
    batches = [synth.make_targets(panel, 50 + i, n, kset=[64, 200, 300], wmin=200, wmax=4000)
               for i, n in enumerate((9, 3, 14))]
    sts = [eng.stream(16, 4000, 300, 128 * MB) for _ in batches]
    solo = []
    for st, tg in zip(sts, batches):
        st.submit(tg.gt, tg.s0, tg.W, tg.K)
        res, nd, _ = st.wait()
        assert nd == 0
        solo.append(res)
    for rnd in range(3):
        say("mix round", rnd)
        for st, tg in zip(sts, batches):
            st.submit(tg.gt, tg.s0, tg.W, tg.K)
        for st, ref in zip(sts, solo):
            res, nd, _ = st.wait()
            assert nd == 0
            for f in ("S", "Mp", "ncalls", "dig_pbwt", "dig_count0"):
                assert np.array_equal(getattr(res, f), getattr(ref, f)), f
    for st in sts:
        st.close()
    eng.close()


def test_stream_argument_errors(built, small):
    panel = small
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 60, 3, K=100, wmin=100, wmax=2000)
    with pytest.raises(EimpError):
        eng.stream(2000, 4000, 300, MB)  # more than 1024 targets
    st = eng.stream(2, 2000, 100, 16 * MB)
    with pytest.raises(EimpError):
        st.submit(tg.gt, tg.s0, tg.W, tg.K)  # 3 > max_targets
    with pytest.raises(EimpError):
        st.submit([tg.gt[0]], [0], [3000], [100])  # window beyond max_window
    with pytest.raises(EimpError):
        st.submit([tg.gt[0]], [0], [tg.W[0]], [200])  # K beyond max_K
    with pytest.raises(EimpError):
        st.wait()  # nothing submitted
    st.submit([tg.gt[0]], [tg.s0[0]], [tg.W[0]], [100])
    with pytest.raises(EimpError):
        st.submit([tg.gt[0]], [tg.s0[0]], [tg.W[0]], [100])  # not collected yet
    st.wait()
    st.close()
    eng.close()


@pytest.mark.parametrize("mode", ["graph", "timed"])
def test_shared_pool_contexts_take_turns(built, small, mode):
    """Two contexts on one output pool, submissions interleaved (the second
    context's preparation beside the first one's build): every submission
    equals the oracle; a context's pool outputs are refused once the other
    context has submitted again; the pool outlives the context that made it."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 53, 24, kset=[300, 64, 150, 299, 257], wmin=80, wmax=4000,
                            miss=0.0, miss_hi=0.05)
    kw = dict(timed=mode == "timed", graph=mode == "graph")
    a = eng.stream(12, 4000, 300, 128 * MB, **kw)
    b = eng.stream(12, 4000, 300, 0, share=a, **kw)
    batches = [list(range(0, 10)), list(range(10, 20)), list(range(20, 24)), [5, 15, 0]]
    ctx = [a, b, a, b]

    def submit(st, idx):
        st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx], [tg.W[t] for t in idx],
                  [tg.K[t] for t in idx])

    say("shared pool", mode, "submissions", len(batches))
    submit(a, batches[0])
    submit(b, batches[1])  # prepares while a builds
    for i in range(len(batches)):
        st, idx = ctx[i], batches[i]
        res, nd, _ = st.wait()
        assert nd == 0
        refs = _refs(pv, tg, idx)
        for j, t in enumerate(idx):
            assert int(res.Mp[j]) == refs[j]["Mp"], f"submission {i} target {j}: Mp"
            assert int(res.dig_pbwt[j]) == refs[j]["dig_pbwt"], f"submission {i} target {j}"
            assert int(res.dig_count0[j]) == refs[j]["dig_count0"], f"submission {i} target {j}"
        if i + 2 < len(batches):
            # the other context has submitted since: these pool outputs are gone
            with pytest.raises(EimpError):
                st.copy_target(0)
            with pytest.raises(EimpError):
                st.digest_outputs()
            submit(st, batches[i + 2])
    # the last submission's outputs are current: rows equal the oracle's
    _compare(ctx[-1], res, _refs(pv, tg, batches[-1]), batches[-1], "gpu", "shared pool, last")
    # the pool outlives the context that allocated it
    a.close()
    submit(b, [2, 3])
    res, nd, _ = b.wait()
    refs = _refs(pv, tg, [2, 3])
    assert [int(x) for x in res.dig_pbwt[:2]] == [r["dig_pbwt"] for r in refs]
    b.close()
    eng.close()


def test_shared_pool_contexts_on_two_threads(built, small):
    """Two contexts on one output pool, each driven by its own host thread
    (the library calls release the interpreter lock, so the submissions and
    reads interleave freely): every submission's digests equal the oracle's,
    and a read of pool rows either returns the oracle's rows or is refused
    because the other context has built since -- never rows half overwritten."""
    import threading
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    tg = synth.make_targets(panel, 59, 24, kset=[300, 64, 150, 299, 257], wmin=80, wmax=4000,
                            miss=0.0, miss_hi=0.05)
    refs = _refs(pv, tg, range(24))
    a = eng.stream(6, 4000, 300, 64 * MB)
    b = eng.stream(6, 4000, 300, 0, share=a)
    plans = {0: [list(range(0, 6)), list(range(12, 18)), [1, 3, 5]] * 2,
             1: [list(range(6, 12)), list(range(18, 24)), [20, 2]] * 2}
    tally = {0: [0, 0], 1: [0, 0]}  # (row reads that matched, reads refused)
    errors = []

    def drive(k, st):
        try:
            for i, idx in enumerate(plans[k]):
                st.submit([tg.gt[t] for t in idx], [tg.s0[t] for t in idx],
                          [tg.W[t] for t in idx], [tg.K[t] for t in idx])
                res, nd, _ = st.wait()
                assert nd == 0
                for j, t in enumerate(idx):
                    assert int(res.dig_pbwt[j]) == refs[t]["dig_pbwt"], f"thread {k} sub {i} tgt {j}"
                    assert int(res.dig_count0[j]) == refs[t]["dig_count0"], f"thread {k} sub {i}"
                j = i % len(idx)
                try:
                    got = st.copy_target(j, cref=True)
                except EimpError:
                    tally[k][1] += 1
                    continue
                check_target(refs[idx[j]], got, int(refs[idx[j]]["K"]), "gpu", f"thread {k} sub {i}")
                tally[k][0] += 1
        except BaseException as exc:  # reported by the main thread
            errors.append(exc)

    say("shared pool on two threads, submissions", len(plans[0]) + len(plans[1]))
    th = [threading.Thread(target=drive, args=(k, st)) for k, st in ((0, a), (1, b))]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=100)
    assert not any(x.is_alive() for x in th), "a driving thread did not finish"
    if errors:
        raise errors[0]
    say("row reads matched / refused", tally)
    assert sum(tally[0]) == len(plans[0]) and sum(tally[1]) == len(plans[1])
    a.close()
    b.close()
    eng.close()
