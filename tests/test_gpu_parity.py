"""Device engine vs the CPU oracle, bit for bit, through the C ABI (libeimp.so).

Every intermediate of the pre-stage is compared: best haps and their bitmap,
split sites / bitmap / cM differences, count0, the condensed reference rows
and the forward/backward compressed PBWT (drop-in GPU layout and compact CPU
layout), plus the device digests.
"""
import numpy as np
import pytest

import oracle
from eagleimp_amd import synth
from eagleimp_amd.engine import Engine, EimpError, unit_words

from conftest import say
from parity import check_target as _check_target, panel_engine as _panel_engine

pytestmark = pytest.mark.gpu


def _run_and_compare(panel, pv, eng, gts, s0s, Ws, Ks, ignore=None, compact=False, cref=True,
                     tag="", raw=False):
    say(tag, "targets", len(gts), "K", list(Ks), "W", list(Ws), "s0", list(s0s))
    res = eng.run_batch(gts, s0s, Ws, Ks, ignore=ignore, compact=compact, raw=raw, digest=True)
    for t in range(len(gts)):
        r = oracle.process_target(pv if not callable(pv) else pv(t), gts[t], s0s[t], Ws[t], Ks[t])
        what = f"{tag} target {t} (K={Ks[t]}, W={Ws[t]}, s0={s0s[t]})"
        assert int(res.S[t]) == r["S"], f"{what}: S"
        assert int(res.Mp[t]) == r["Mp"], f"{what}: Mp"
        assert int(res.ncalls[t]) == r["ncalls"], f"{what}: ncalls"
        got = eng.copy_target(t, cref=cref)
        _check_target(r, got, int(r["K"]), "raw" if raw else ("compact" if compact else "gpu"), what,
                      monos=eng.copy_monos(t))
        assert int(res.dig_pbwt[t]) == r["dig_pbwt"], f"{what}: pbwt digest"
        assert int(res.dig_count0[t]) == r["dig_count0"], f"{what}: count0 digest"
    return res


@pytest.fixture(scope="module")
def small():
    panel = synth.make_panel(11, N=300, M=4000, F=20, T=16)
    return panel


def test_topk_windows_and_identity(built, small):
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel)
    cases = [  # (s0, W, K)
        (0, 4000, 300), (0, 4000, 100), (137, 2500, 64), (1, 63, 33), (511, 1025, 2),
        (3000, 1000, 299), (64, 640, 150), (0, 4000, 1), (2999, 1001, 257), (77, 3923, 31),
    ]
    tg = synth.make_targets(panel, 11, len(cases), K=1, miss=0.02)
    gts, s0s, Ws, Ks = [], [], [], []
    for i, (s0, W, K) in enumerate(cases):
        gt = np.zeros(2 * unit_words(W), np.uint64)
        synth._lib().synth_target_gt(synth.ptr(panel.tgt_hap[2 * i]), synth.ptr(panel.tgt_hap[2 * i + 1]),
                                     s0, W, 0.02, 11, i, synth.ptr(gt))
        gts.append(gt); s0s.append(s0); Ws.append(W); Ks.append(K)
    del tg
    _run_and_compare(panel, pv, eng, gts, s0s, Ws, Ks, tag="windows")
    eng.close()


def test_compact_layout_and_given_hap(built, small):
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel, give_hap=True)
    tg = synth.make_targets(panel, 5, 6, kset=[50, 97, 300, 129], wmin=300, wmax=4000, miss=0.0,
                            miss_hi=0.05)
    _run_and_compare(panel, pv, eng, tg.gt, tg.s0, tg.W, tg.K, compact=True, tag="compact")
    eng.close()


def test_raw_layout(built, small):
    """EIMP_BATCH_RAW: data words only, rows of roundToMultiple(K,512)/32 words
    (the FPGA handler's buffer), K on both sides of a 512 boundary."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel)
    tg = synth.make_targets(panel, 7, 5, kset=[31, 64, 200, 299, 300], wmin=300, wmax=4000)
    _run_and_compare(panel, pv, eng, tg.gt, tg.s0, tg.W, tg.K, raw=True, tag="raw")
    with pytest.raises(EimpError):  # the C ABI rejects two layouts at once
        eng.run_batch(tg.gt, tg.s0, tg.W, tg.K, compact=True, raw=True)
    eng.close()


def test_ties_and_all_missing(built, small):
    """Many equal distances (all-missing / all-het targets) exercise the index tie-break."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel)
    W = 1500
    uw = unit_words(W)
    allmiss = np.zeros(2 * uw, np.uint64)
    full = np.uint64(0xFFFFFFFFFFFFFFFF)
    for i in range(W // 64):
        allmiss[2 * i] = full
        allmiss[2 * i + 1] = full
    allhet = np.zeros(2 * uw, np.uint64)
    # a target equal to panel haplotype 7 twice (homozygous everywhere)
    h = panel.hap[7]
    hom = np.zeros(2 * uw, np.uint64)
    for i in range((W + 63) // 64):
        w = int(h[i]) if i < len(h) else 0
        valid = (1 << (W - 64 * i)) - 1 if W - 64 * i < 64 else (1 << 64) - 1
        hom[2 * i] = np.uint64(~w & valid)
        hom[2 * i + 1] = np.uint64(w & valid)
    gts = [allmiss, allhet, hom, hom]
    _run_and_compare(panel, pv, eng, gts, [0, 0, 0, 0], [W] * 4, [10, 10, 5, 299], tag="ties")
    eng.close()


def test_haploid_and_ignored_sample(built, small):
    panel = small
    Nad = panel.N
    hapl = np.zeros(Nad // 2, np.uint8)
    hapl[[3, 10, 11, 77]] = 1
    ignore = [5, 20, 149, 0]
    Ncorr = Nad - int(hapl.sum()) - 2
    eng = _panel_engine(panel, Ncorr=Ncorr, haploid=hapl)
    tg = synth.make_targets(panel, 9, 4, kset=[40, 200, Ncorr, Ncorr + 5], wmin=500, wmax=4000)
    pvs = [oracle.PanelView(panel.hap, panel.site, panel.cM, haploid_ref=hapl, ignore_sample=ig,
                            N_corr=Ncorr) for ig in ignore]
    _run_and_compare(panel, lambda t: pvs[t], eng, tg.gt, tg.s0, tg.W, tg.K, ignore=ignore,
                     tag="haploid")
    eng.close()


def test_dropin_run_cref_pbwt(built, small):
    """GPUEngine::runCrefPBWTKernel: host prep (oracle) in, GPU layout out."""
    panel = small
    for K in (300, 120):
        pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
        tg = synth.make_targets(panel, 21, 5, K=K, miss=0.01)
        refs = [oracle.process_target(pv, tg.gt[t], 0, panel.M, K) for t in range(5)]
        eng = Engine(0)
        eng.initialize(K, panel.M, panel.N, panel.N, panel.site)
        offs = np.zeros(6, np.uint32)
        for t in range(5):
            offs[t + 1] = offs[t] + refs[t]["Mp"]
        # the reference copies roundToMultiple(2M,512)/8 bytes of each GenotypeVector
        say("drop-in K", K, "Mp", [r["Mp"] for r in refs])
        outs, c0s = eng.run_cref_pbwt([g for g in tg.gt], [r["splits"] for r in refs],
                                      [r["flags"] for r in refs], offs)
        for t in range(5):
            Mp, KW = refs[t]["Mp"], oracle.pbwt_row_words(K)
            gw = oracle.gpu_row_words(K)
            got = outs[t].reshape(2 * Mp, gw)
            assert np.array_equal(got[:, :KW], refs[t]["pbwt"]), f"drop-in K={K} target {t}"
            assert not got[:, KW:].any()
            assert np.array_equal(c0s[t], refs[t]["count0"])
            exp = np.zeros((2 * Mp, gw), np.uint32)
            oracle.lib().eo_pbwt_to_gpu_layout(oracle._p(refs[t]["pbwt"]), K, Mp, oracle._p(exp))
            assert np.array_equal(got, exp)
        eng.close()


def test_dropin_rejects_inconsistent_offsets(built, small):
    panel = small
    K = 100
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    tg = synth.make_targets(panel, 3, 1, K=K)
    r = oracle.process_target(pv, tg.gt[0], 0, panel.M, K)
    eng = Engine(0)
    eng.initialize(K, panel.M, panel.N, panel.N, panel.site)
    offs = np.array([0, r["Mp"] + 2], np.uint32)
    with pytest.raises(EimpError):
        eng.run_cref_pbwt([tg.gt[0]], [r["splits"]], [r["flags"]], offs)
    eng.close()


@pytest.mark.parametrize("variant", [None, "all"])
def test_pbwt_variants_agree(built, small, variant):
    """Every PBWT kernel variant that fits K gives the same digests."""
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel)
    tg = synth.make_targets(panel, 33, 3, kset=[200, 300, 64], wmin=1000, wmax=4000)
    refs = [oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t], keep=False)
            for t in range(3)]
    from eagleimp_amd.engine import pbwt_variants
    vs = [v["variant"] for v in pbwt_variants() if v["max_k"] >= 320] if variant == "all" else [-1]
    for v in vs:
        eng.set_pbwt_variant(v)
        say("variant", v)
        res = eng.run_batch(tg.gt, tg.s0, tg.W, tg.K)
        for t in range(3):
            assert int(res.dig_pbwt[t]) == refs[t]["dig_pbwt"], f"variant {v} target {t}"
            assert int(res.dig_count0[t]) == refs[t]["dig_count0"]
    eng.close()


def test_wide_panel_cref(built):
    """N > 32,768 haplotypes: whole-panel targets leave the transpose kernel
    for the haplotype-major one, subset targets take K up to 24,000."""
    panel = synth.make_panel(17, N=33500, M=700, F=40, T=6)
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = _panel_engine(panel)
    tg = synth.make_targets(panel, 17, 4, kset=[40, 3000, 20000, 24000], wmin=300, wmax=700,
                            miss=0.0, miss_hi=0.03)
    _run_and_compare(panel, pv, eng, tg.gt, tg.s0, tg.W, tg.K, tag="wide")
    st = eng.stream(4, 700, 24000, 1 << 30)
    st.submit(tg.gt, tg.s0, tg.W, tg.K)
    res, nd, _ = st.wait()
    assert nd == 0
    for t in range(4):
        r = oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t], keep=False)
        assert (int(res.Mp[t]), int(res.dig_pbwt[t]), int(res.dig_count0[t])) == \
            (r["Mp"], r["dig_pbwt"], r["dig_count0"]), f"stream wide target {t}"
    st.close()
    eng.close()


def test_subset_cref_edge_windows(built, small):
    """Subset targets through the haplotype-major cref kernel, windows chosen
    for its edge cases: no split at all (S = 0), every S mod 8 (S % 8 == 0
    leaves the last 16-row block holding only the final segment), windows
    ending at the panel's last site, unaligned starts and all-missing
    stretches (is0 and is2 both set: no inconsistency); synchronous engine
    and stream context."""
    from phase_cases import encode_gt
    panel = small
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    M = panel.M
    bits = lambda w: np.unpackbits(w.view(np.uint8), bitorder="little")[:M]  # noqa: E731
    codes = (bits(panel.tgt_hap[0]).astype(np.uint8) + bits(panel.tgt_hap[1])).astype(np.uint8)
    # 0 HomRef, 1 Het, 2 HomAlt
    rs = np.random.default_rng(5)
    cases = [(100, 50, 150, "nohet"), (3940, 60, 150, None), (3999, 1, 64, None),
             (M - 333, 333, 200, "miss"), (5, 700, 299, "miss")]
    cases += [(17 + 7 * i, 40 + 3 * i, 100 + i, None) for i in range(30)]
    gts, s0s, Ws, Ks = [], [], [], []
    for s0, W, K, mod in cases:
        c = codes[s0:s0 + W].copy()
        if mod == "nohet":
            c[c == 1] = 0
        elif mod == "miss":
            a = int(rs.integers(0, max(1, W - 40)))
            c[a:a + 40] = 3
        gts.append(encode_gt(c)); s0s.append(s0); Ws.append(W); Ks.append(K)
    eng = _panel_engine(panel)
    res = _run_and_compare(panel, pv, eng, gts, s0s, Ws, Ks, tag="cref edges")
    S = [int(x) for x in res.S]
    assert 0 in S and {s % 8 for s in S if s} == set(range(8)), f"edge classes not reached: {S}"
    st = eng.stream(len(gts), max(Ws), max(Ks), 1 << 30)
    st.submit(gts, s0s, Ws, Ks)
    sres, nd, _ = st.wait()
    assert nd == 0
    for t in range(len(gts)):
        assert (int(sres.Mp[t]), int(sres.dig_pbwt[t]), int(sres.dig_count0[t])) == \
            (int(res.Mp[t]), int(res.dig_pbwt[t]), int(res.dig_count0[t])), f"stream target {t}"
    st.close()
    eng.close()
