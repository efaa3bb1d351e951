"""Pins of the oracle where the reference has no runnable code here.

Target.cpp (findBestHaps, findMonos, findCallLocs, condensedReferenceT)
needs TBB/HTSlib headers this image lacks, so those functions are checked
against brute force written from the definitions (not from the oracle's
code), and findCallLocs against known-answer cases written out by hand.  The
PBWT is also checked against Durbin's formulation (order = stable sort by the
reversed prefix) and the interval mapping against its closed form.
"""
import numpy as np
import pytest

import oracle
from eagleimp_amd import synth


def bits(words, n):
    return np.unpackbits(np.ascontiguousarray(words).view(np.uint8), bitorder="little")[:n]


def gt_codes(gt, W):
    is0 = bits(gt[0::2], W)
    is2 = bits(gt[1::2], W)
    return np.where(is0 & is2, 3, np.where(is0, 0, np.where(is2, 2, 1))).astype(np.uint8)


def words_of(bitarr, nwords):
    padded = np.zeros(nwords * 64, np.uint8)
    padded[:len(bitarr)] = bitarr
    return np.packbits(padded, bitorder="little").view(np.uint64)


@pytest.fixture(scope="module")
def data(built):
    p = synth.make_panel(17, N=160, M=1800, F=10, T=8)
    tg = synth.make_targets(p, 17, 8, kset=[2, 7, 40, 100, 159, 160], wmin=100, wmax=1800,
                            miss=0.0, miss_hi=0.08)
    return p, tg, oracle.PanelView(p.hap, p.site, p.cM)


def test_best_haps_brute_force(data):
    p, tg, pv = data
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)
    for t in range(len(tg)):
        s0, W, K = int(tg.s0[t]), int(tg.W[t]), min(int(tg.K[t]), p.N)
        codes = gt_codes(tg.gt[t], W)
        h = H[:, s0:s0 + W]
        # error = hap differs from a homozygous genotype, or genotype missing
        err = ((codes == 0) & (h == 1)) | ((codes == 2) & (h == 0)) | (codes == 3)
        nerr = err.sum(axis=1)
        order = np.lexsort((np.arange(p.N), nerr))
        exp = np.sort(order[:K]) if K < p.N else np.arange(p.N)
        r = oracle.process_target(pv, tg.gt[t], s0, W, K)
        assert np.array_equal(r["besthaps"], exp), t
        assert np.array_equal(bits(r["flags"], p.N), np.isin(np.arange(p.N), exp).astype(np.uint8))


def test_best_haps_haploid_and_ignored_brute_force(data):
    """findBestHaps with haploid reference samples and an ignored sample
    (src/Target.cpp:791-853): the candidates are every haplotype except both of
    the ignored sample and the second of each haploid sample; K >= that count
    takes all of them in index order, otherwise the K smallest (nerr, index)
    pairs, reported by index.  Written from that rule, not from the oracle;
    all-missing targets make every distance equal, so the index tie-break
    decides among the candidates alone."""
    p, tg, _ = data
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)
    rng = np.random.default_rng(2024)
    for trial in range(6):
        hapl = (rng.random(p.N // 2) < (0.0 if trial == 0 else 0.25)).astype(np.uint8)
        ignore = -1 if trial == 1 else int(rng.integers(0, p.N // 2))
        if trial == 2:
            hapl[ignore] = 1  # an ignored sample that is itself haploid
        allowed = np.ones(p.N, bool)
        allowed[1::2][hapl == 1] = False
        if ignore >= 0:
            allowed[2 * ignore:2 * ignore + 2] = False
        cand = np.nonzero(allowed)[0]
        ncorr = len(cand)
        pv = oracle.PanelView(p.hap, p.site, p.cM, haploid_ref=hapl, ignore_sample=ignore)
        assert pv.N_corr == ncorr
        for t in range(len(tg)):
            s0, W = int(tg.s0[t]), int(tg.W[t])
            gt = tg.gt[t]
            if t == 0:  # all missing: every candidate has nerr = W
                gt = np.zeros_like(gt)
                for i in range((W + 63) // 64):
                    v = (1 << min(64, W - 64 * i)) - 1
                    gt[2 * i] = gt[2 * i + 1] = np.uint64(v)
            codes = gt_codes(gt, W)
            h = H[:, s0:s0 + W]
            nerr = (((codes == 0) & (h == 1)) | ((codes == 2) & (h == 0)) | (codes == 3)).sum(axis=1)
            for K in (1, 5, ncorr // 2, ncorr - 1, ncorr, ncorr + 3):
                Kc = min(K, ncorr)
                if Kc >= ncorr:
                    exp = cand
                else:
                    order = np.lexsort((cand, nerr[cand]))
                    exp = np.sort(cand[order[:Kc]])
                r = oracle.process_target(pv, gt, s0, W, K)
                what = f"trial {trial} target {t} K {K}"
                assert np.array_equal(r["besthaps"], exp), what
                assert np.array_equal(bits(r["flags"], p.N),
                                      np.isin(np.arange(p.N), exp).astype(np.uint8)), what


def test_monos_and_cref_brute_force(data):
    p, tg, pv = data
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)
    for t in range(len(tg)):
        s0, W, K = int(tg.s0[t]), int(tg.W[t]), min(int(tg.K[t]), p.N)
        r = oracle.process_target(pv, tg.gt[t], s0, W, K)
        sel = H[r["besthaps"]][:, s0:s0 + W]
        mono = (sel == sel[0]).all(axis=0)
        assert np.array_equal(bits(r["monos"], W), mono.astype(np.uint8)), t
        codes = gt_codes(tg.gt[t], W)
        ss = list(r["splitsites"])
        S = len(ss)
        rows = []
        bounds = [-1] + ss + [W]
        for s in range(S + 1):
            seg = np.arange(bounds[s] + 1, bounds[s + 1])
            inc = np.zeros(K, bool)
            for m in seg:
                if codes[m] == 0:
                    inc |= sel[:, m] == 1
                elif codes[m] == 2:
                    inc |= sel[:, m] == 0
            rows.append(inc)
            if s < S:
                rows.append(sel[:, ss[s]] == 1)
        rows = np.array(rows, np.uint8)
        got = np.unpackbits(r["cref"].view(np.uint8), bitorder="little").reshape(r["Mp"], -1)[:, :K]
        assert np.array_equal(got, rows), t
        assert np.array_equal(r["count0"], K - rows.sum(axis=1)), t
        assert not np.unpackbits(r["cref"].view(np.uint8), bitorder="little").reshape(
            r["Mp"], -1)[:, K:].any()


def _durbin_pbwt(x):
    """x: (Mp, K) bits.  Returns the logical rows, forward then backward."""
    Mp, K = x.shape
    KW = oracle.pbwt_row_words(K)
    out = np.zeros((2 * Mp, KW), np.uint32)
    for d, order in enumerate((range(Mp), range(Mp - 1, -1, -1))):
        order = list(order)
        for i, r in enumerate(order):
            prev = order[:i]
            # stable sort of the indices by (x_prev[-1], ..., x_prev[0])
            keys = [np.arange(K)] + [x[q] for q in prev]
            perm = np.lexsort(keys) if prev else np.arange(K)
            y = x[r][perm]
            ypad = np.ones(((K + 31) // 32) * 32, np.uint8)
            ypad[:K] = y
            data = np.packbits(ypad, bitorder="little").view(np.uint32)
            zeros = np.cumsum((1 - ypad).reshape(-1, 32).sum(axis=1))
            row = np.empty(KW, np.uint32)
            row[0::2] = data
            row[1::2] = zeros
            out[d * Mp + i] = row
    return out


@pytest.mark.parametrize("K,Mp,dens", [(1, 3, 0.5), (5, 6, 0.5), (32, 9, 0.3), (33, 8, 0.6),
                                       (70, 12, 0.5), (130, 10, 0.1)])
def test_pbwt_durbin(built, K, Mp, dens):
    rng = np.random.default_rng(K * 100 + Mp)
    x = (rng.random((Mp, K)) < dens).astype(np.uint8)
    kw = synth.unit_words(K)
    cref = np.stack([words_of(row, kw) for row in x])
    out = oracle.build_pbwt(cref, K, (K - x.sum(axis=1)).astype(np.int32))
    assert np.array_equal(out, _durbin_pbwt(x))


def test_map_interval_closed_form(built):
    rng = np.random.default_rng(5)
    K, Mp = 77, 9
    x = (rng.random((Mp, K)) < 0.45).astype(np.uint8)
    kw = synth.unit_words(K)
    cref = np.stack([words_of(row, kw) for row in x])
    c0 = (K - x.sum(axis=1)).astype(np.int32)
    pb = oracle.build_pbwt(cref, K, c0)
    KW = oracle.pbwt_row_words(K)
    for rev in (0, 1):
        for m in range(Mp):
            row = pb[(2 * Mp - 1 - m) if rev else m]
            y = bits(row[0::2].copy(), K)
            cnt0 = int(c0[m])
            for l in range(K):
                for r in range(l, K, 7):
                    z_lt = int((y[:l] == 0).sum())
                    z_le = int((y[:r + 1] == 0).sum())
                    if l == r:
                        exp = (z_lt, z_lt, 0, -1) if y[l] == 0 else (0, -1, cnt0 + l - z_lt,
                                                                      cnt0 + l - z_lt)
                    else:
                        exp = (z_lt, z_le - 1, cnt0 + l - z_lt, cnt0 + (r + 1 - z_le) - 1)
                    assert oracle.map_interval(pb, K, c0, rev, m, l, r) == exp
    assert pb.shape[1] == KW


def _call_locs(codes, cM, monos=(), thr=0.125):
    W = len(codes)
    uw = synth.unit_words(W)
    gt = np.zeros(2 * uw, np.uint64)
    c = np.asarray(codes, np.uint8)
    gt[0::2] = words_of(((c == 0) | (c == 3)).astype(np.uint8), uw)
    gt[1::2] = words_of(((c == 2) | (c == 3)).astype(np.uint8), uw)
    mono = np.zeros(W, np.uint8)
    mono[list(monos)] = 1
    mw = words_of(mono, uw)
    cMf = np.asarray(cM, np.float32)
    dummy = np.zeros((1, 8), np.uint64)
    pv = oracle.PanelView(dummy, dummy, cMf, N_ad=2, N_corr=2, M=W)
    t = oracle.make_target(gt, 0, W, 2, thr)
    splits = np.zeros(uw, np.uint64)
    ss = np.zeros(W + 1, np.uint32)
    cl = np.zeros(W + 1, np.uint32)
    nc = np.zeros(1, np.uint64)
    diffs = np.zeros(W + 1, np.float32)
    import ctypes as C
    S = oracle.lib().eo_find_call_locs(C.byref(pv.c), C.byref(t), oracle._p(mw), oracle._p(splits),
                                       oracle._p(ss), oracle._p(cl), oracle._p(nc), oracle._p(diffs))
    return list(ss[:S]), bits(splits, W).nonzero()[0].tolist(), diffs[:S + 1 if S else 0]


def f32(*v):
    return np.float32(v[0]) if len(v) == 1 else [np.float32(x) for x in v]


def test_call_locs_kat_basic():
    cM = [0.0, 0.05, 0.1, 0.2, 0.3, 0.35, 0.4, 0.5, 0.6, 0.7]
    ss, bm, d = _call_locs([0, 0, 1, 0, 0, 2, 1, 0, 0, 0], cM)
    assert ss == [2, 4, 6] and bm == [2, 4, 6]        # 8 split then truncated (trailing hom)
    c = np.asarray(cM, np.float32)
    assert np.array_equal(d.view(np.uint32), np.array([c[2] - c[0], c[4] - c[2], c[6] - c[4],
                                                       c[9] - c[6]], np.float32).view(np.uint32))


def test_call_locs_kat_mono_first_het_keeps_initial_cm():
    # the first het (mono) is no split; the last split cM stays cM[0] = 0, so
    # site 2 (0.2 cM away from cM[0], only 0.1 from the first het) is a split
    ss, _, _ = _call_locs([0, 1, 0, 0, 1], [0.0, 0.1, 0.2, 0.21, 0.5], monos=[1])
    assert ss == [2, 4]


def test_call_locs_kat_missing_never_split():
    ss, _, _ = _call_locs([1, 3, 3, 1], [0.0, 1.0, 2.0, 3.0])
    assert ss == [0, 3]


def test_call_locs_kat_no_het_no_call():
    assert _call_locs([0, 2, 0, 2], [0.0, 1.0, 2.0, 3.0])[0] == []
    # only hom splits behind a mono first het: no call -> no split at all
    ss, bm, d = _call_locs([1, 0, 0, 0], [0.0, 1.0, 2.0, 3.0], monos=[0])
    assert ss == [] and bm == [] and len(d) == 0


def test_call_locs_kat_threshold_is_strict():
    ss, _, _ = _call_locs([1, 0, 0, 1], [0.0, 0.125, 0.25, 0.375])
    assert ss == [0, 2, 3]


def test_call_locs_kat_mono_het_inside():
    ss, _, _ = _call_locs([1, 1, 1], [0.0, 0.0, 0.0], monos=[1])
    assert ss == [0, 2]


def test_call_locs_kat_long_runs_cross_words():
    W = 200
    codes = [0] * W
    codes[70] = 1
    codes[190] = 1
    cM = (np.arange(W) * np.float32(0.1)).astype(np.float32)
    ss, bm, _ = _call_locs(codes, cM)
    assert ss == list(range(70, 191, 2)) and bm == ss


def test_window_equals_sliced_panel(data):
    """A window is the same algorithm on the column slice (chunk semantics)."""
    p, tg, pv = data
    t = 3
    s0, W, K = int(tg.s0[t]), int(tg.W[t]), min(int(tg.K[t]), p.N)
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)[:, s0:s0 + W]
    hw = synth.unit_words(W)
    hap2 = np.stack([words_of(h, hw) for h in H])
    site2 = p.site[s0:s0 + W].copy()
    pv2 = oracle.PanelView(hap2, site2, p.cM[s0:s0 + W])
    a = oracle.process_target(pv, tg.gt[t], s0, W, K)
    b = oracle.process_target(pv2, tg.gt[t], 0, W, K)
    for k in ("besthaps", "splitsites", "count0", "pbwt", "cref"):
        assert np.array_equal(a[k], b[k]), k


def test_gpu_layout_conversion(data):
    p, tg, pv = data
    r = oracle.process_target(pv, tg.gt[0], tg.s0[0], tg.W[0], tg.K[0])
    K, Mp = int(r["K"]), int(r["Mp"])
    g = np.zeros((2 * Mp, oracle.gpu_row_words(K)), np.uint32)
    oracle.lib().eo_pbwt_to_gpu_layout(oracle._p(r["pbwt"]), K, Mp, oracle._p(g))
    KW = oracle.pbwt_row_words(K)
    assert np.array_equal(g[:, :KW], r["pbwt"]) and not g[:, KW:].any()
    assert oracle.digest(g, K, r["count0"]) == (r["dig_pbwt"], r["dig_count0"])


def test_digest_detects_single_word_change(data):
    p, tg, pv = data
    r = oracle.process_target(pv, tg.gt[1], tg.s0[1], tg.W[1], tg.K[1])
    K = int(r["K"])
    base = oracle.digest(r["pbwt"], K, r["count0"])
    pb = r["pbwt"].copy()
    pb[pb.shape[0] // 2, 1] ^= np.uint32(1)
    assert oracle.digest(pb, K, r["count0"])[0] != base[0]
    pb = r["pbwt"].copy()
    pb[[0, 1]] = pb[[1, 0]]  # swapped rows
    if not np.array_equal(pb, r["pbwt"]):
        assert oracle.digest(pb, K, r["count0"])[0] != base[0]


def test_digest_formula():
    """oracle.digest against a numpy evaluation of its definition: sum over
    logical words of rotl32(v ^ g*0x9E3779B1, g mod 32), and over count0 of
    fmix32(c0 + r*0x7FEB352D + 0x68E31DA4)."""
    rng = np.random.default_rng(7)
    K, Mp = 77, 9
    KW = oracle.pbwt_row_words(K)
    rows = rng.integers(0, 2**32, size=(2 * Mp, KW + 3), dtype=np.uint64).astype(np.uint32)
    c0 = rng.integers(0, K + 1, size=Mp).astype(np.int32)
    g = np.arange(2 * Mp * KW, dtype=np.uint64)
    v = rows[:, :KW].reshape(-1).astype(np.uint64)
    x = (v ^ ((g * 0x9E3779B1) & 0xFFFFFFFF)) & 0xFFFFFFFF
    r = g & 31
    t = ((x << r) | (x >> ((32 - r) & 63))) & 0xFFFFFFFF
    t = np.where(r == 0, x, t)

    def fmix32(h):
        h = h & 0xFFFFFFFF
        h ^= h >> 16
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & 0xFFFFFFFF
        h ^= h >> 16
        return h

    rr = np.arange(Mp, dtype=np.uint64)
    c = fmix32((c0.astype(np.uint64) & 0xFFFFFFFF) + ((rr * 0x7FEB352D) & 0xFFFFFFFF) + 0x68E31DA4)
    want = (int(t.sum()) % 2**64, int(c.sum()) % 2**64)
    assert oracle.digest(rows, K, c0) == want
