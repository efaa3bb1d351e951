"""The oracle against golden vectors written by the reference's own code.

tests/golden/ref_pbwt.npz and ref_genotype.npz come from the reference's
PBWT.cpp / Datatypes.cpp (compiled in place by oracle/Makefile, driven by
tests/golden/make_golden.py).  Where the reference sources are present, the
same comparisons also run live on fresh random inputs.
"""
import os

import numpy as np
import pytest

import oracle
from eagleimp_amd import synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def encode_genotypes(codes):
    """GenotypeVector words for genotype codes 0 HomRef, 1 Het, 2 HomAlt, 3 Miss:
    an is0 word (HomRef or Miss) and an is2 word (HomAlt or Miss) per 64 sites,
    grown in 512-site units."""
    codes = np.asarray(codes, np.uint8)
    W = len(codes)
    units = (W + 511) // 512
    is0 = np.zeros(units * 512, np.uint8)
    is2 = np.zeros(units * 512, np.uint8)
    is0[:W] = (codes == 0) | (codes == 3)
    is2[:W] = (codes == 2) | (codes == 3)
    w0 = np.packbits(is0, bitorder="little").view(np.uint64)
    w2 = np.packbits(is2, bitorder="little").view(np.uint64)
    out = np.empty(2 * len(w0), np.uint64)
    out[0::2] = w0
    out[1::2] = w2
    return out


def test_pbwt_golden(built):
    g = np.load(os.path.join(GOLD, "ref_pbwt.npz"))
    for i in range(int(g["ncases"])):
        K = int(g[f"c{i}_K"])
        out = oracle.build_pbwt(g[f"c{i}_cref"], K, g[f"c{i}_count0"])
        assert np.array_equal(out, g[f"c{i}_pbwt"]), f"case {i} K={K}"


def test_map_interval_golden(built):
    g = np.load(os.path.join(GOLD, "ref_pbwt.npz"))
    n = 0
    for i in range(int(g["ncases"])):
        K = int(g[f"c{i}_K"])
        pbwt, c0 = g[f"c{i}_pbwt"], g[f"c{i}_count0"]
        for q, exp in zip(g[f"c{i}_queries"], g[f"c{i}_qres"]):
            got = oracle.map_interval(pbwt, K, c0, int(q[0]), int(q[1]), int(q[2]), int(q[3]))
            assert got == tuple(int(x) for x in exp), f"case {i} query {tuple(q)}"
            n += 1
    assert n > 200


def test_genotype_encoding_and_distance_golden(built):
    g = np.load(os.path.join(GOLD, "ref_genotype.npz"))
    for i in range(int(g["ncases"])):
        codes, hap = g[f"g{i}_codes"], g[f"g{i}_hap"]
        words = g[f"g{i}_words"]
        assert np.array_equal(encode_genotypes(codes), words), f"case {i}: encoding"
        nerr = oracle.compare_with_preinit(words, hap, hap.shape[0])
        assert nerr == int(g[f"g{i}_nerr"]), f"case {i}: compare_withPreInit"


def test_generator_encoding_matches_reference_encoding(built):
    """synth_target_gt writes what GenotypeVector::push_back would."""
    p = synth.make_panel(3, N=20, M=1300, F=4, T=2)
    for (s0, W) in [(0, 1300), (5, 700), (64, 513), (1299, 1)]:
        gt = np.zeros(2 * synth.unit_words(W), np.uint64)
        synth._lib().synth_target_gt(synth.ptr(p.tgt_hap[0]), synth.ptr(p.tgt_hap[1]), s0, W, 0.0,
                                     3, 0, synth.ptr(gt))
        a = np.unpackbits(p.tgt_hap[0].view(np.uint8), bitorder="little")[s0:s0 + W]
        b = np.unpackbits(p.tgt_hap[1].view(np.uint8), bitorder="little")[s0:s0 + W]
        codes = (a.astype(np.uint8) + b).astype(np.uint8)
        assert np.array_equal(gt, encode_genotypes(codes)), (s0, W)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_pbwt_live_against_reference(built):
    rng = np.random.default_rng(1234)
    for K, Mp in [(2, 5), (47, 31), (128, 17), (300, 63), (1531, 9)]:
        bits = (rng.random((Mp, K)) < rng.random()).astype(np.uint8)
        kw = synth.unit_words(K)
        padded = np.zeros((Mp, kw * 64), np.uint8)
        padded[:, :K] = bits
        cref = np.packbits(padded, axis=1, bitorder="little").view(np.uint64)
        c0 = (K - bits.sum(axis=1)).astype(np.int32)
        ref, _ = oracle.ref_build_pbwt(cref, K, c0)
        assert np.array_equal(oracle.build_pbwt(cref, K, c0), ref), (K, Mp)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_oracle_pipeline_pbwt_against_reference(built):
    """Condensed references produced by the oracle on synthetic targets, fed to
    both PBWT builders."""
    p = synth.make_panel(4, N=200, M=2500, F=12, T=3)
    tg = synth.make_targets(p, 4, 3, kset=[64, 150, 200], wmin=800, wmax=2500, miss=0.02)
    pv = oracle.PanelView(p.hap, p.site, p.cM)
    for t in range(3):
        r = oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t])
        ref, _ = oracle.ref_build_pbwt(r["cref"], int(r["K"]), r["count0"])
        assert np.array_equal(r["pbwt"], ref)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_raw_layout_ingested_by_reference(built):
    """The raw layout (data words only, rows of roundToMultiple(K,512)/32
    words, EIMP_BATCH_RAW) fed to the reference's own raw-buffer constructor
    (src/PBWT.cpp:72-110) gives back the CPU-layout rows and count0."""
    rng = np.random.default_rng(99)
    for K, Mp in [(2, 5), (47, 31), (512, 7), (513, 4), (1531, 9)]:
        bits = (rng.random((Mp, K)) < rng.random()).astype(np.uint8)
        kw = synth.unit_words(K)
        padded = np.zeros((Mp, kw * 64), np.uint8)
        padded[:, :K] = bits
        cref = np.packbits(padded, axis=1, bitorder="little").view(np.uint64)
        c0 = (K - bits.sum(axis=1)).astype(np.int32)
        rows = oracle.build_pbwt(cref, K, c0)
        got, gc0 = oracle.ref_ingest_raw(oracle.raw_rows(rows, K), K)
        nw = (K + 31) // 32
        assert np.array_equal(got[:, :2 * nw], rows), (K, Mp)
        assert np.array_equal(gc0, c0), (K, Mp)


def test_panel_pbwt_golden(built):
    """The oracle's panel PBWT for imputation against the reference's
    PBWT(refT, true) outputs: rows, count0 and every absolute permutation."""
    g = np.load(os.path.join(GOLD, "ref_panel_pbwt.npz"))
    for i in range(int(g["ncases"])):
        K = int(g[f"p{i}_K"])
        rows, c0, ap = oracle.build_ref_pbwt(g[f"p{i}_rows"], K)
        assert np.array_equal(rows, g[f"p{i}_pbwt"]), (i, K)
        assert np.array_equal(c0, g[f"p{i}_count0"]), (i, K)
        assert np.array_equal(ap, g[f"p{i}_absperm"]), (i, K)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built here")
def test_panel_pbwt_live_against_reference(built):
    p = synth.make_panel(5, N=333, M=700, F=10, T=2)
    a = oracle.build_ref_pbwt(p.site, p.N)
    b = oracle.ref_build_ref_pbwt(p.site, p.N)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
