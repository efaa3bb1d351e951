"""Phasing test cases: synthetic targets made hard for the beam search
(genotype noise, runs of homozygosity that create many homozygous split
sites, missing genotypes, panels with few founders so that many haplotypes
are identical and probabilities tie), with the engine's inputs computed by
the oracle."""
from __future__ import annotations

import hashlib

import numpy as np

import oracle
from eagleimp_amd import phase, synth


def encode_gt(codes: np.ndarray) -> np.ndarray:
    """Genotype codes (0 HomRef, 1 Het, 2 HomAlt, 3 Miss) -> GenotypeVector
    words: (is0, is2) word pairs (src/Datatypes.cpp:27-56)."""
    codes = np.asarray(codes, np.uint8)
    W = codes.shape[0]
    uw = synth.unit_words(W)

    def pack(bits):
        b = np.zeros(uw * 64, np.uint8)
        b[:W] = bits
        return np.packbits(b, bitorder="little").view(np.uint64)

    gt = np.zeros(2 * uw, np.uint64)
    gt[0::2] = pack((codes == 0) | (codes == 3))
    gt[1::2] = pack((codes == 2) | (codes == 3))
    return gt


def decode_gt(gt: np.ndarray, W: int) -> np.ndarray:
    gt = np.asarray(gt, np.uint64)
    is0 = np.unpackbits(np.ascontiguousarray(gt[0::2]).view(np.uint8), bitorder="little")[:W]
    is2 = np.unpackbits(np.ascontiguousarray(gt[1::2]).view(np.uint8), bitorder="little")[:W]
    codes = np.full(W, 1, np.uint8)
    codes[(is0 == 1) & (is2 == 0)] = 0
    codes[(is0 == 0) & (is2 == 1)] = 2
    codes[(is0 == 1) & (is2 == 1)] = 3
    return codes


# name: panel (N, M, founders), K, cMmaxSplit, noise, runs of homozygosity
# (count, min, max length), missing rate, phasing parameter overrides, start phase
CASES = {
    "default": dict(N=400, M=2000, F=16, K=200, split=0.125, noise=0.03, roh=(0, 0, 0),
                    miss=0.01, params={}, start=0),
    "few_founders_full_panel": dict(N=300, M=1500, F=3, K=300, split=0.125, noise=0.05,
                                    roh=(0, 0, 0), miss=0.01, params={}, start=1),
    "tiny_beam_small_K": dict(N=200, M=1500, F=6, K=4, split=0.125, noise=0.08, roh=(0, 0, 0),
                              miss=0.02,
                              params=dict(pre_phasing=1, beam_width=4, beam_width_fast=3,
                                          p_limit=0.3), start=0),
    "roh_hom_splits_prephase": dict(N=300, M=3000, F=8, K=120, split=0.005, noise=0.02,
                                    roh=(6, 100, 300), miss=0.01,
                                    params=dict(pre_phasing=1), start=0),
    "missing_imputed_no_reverse": dict(N=250, M=1500, F=10, K=100, split=0.125, noise=0.03,
                                       roh=(0, 0, 0), miss=0.12,
                                       params=dict(imp_missing=1, rev_phasing=0), start=1),
    "short_history": dict(N=300, M=2000, F=12, K=150, split=0.02, noise=0.05, roh=(3, 80, 160),
                          miss=0.01,
                          params=dict(hist=0.5, delta=5, delta_fast=3, pre_phasing=1), start=0),
    "noisy_wide_beam": dict(N=400, M=1200, F=30, K=400, split=0.125, noise=0.25, roh=(0, 0, 0),
                            miss=0.0, params=dict(beam_width=120, pre_phasing=1), start=0),
    # calls up to ~150 split sites apart: the beam reads its haplotype
    # histories past their 100 (and past 128) recorded sites
    "long_roh_deep_history": dict(N=300, M=4000, F=8, K=120, split=0.005, noise=0.02,
                                  roh=(1, 700, 700), miss=0.01, params={}, start=0),
}

# a run of homozygosity so long that a call looks further back than the
# reference's history words reach (rejected)
TOO_DEEP = dict(N=300, M=4000, F=8, K=120, split=0.005, noise=0.02, roh=(1, 1100, 1100),
                miss=0.01, params={}, start=0)


def make_case(name: str, seed: int = 0, spec=None):
    """Returns (panel view, codes, gt words, oracle result, phasing overrides,
    allele frequencies, start phase)."""
    c = CASES[name] if spec is None else spec
    rs = np.random.default_rng(1000 + seed * 31 + len(name))
    panel = synth.make_panel(7 + seed, N=c["N"], M=c["M"], F=c["F"], T=2)
    tg = synth.make_targets(panel, 7 + seed, 1, K=c["K"], miss=0.0)
    W = c["M"]
    codes = decode_gt(tg.gt[0], W)
    h1 = np.unpackbits(panel.tgt_hap[0].view(np.uint8), bitorder="little")[:W]
    flip = rs.random(W) < c["noise"]
    codes[flip] = rs.integers(0, 3, int(flip.sum()))
    miss = rs.random(W) < c["miss"]
    codes[miss] = 3
    n_roh, lo, hi = c["roh"]
    for _ in range(n_roh):  # no heterozygous site inside a run
        L = int(rs.integers(lo, hi + 1))
        a = int(rs.integers(0, max(1, W - L)))
        codes[a:a + L] = np.where(h1[a:a + L] == 1, 2, 0)
    gt = encode_gt(codes)
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    r = oracle.process_target(pv, gt, 0, W, c["K"], cMmaxSplit=c["split"])
    af = np.clip(rs.random(W).astype(np.float32), 0.01, 0.99).astype(np.float32)
    return dict(panel=panel, pv=pv, codes=codes, gt=gt, W=W, r=r, params=c["params"], af=af,
                start=c["start"], split=c["split"])


def params_for(c) -> phase.PhaseParams:
    return phase.default_params(**c["params"])


def job_for(c, pbwt=None, count0=None):
    """Phasing job over the oracle's compact cPBWT (or the given one)."""
    r = c["r"]
    return phase.job(r["pbwt"] if pbwt is None else pbwt,
                     r["count0"] if count0 is None else count0, r["K"], c["gt"], c["W"],
                     r["splitsites"], r["cMdiffs"], af=c["af"], start_phase=c["start"])


def input_hash(c) -> str:
    """SHA-256 of a case's phasing inputs (detects drift of the generators)."""
    r = c["r"]
    h = hashlib.sha256()
    for a in (r["pbwt"], r["count0"], c["gt"], r["splitsites"], r["cMdiffs"], c["af"]):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(repr((int(r["K"]), int(c["W"]), int(c["start"]), sorted(c["params"].items()))).encode())
    return h.hexdigest()
