"""Writes the golden vectors in this folder from the reference's own code.

Runs in the development container only (it needs /root/reference to build
oracle/_ref/libeimp_ref.so, the reference's PBWT.cpp, Datatypes.cpp and
RingBuffer.cpp compiled in place with oracle/ref_harness.cpp).  The fixtures
it writes are plain data (inputs and the reference's outputs) in .npz files:

  ref_pbwt.npz      condensed rows + count0 -> the reference's compressed
                    forward/backward PBWT (src/PBWT.cpp:207-285) and
                    mapInterval results (src/PBWT.cpp:288-382)
  ref_genotype.npz  genotype codes -> the reference GenotypeVector words
                    (src/Datatypes.cpp:27-56) and compare_withPreInit error
                    counts (src/Datatypes.cpp:107-116)
  ref_panel_pbwt.npz  site rows -> the reference's panel PBWT for imputation,
                    PBWT(refT, true): forward rows, count0 and the absolute
                    permutations (src/PBWT.cpp:53-71,131-153,207-285)
  ref_rle.npz       haplotype bit vectors -> the reference's Qref run-length
                    code (src/utils.h:133-190) or its refusal (code not
                    shorter than the raw vector)
  ref_rle_masked.npz  run-length coded haplotype records, haploid flags and
                    masked samples -> the reference's decoder with sample
                    masking (src/utils.h:193-315): the kept haplotypes and
                    the count of the masked samples' variant alleles
  ref_phase.npz     the reference's beam search (src/Beam.cpp,
                    src/BeamPath.cpp, src/HaplotypePath.cpp) driven as
                    Target::phaseExtPBWT (src/Target.cpp:291-777) on the
                    phasing cases of tests/phase_cases.py: phased haplotypes,
                    dosages, call confidences and totals, with a hash of the
                    inputs (regenerated from their seeds by the test)

Usage:  python tests/golden/make_golden.py [fixture ...]   (default: all;
        a fixture is named by its file without "ref_" and ".npz")
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import oracle  # noqa: E402
import phase_cases  # noqa: E402
from eagleimp_amd import phase  # noqa: E402

PBWT_CASES = [  # (K, Mp, density, seed)
    (2, 1, 0.5, 1), (2, 9, 0.5, 2), (31, 7, 0.3, 3), (32, 33, 0.5, 4), (33, 33, 0.5, 5),
    (63, 17, 0.2, 6), (64, 64, 0.5, 7), (65, 64, 0.7, 8), (100, 41, 0.5, 9),
    (1000, 64, 0.5, 10), (513, 25, 0.05, 11), (777, 31, 0.95, 12), (96, 1, 0.5, 13),
]


def random_cref(K, Mp, dens, rng):
    bits = (rng.random((Mp, K)) < dens).astype(np.uint8)
    # a few structured rows: all zero, all one, alternating
    if Mp > 3:
        bits[1] = 0
        bits[2] = 1
        bits[3] = np.arange(K) % 2
    kw = ((K + 511) // 512) * 8
    padded = np.zeros((Mp, kw * 64), np.uint8)
    padded[:, :K] = bits
    cref = np.packbits(padded, axis=1, bitorder="little").view(np.uint64)
    count0 = (K - bits.sum(axis=1)).astype(np.int32)
    return cref, count0


def pbwt_fixtures():
    out = {}
    for i, (K, Mp, dens, seed) in enumerate(PBWT_CASES):
        rng = np.random.default_rng(seed)
        cref, count0 = random_cref(K, Mp, dens, rng)
        qs = []
        for _ in range(24):
            rev = int(rng.integers(0, 2))
            m = int(rng.integers(0, Mp))
            a, b = sorted(int(x) for x in rng.integers(0, K, size=2))
            if rng.random() < 0.25:
                b = a  # single element
            if rng.random() < 0.1:
                a, b = b + 1, b  # empty
            qs.append((rev, m, a, b))
        qs = np.asarray(qs, np.int32)
        pbwt, qres = oracle.ref_build_pbwt(cref, K, count0, qs)
        out[f"c{i}_K"] = np.int64(K)
        out[f"c{i}_cref"] = cref
        out[f"c{i}_count0"] = count0
        out[f"c{i}_pbwt"] = pbwt
        out[f"c{i}_queries"] = qs
        out[f"c{i}_qres"] = qres
    out["ncases"] = np.int64(len(PBWT_CASES))
    np.savez_compressed(os.path.join(HERE, "ref_pbwt.npz"), **out)


def genotype_fixtures():
    out = {}
    Ws = [1, 63, 64, 65, 511, 512, 513, 1000, 1537]
    for i, W in enumerate(Ws):
        rng = np.random.default_rng(100 + i)
        codes = rng.integers(0, 4, size=W).astype(np.uint8)
        if W > 10:
            codes[:5] = [0, 1, 2, 3, 3]
        hbits = (rng.random(W) < 0.5).astype(np.uint8)
        hw = ((W + 511) // 512) * 8
        padded = np.zeros(hw * 64, np.uint8)
        padded[:W] = hbits
        hap = np.packbits(padded, bitorder="little").view(np.uint64)
        words, nerr = oracle.ref_genotype_compare(codes, hap, W)
        out[f"g{i}_codes"] = codes
        out[f"g{i}_hap"] = hap
        out[f"g{i}_words"] = words
        out[f"g{i}_nerr"] = np.int64(nerr)
    out["ncases"] = np.int64(len(Ws))
    np.savez_compressed(os.path.join(HERE, "ref_genotype.npz"), **out)


PANEL_CASES = [  # (K, M, density, seed)
    (1, 3, 0.5, 21), (31, 20, 0.3, 22), (64, 33, 0.5, 23), (77, 60, 0.5, 24), (300, 17, 0.1, 25),
]


def panel_pbwt_fixtures():
    out = {}
    for i, (K, M, dens, seed) in enumerate(PANEL_CASES):
        rng = np.random.default_rng(seed)
        rows, _ = random_cref(K, M, dens, rng)
        pbwt, count0, absperm = oracle.ref_build_ref_pbwt(rows, K)
        out[f"p{i}_K"] = np.int64(K)
        out[f"p{i}_rows"] = rows
        out[f"p{i}_pbwt"] = pbwt
        out[f"p{i}_count0"] = count0
        out[f"p{i}_absperm"] = absperm
    out["ncases"] = np.int64(len(PANEL_CASES))
    np.savez_compressed(os.path.join(HERE, "ref_panel_pbwt.npz"), **out)


PHASE_GOLDEN = [("tiny_beam_small_K", 1), ("missing_imputed_no_reverse", 0),
                ("roh_hom_splits_prephase", 0), ("long_roh_deep_history", 2),
                ("few_founders_full_panel", 0), ("noisy_wide_beam", 1)]


def phase_fixtures():
    out = {}
    for i, (name, seed) in enumerate(PHASE_GOLDEN):
        c = phase_cases.make_case(name, seed)
        j = phase_cases.job_for(c)
        res = oracle.ref_phase_target(phase_cases.params_for(c), j, c["r"]["monos"])
        out[f"c{i}_name"] = np.array(name)
        out[f"c{i}_seed"] = np.int64(seed)
        out[f"c{i}_inhash"] = np.array(phase_cases.input_hash(c))
        out[f"c{i}_mat"] = np.packbits(res.mat, bitorder="little")
        out[f"c{i}_pat"] = np.packbits(res.pat, bitorder="little")
        out[f"c{i}_dm"] = res.dosage_mat
        out[f"c{i}_dp"] = res.dosage_pat
        out[f"c{i}_cc"] = res.call_conf
        out[f"c{i}_tot"] = np.float32(res.total_conf)
        out[f"c{i}_counts"] = np.array([res.ncalls, res.ncallsites, res.nmissing, res.nmonhets],
                                       np.int64)
    out["ncases"] = np.int64(len(PHASE_GOLDEN))
    np.savez_compressed(os.path.join(HERE, "ref_phase.npz"), **out)


def rle_vectors(rng):
    """Haplotype vectors with sparse, dense, long-run and leading-one
    structure, including runs longer than 32767 and 65534 bits."""
    out = []
    for t in range(48):
        nw = int(rng.choice([8, 16, 64, 256]))
        n = nw * 64
        kind = t % 6
        if kind == 0:
            bits = rng.random(n) < rng.random() * 0.05
        elif kind == 1:
            bits = rng.random(n) < 0.5
        elif kind == 2:
            bits = np.zeros(n, bool)
            a = int(rng.integers(0, n))
            bits[a:a + int(rng.integers(1, n))] = True
        elif kind == 3:
            bits = np.ones(n, bool)
            bits[int(rng.integers(0, n))] = False
        elif kind == 4:
            bits = np.repeat(rng.random(n // 128) < 0.5, 128)
        else:
            bits = rng.random(n) < 0.01
            bits[0] = True
        out.append(bits)
    for n, ones in ((1100 * 64, True), (1100 * 64, False), (520 * 64, True)):
        bits = np.full(n, ones)
        bits[n // 3] = not ones
        out.append(bits)
    return [np.packbits(b.astype(np.uint8), bitorder="little").view(np.uint64) for b in out]


def rle_fixtures():
    rng = np.random.default_rng(77)
    out = {}
    vecs = rle_vectors(rng)
    for i, w in enumerate(vecs):
        code = oracle.ref_rle_encode(w)
        out[f"v{i}_words"] = w
        out[f"v{i}_code"] = np.frombuffer(code, np.uint8) if code is not None else np.zeros(0, np.uint8)
        out[f"v{i}_coded"] = np.int64(code is not None)
    out["nvec"] = np.int64(len(vecs))
    np.savez_compressed(os.path.join(HERE, "ref_rle.npz"), **out)


def rle_masked_cases(rng):
    """(samples, haploid flags, record bits, masked samples): masks at the
    first and last sample, across word and 512-bit boundaries, runs of
    neighbours, all but one and all samples; haploid samples stored
    homozygous and, a few, heterozygous."""
    cases = []
    for t in range(40):
        N = int(rng.choice([1, 2, 31, 32, 33, 200, 256, 257, 700]))
        p = float(rng.choice([0.01, 0.05, 0.95, 0.99]))
        bits = (rng.random(2 * N) < p).astype(np.uint8)
        if t % 3 == 0:  # runs of samples sharing an allele
            bits = np.repeat(rng.random(2 * N // 8 + 1) < 0.5, 8)[:2 * N].astype(np.uint8)
        hap = (rng.random(N) < float(rng.choice([0.0, 0.2, 1.0]))).astype(np.uint8)
        for s in np.flatnonzero(hap):
            if rng.random() < 0.85:
                bits[2 * s + 1] = bits[2 * s]
        kind = t % 8
        if kind == 0:
            ms = [0]
        elif kind == 1:
            ms = [N - 1]
        elif kind == 2:
            ms = [s for s in (15, 16, 31, 32, 255, 256) if s < N]
        elif kind == 3:
            a = int(rng.integers(0, N))
            ms = list(range(a, min(N, a + int(rng.integers(1, 40)))))
        elif kind == 4:
            keep = int(rng.integers(0, N))
            ms = [s for s in range(N) if s != keep]
        elif kind == 5:
            ms = list(range(N))
        else:
            ms = sorted(rng.choice(N, int(rng.integers(0, N + 1)), replace=False).tolist())
        cases.append((N, hap, bits, np.asarray(sorted(set(ms)), np.uint64)))
    return cases


def rle_masked_fixtures():
    rng = np.random.default_rng(91)
    out = {}
    n = 0
    for N, hap, bits, ms in rle_masked_cases(rng):
        hw = (2 * N + 511) // 512 * 8
        full = np.zeros(hw * 64, np.uint8)
        full[:2 * N] = bits
        w = np.packbits(full, bitorder="little").view(np.uint64)
        code = oracle.ref_rle_encode(w)
        if code is None:
            continue
        Nk = N - ms.size
        nwk = (2 * Nk + 511) // 512 * 8
        got, cnt = oracle.ref_rle_decode_masked(code, nwk, ms, hap[ms.astype(np.int64)])
        out[f"c{n}_N"] = np.int64(N)
        out[f"c{n}_haploid"] = hap
        out[f"c{n}_words"] = w
        out[f"c{n}_code"] = np.frombuffer(code, np.uint8)
        out[f"c{n}_mask"] = ms
        out[f"c{n}_rows"] = got
        out[f"c{n}_masked_alleles"] = np.int64(cnt)
        n += 1
    out["ncases"] = np.int64(n)
    np.savez_compressed(os.path.join(HERE, "ref_rle_masked.npz"), **out)


FIXTURES = {"pbwt": pbwt_fixtures, "genotype": genotype_fixtures,
            "panel_pbwt": panel_pbwt_fixtures, "phase": phase_fixtures, "rle": rle_fixtures,
            "rle_masked": rle_masked_fixtures}

if __name__ == "__main__":
    oracle.build()
    if not oracle.ref_available():
        sys.exit("reference harness not built (is /root/reference present?)")
    names = sys.argv[1:] or list(FIXTURES)
    for name in names:
        FIXTURES[name]()
    for name in names:
        f = f"ref_{name}.npz"
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")
