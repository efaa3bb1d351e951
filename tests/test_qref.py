"""Qref panel reader (eimp_qref_*, host code): the run-length code matches
the reference's coder (golden vectors written by src/utils.h:133-190, and
live where the harness is built), files written as the reference writes them
read back to the exact panel rows and metadata, and damaged files fail."""
import os
import struct

import numpy as np
import pytest

import oracle
import qref_writer as qw
from eagleimp_amd import synth
from eagleimp_amd.engine import EimpError
from eagleimp_amd.qref import Qref

RLE = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_rle.npz"))


def test_rle_coder_matches_reference_golden():
    for i in range(int(RLE["nvec"])):
        w = RLE[f"v{i}_words"]
        code = qw.rle_encode(w)
        if int(RLE[f"v{i}_coded"]):
            assert code == RLE[f"v{i}_code"].tobytes(), i
            assert np.array_equal(qw.rle_decode(code, w.size), w), i
        else:
            assert code is None, i


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built")
def test_rle_coder_matches_reference_live():
    rng = np.random.default_rng(3)
    for t in range(120):
        nw = int(rng.choice([8, 24, 64, 300]))
        p = rng.random() ** 3
        bits = (rng.random(nw * 64) < p).astype(np.uint8)
        if t % 7 == 0:
            bits[: rng.integers(1, nw * 64)] ^= 1
        w = np.packbits(bits, bitorder="little").view(np.uint64)
        code = qw.rle_encode(w)
        assert code == oracle.ref_rle_encode(w), t
        if code is not None:
            assert np.array_equal(oracle.ref_rle_decode(code, nw), w)


def test_spec_example_decodes():
    # the run-length example of the format description, completed to a word
    # boundary by a final run of 42 zeros
    code = bytes([0x0A, 0x03, 0xAA, 0x02, 0xFF, 0xFF, 0x00, 0x20, 42])
    bits = np.unpackbits(qw.rle_decode(code, 518).view(np.uint8), bitorder="little")
    want = np.concatenate([np.zeros(10), np.ones(3), np.zeros(298), np.ones(32799), np.zeros(42)])
    assert np.array_equal(bits, want)


def make_panel_file(tmp_path, seed=1, N=300, M=700, chrom=1, haploid=None, force_raw=False,
                    name="p.qref"):
    panel = synth.make_panel(seed, N=2 * N, M=M, F=12, T=1)  # 2N haplotypes = N samples
    rng = np.random.default_rng(seed)
    rows = panel.site.copy()
    if haploid is not None:  # haploid samples are stored homozygous
        bits = np.unpackbits(rows.view(np.uint8), axis=1, bitorder="little")[:, :2 * N]
        for s in np.flatnonzero(haploid):
            bits[:, 2 * s + 1] = bits[:, 2 * s]
        full = np.zeros((M, rows.shape[1] * 64), np.uint8)
        full[:, :2 * N] = bits
        rows = np.packbits(full, axis=1, bitorder="little").view(np.uint64)
    pos = np.cumsum(rng.integers(1, 500, M)).astype(np.int64)
    af = rng.random(M).astype(np.float32)
    alleles = [("ACGT"[m % 4], "ACGT"[(m + 1 + m // 4) % 4] * (1 + m % 3)) for m in range(M)]
    ids = [f"rs{m}" if m % 5 else "" for m in range(M)]
    ma = rng.random(M) < 0.05
    path = str(tmp_path / name)
    qw.write_qref(path, rows, N, pos, af, alleles, ids, chrom=chrom, haploid=haploid,
                  ma_flags=ma, force_raw=force_raw)
    return path, dict(rows=rows, pos=pos, af=af, alleles=alleles, ids=ids, ma=ma, N=N, M=M)


@pytest.mark.parametrize("force_raw", [False, True])
def test_round_trip(tmp_path, force_raw):
    path, w = make_panel_file(tmp_path, force_raw=force_raw)
    with Qref(path) as q:
        assert (q.Nref, q.Mref, q.info.NrefHap, q.info.chrom) == (w["N"], w["M"], 0, 1)
        assert (q.info.v_major, q.info.v_minor, q.info.has_haploid_flags) == (1, 1, 0)
        assert q.info.MrefMA == int(w["ma"].sum())
        rows = q.site_rows()
        nw = (2 * w["N"] + 63) // 64
        assert rows.shape == (w["M"], (2 * w["N"] + 511) // 512 * 8)
        assert np.array_equal(rows[:, :nw], w["rows"][:, :nw])
        assert not rows[:, nw:].any()
        md = q.metadata()
        assert np.array_equal(md["positions"], w["pos"])
        assert np.array_equal(md["af"].view(np.uint32), w["af"].view(np.uint32))
        assert np.array_equal(md["multiallelic"], w["ma"])
        assert not md["haploid"].any()
        for m in (0, 1, 17, w["M"] - 1):
            assert q.variant(m) == (w["alleles"][m][0], w["alleles"][m][1], w["ids"][m])
        # a sub-range into wider rows, on one thread
        part = q.site_rows(100, 50, row_words=nw + 3, threads=1)
        assert np.array_equal(part[:, :nw], w["rows"][100:150, :nw]) and not part[:, nw:].any()


def test_haploid_samples_and_chrx(tmp_path):
    hap = np.zeros(300, np.uint8)
    hap[[0, 7, 150, 299]] = 1
    path, w = make_panel_file(tmp_path, chrom=23, haploid=hap)
    with Qref(path) as q:
        assert q.info.has_haploid_flags == 1 and q.info.NrefHap == 4
        assert np.array_equal(q.haploid_flags(), hap)
        assert q.Ncorr == 600 - 4 and q.Nad == 600
        nw = (600 + 63) // 64
        assert np.array_equal(q.site_rows()[:, :nw], w["rows"][:, :nw])
    # chromosome X stores the flags even when no sample is haploid
    path2, _ = make_panel_file(tmp_path, chrom=23, name="x.qref")
    with Qref(path2) as q:
        assert q.info.has_haploid_flags == 1 and not q.haploid_flags().any()
    # all haploid: no flags stored, every sample haploid
    path3, _ = make_panel_file(tmp_path, chrom=24, haploid=np.ones(300, np.uint8), name="y.qref")
    with Qref(path3) as q:
        assert q.info.has_haploid_flags == 0 and q.haploid_flags().all()


def test_damaged_files_fail(tmp_path):
    path, w = make_panel_file(tmp_path, M=50)
    data = open(path, "rb").read()

    def variant(b, name):
        p = str(tmp_path / name)
        open(p, "wb").write(b)
        return p

    for bad in (b"QRFF" + data[4:],                       # magic
                data[:4] + bytes([2, 1]) + data[6:],      # major version
                data[:4] + bytes([1, 2]) + data[6:],      # newer minor version
                data[:-5],                                # truncated haplotypes
                data[:30]):                               # truncated constants
        with pytest.raises(EimpError):
            Qref(variant(bad, "bad.qref"))
    with pytest.raises(EimpError):
        Qref(str(tmp_path / "missing.qref"))
    # older minor version is accepted
    with Qref(variant(data[:4] + bytes([1, 0]) + data[6:], "old.qref")) as q:
        assert q.Mref == 50


def test_truncated_run_length_code_fails(tmp_path):
    """A code that stops inside a word is rejected when the site is read."""
    N, M = 40, 3
    rows = np.zeros((M, 2), np.uint64)
    rows[:, 0] = 0x00FF00FF00FF00FF
    path = str(tmp_path / "t.qref")
    qw.write_qref(path, rows, N, np.arange(M), np.zeros(M, np.float32),
                  [("A", "C")] * M, [""] * M)
    data = bytearray(open(path, "rb").read())
    # the last variant's code: shorten it by its final run (ends inside a word)
    off = len(data)
    code_len = None
    for L in range(1, 64):
        if struct.unpack("<Q", bytes(data[off - L - 8:off - L]))[0] == L:
            code_len = L
            break
    assert code_len is not None
    start = off - code_len - 8
    short = data[:start] + struct.pack("<Q", code_len - 1) + data[start + 8:off - 1]
    p2 = str(tmp_path / "t2.qref")
    open(p2, "wb").write(bytes(short))
    with Qref(p2) as q:
        q.site_rows(0, 2)
        with pytest.raises(EimpError):
            q.site_rows(2, 1)


# ---------------------------------------------------------------------------
# sample masking (participant withdrawals)

MASKED = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                              "ref_rle_masked.npz"))


def _one_record_file(tmp_path, N, hap, words, name, force_raw=False, af=0.25):
    """A chromosome-X panel (haploid flags stored) of one variant."""
    path = str(tmp_path / name)
    qw.write_qref(path, words[None, :], N, [5], np.array([af], np.float32), [("A", "C")], ["v"],
                  chrom=23, haploid=hap, force_raw=force_raw)
    return path


def test_masked_decoder_matches_reference_golden(tmp_path):
    """Coded records read with a mask give the reference decoder's kept
    haplotypes and masked allele count (golden vectors of src/utils.h:193-315)."""
    from oracle import qref_mask
    for i in range(int(MASKED["ncases"])):
        N, hap, w = int(MASKED[f"c{i}_N"]), MASKED[f"c{i}_haploid"], MASKED[f"c{i}_words"]
        ms, want = MASKED[f"c{i}_mask"], MASKED[f"c{i}_rows"]
        assert qw.rle_encode(w) == MASKED[f"c{i}_code"].tobytes(), i
        path = _one_record_file(tmp_path, N, hap, w, f"g{i}.qref")
        with Qref(path) as q:
            q.mask(ms)
            Nk = N - ms.size
            assert q.Nref == Nk and q.info.hap_words == (2 * Nk + 511) // 512 * 8
            assert q.info.NrefHap == int(hap.sum()) - int(hap[ms.astype(np.int64)].sum())
            rows = q.site_rows()
            assert rows.shape == (1, want.size) and np.array_equal(rows[0], want), i
            _, cnt = q.allele_freqs()
            assert int(cnt[0]) == int(MASKED[f"c{i}_masked_alleles"]), i
            assert np.array_equal(q.haploid_flags(), qref_mask.kept_haploid(hap, ms))
        # the restated rule agrees with the reference's decoder
        r, c = qref_mask.masked_record(w, N, ms, hap[ms.astype(np.int64)], False, want.size)
        assert np.array_equal(r, want) and c == int(MASKED[f"c{i}_masked_alleles"]), i


@pytest.mark.skipif(not oracle.ref_available(), reason="reference harness not built")
def test_masked_decoder_matches_reference_live(tmp_path):
    from oracle import qref_mask
    rng = np.random.default_rng(17)
    done = 0
    for t in range(150):
        N = int(rng.integers(1, 900))
        hap = (rng.random(N) < rng.choice([0.0, 0.3, 1.0])).astype(np.uint8)
        bits = np.repeat(rng.random(2 * N // 4 + 1) < rng.random(), 4)[:2 * N]
        bits = bits ^ (rng.random(2 * N) < 0.01)
        for s in np.flatnonzero(hap):
            if rng.random() < 0.9:
                bits[2 * s + 1] = bits[2 * s]
        full = np.zeros((2 * N + 511) // 512 * 512, np.uint8)
        full[:2 * N] = bits
        w = np.packbits(full, bitorder="little").view(np.uint64)
        code = qw.rle_encode(w)
        if code is None:
            continue
        ms = np.sort(rng.choice(N, int(rng.integers(0, min(N, 60) + 1)), replace=False))
        ms = ms.astype(np.uint64)
        nwk = (2 * (N - ms.size) + 511) // 512 * 8
        want, wc = oracle.ref_rle_decode_masked(code, nwk, ms, hap[ms.astype(np.int64)])
        r, c = qref_mask.masked_record(w, N, ms, hap[ms.astype(np.int64)], False, nwk)
        assert np.array_equal(r, want) and c == wc, t
        if t % 5 == 0:
            path = _one_record_file(tmp_path, N, hap, w, "live.qref")
            with Qref(path) as q:
                q.mask(ms)
                assert np.array_equal(q.site_rows()[0], want) and int(q.allele_freqs()[1][0]) == wc
        done += 1
    assert done > 50


@pytest.mark.parametrize("force_raw", [False, True])
@pytest.mark.parametrize("chrom,hapmode", [(1, "none"), (23, "some"), (24, "all")])
def test_masked_panel_rows_counts_and_af(tmp_path, force_raw, chrom, hapmode):
    """Whole panels, coded and raw: rows equal the records with the masked bit
    pairs removed, counts follow each record kind's rule, frequencies the
    reference's correction, haploid flags and counts shrink."""
    from oracle import qref_mask
    N, M = 300, 400
    hap = {"none": None, "some": (np.arange(N) % 7 == 3).astype(np.uint8),
           "all": np.ones(N, np.uint8)}[hapmode]
    path, w = make_panel_file(tmp_path, seed=5, N=N, M=M, chrom=chrom, haploid=hap,
                              force_raw=force_raw)
    hapf = np.zeros(N, np.uint8) if hap is None else hap
    ms = np.array([0, 1, 3, 31, 32, 33, 100, 255, 256, 299], np.uint64)
    with Qref(path) as q:
        q.mask(ms)
        Nk = N - ms.size
        assert q.Nref == Nk and q.Nad == 2 * Nk
        assert q.info.NrefHap == int(hapf.sum()) - int(hapf[ms.astype(np.int64)].sum())
        assert np.array_equal(q.haploid_flags(), qref_mask.kept_haploid(hapf, ms))
        rows = q.site_rows(threads=3)
        af, cnt = q.allele_freqs(threads=2)
        old_an, new_an = 2 * N - int(hapf.sum()), 2 * Nk - int(q.info.NrefHap)
        for m in range(M):
            r, c = qref_mask.masked_record(w["rows"][m], N, ms, hapf[ms.astype(np.int64)],
                                           force_raw, rows.shape[1])
            assert np.array_equal(rows[m], r), m
            assert int(cnt[m]) == c, m
            want = qref_mask.corrected_af(w["af"][m], old_an, new_an, c)
            assert af[m].view(np.uint32) == want.view(np.uint32), m
        # sub-ranges agree with the whole
        assert np.array_equal(q.site_rows(37, 50, threads=1), rows[37:87])
        a2, c2 = q.allele_freqs(37, 50, threads=1)
        assert np.array_equal(a2.view(np.uint32), af[37:87].view(np.uint32))
        # clearing the mask restores the file
        q.mask([])
        assert q.Nref == N and np.array_equal(q.allele_freqs()[0].view(np.uint32),
                                              w["af"].view(np.uint32))
        assert not q.allele_freqs()[1].any()
        nw = (2 * N + 63) // 64
        assert np.array_equal(q.site_rows()[:, :nw], w["rows"][:, :nw])


def test_raw_and_coded_haploid_counts(tmp_path):
    """A haploid sample stored heterozygous counts differently in the two
    record kinds: coded records only count (1,1), raw records any set bit."""
    N = 64
    hap = np.zeros(N, np.uint8)
    hap[[2, 5, 9]] = 1
    bits = np.zeros(512, np.uint8)
    bits[[4, 5, 10, 19]] = 1  # sample 2 (1,1), sample 5 (1,0), sample 9 (0,1)
    bits[[20, 21, 40]] = 1    # diploid 10 (1,1), diploid 20 (1,0)
    w = np.packbits(bits, bitorder="little").view(np.uint64)
    ms = np.array([2, 5, 9, 10, 20], np.uint64)
    for raw, want in ((False, 1 + 0 + 0 + 2 + 1), (True, 1 + 1 + 1 + 2 + 1)):
        with Qref(_one_record_file(tmp_path, N, hap, w, f"h{raw}.qref", force_raw=raw)) as q:
            q.mask(ms)
            assert int(q.allele_freqs()[1][0]) == want
            assert not q.site_rows().any()


def test_af_correction_closed_form():
    """With old_an a power of two the file's frequency ac/old_an is exact,
    so the corrected frequency is (ac - masked) / new_an rounded once."""
    from fractions import Fraction
    from oracle import qref_mask
    rng = np.random.default_rng(8)
    for _ in range(2000):
        old_an = 1 << int(rng.integers(1, 20))
        ac = int(rng.integers(0, old_an + 1))
        masked = int(rng.integers(0, ac + 1))
        new_an = int(rng.integers(1, old_an + 1))
        got = qref_mask.corrected_af(np.float32(ac / old_an), old_an, new_an, masked)
        assert got == qref_mask.round_f32(Fraction(ac - masked, new_an))


def test_af_correction_rounds_as_the_reference_build(tmp_path):
    """The frequency expression compiled by g++ with the reference's flags
    (-std=c++11 -O3 -march=native, src/CMakeLists.txt:165) gives the
    restated (fused) rounding on this host."""
    import shutil
    import subprocess
    from oracle import qref_mask
    gxx = shutil.which("g++")
    if gxx is None or "fma" not in open("/proc/cpuinfo").read():
        pytest.skip("needs g++ and an FMA host")
    src = tmp_path / "af.cpp"
    src.write_text(
        "#include <cstdio>\n#include <cstdlib>\n#include <cstring>\n"
        "static float af(std::size_t an0, float f, std::size_t k, std::size_t an1) {\n"
        "  float ac = an0 * f - k;\n  return ac / an1;\n}\n"
        "int main() {\n  unsigned long a, k, b; unsigned int u;\n"
        "  while (std::scanf(\"%lu %x %lu %lu\", &a, &u, &k, &b) == 4) {\n"
        "    float f; std::memcpy(&f, &u, 4); float r = af(a, f, k, b);\n"
        "    std::memcpy(&u, &r, 4); std::printf(\"%08x\\n\", u);\n  }\n}\n")
    exe = tmp_path / "af"
    subprocess.run([gxx, "-std=c++11", "-O3", "-march=native", str(src), "-o", str(exe)],
                   check=True)
    rng = np.random.default_rng(9)
    cases = []
    for _ in range(3000):
        a = int(rng.integers(2, 2_000_000))
        f = np.float32(rng.random())
        k = int(rng.integers(0, 200))
        b = int(rng.integers(1, a + 1))
        cases.append((a, f, k, b))
    inp = "".join(f"{a} {int(f.view(np.uint32)):08x} {k} {b}\n" for a, f, k, b in cases)
    out = subprocess.run([str(exe)], input=inp, capture_output=True, text=True, check=True).stdout
    unfused_differs = 0
    for (a, f, k, b), line in zip(cases, out.split()):
        want = qref_mask.corrected_af(f, a, b, k)
        assert int(line, 16) == int(want.view(np.uint32)), (a, f, k, b)
        unf = np.float32(np.float32(np.float32(a) * f) - np.float32(k)) / np.float32(b)
        unfused_differs += int(unf.view(np.uint32) != want.view(np.uint32))
    assert unfused_differs > 0  # the cases tell the fused from the unfused rounding


def test_sample_lists_and_mask_errors(tmp_path):
    from eagleimp_amd.qref import parse_sample_list
    plain = tmp_path / "s.txt"
    plain.write_text("S1\nW17 x\n-3\n\nW\nWx1\n-\nW12a\nW0042\n   -9  \nS2")
    ms, n = parse_sample_list(str(plain))
    assert n == 11 and ms.tolist() == [1, 2, 8, 9]
    bgen = tmp_path / "s.sample"
    bgen.write_text("ID_1 ID_2 missing\n0 0 0\nA S1 0\nB W5 0\nC -77 0\nD S4 0\n")
    ms, n = parse_sample_list(str(bgen))
    assert n == 4 and ms.tolist() == [1, 2]
    broken = tmp_path / "b.sample"
    broken.write_text("h\nh\nA S1\nB\n")
    with pytest.raises(EimpError):
        parse_sample_list(str(broken))
    path, w = make_panel_file(tmp_path, N=4, M=30)
    with Qref(path) as q:
        q.mask_sample_list(str(bgen))  # four samples listed, four in the panel
        assert q.Nref == 2
        for bad in ([3, 1], [1, 1], [4]):
            with pytest.raises(EimpError):
                q.mask(bad)
        assert q.Nref == 2  # a refused mask leaves the last one
        with pytest.raises(EimpError):
            q.mask_sample_list(str(plain))  # eleven samples listed
