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
