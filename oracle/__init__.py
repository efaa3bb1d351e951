"""ctypes bindings of the CPU oracle (eimp_oracle.c) and of the reference harness.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libeimp_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libeimp_ref.so")

_P = C.c_void_p
_U64 = C.c_uint64


class EoPanel(C.Structure):
    _fields_ = [("N_ad", _U64), ("N_corr", _U64), ("M", _U64), ("hap", _P), ("hap_stride", _U64),
                ("site", _P), ("site_stride", _U64), ("cM", _P), ("haploid_ref", _P),
                ("ignore_sample", C.c_int64)]


class EoTarget(C.Structure):
    _fields_ = [("gt", _P), ("s0", _U64), ("W", _U64), ("K", _U64), ("cMmaxSplit", C.c_float)]


class EoResult(C.Structure):
    _fields_ = [("K", _U64), ("W", _U64), ("S", _U64), ("Mp", _U64), ("ncalls", _U64),
                ("besthaps", _P), ("flags", _P), ("monos", _P), ("splits", _P),
                ("splitsites", _P), ("callLocs", _P), ("cMdiffs", _P), ("cref", _P),
                ("count0", _P), ("pbwt", _P), ("dig_pbwt", _U64), ("dig_count0", _U64)]


_LIB = None
_REF = None


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference harness when the sources exist)."""
    args = ["make", "-s", "-f", os.path.join(HERE, "Makefile"), "all"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: build the oracle first")
        L = C.CDLL(ORACLE_SO)
        L.eo_unit_words.restype = _U64
        L.eo_unit_words.argtypes = [_U64]
        L.eo_pbwt_row_words.restype = _U64
        L.eo_pbwt_row_words.argtypes = [_U64]
        L.eo_gpu_row_words.restype = _U64
        L.eo_gpu_row_words.argtypes = [_U64]
        L.eo_compare_with_preinit.restype = _U64
        L.eo_compare_with_preinit.argtypes = [_P, _U64, _P, _U64]
        L.eo_find_best_haps.argtypes = [_P, _P, _P, _P]
        L.eo_find_monos.argtypes = [_P, _P, _P, _P]
        L.eo_find_call_locs.restype = _U64
        L.eo_find_call_locs.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P]
        L.eo_condensed_ref.argtypes = [_P, _P, _P, _P, _U64, _P, _P]
        L.eo_build_pbwt.argtypes = [_P, _U64, _U64, _U64, _P, _P]
        L.eo_pbwt_to_gpu_layout.argtypes = [_P, _U64, _U64, _P]
        L.eo_map_interval.argtypes = [_P, _U64, _U64, _P, C.c_int, _U64, C.c_int32, C.c_int32, _P]
        L.eo_decode_rel_perm.restype = C.c_int32
        L.eo_decode_rel_perm.argtypes = [_P, C.c_int32, C.c_int32]
        L.eo_mix64.restype = _U64
        L.eo_mix64.argtypes = [_U64]
        L.eo_digest.argtypes = [_P, _U64, _U64, _U64, _P, _P, _P]
        L.eo_build_ref_pbwt.argtypes = [_P, _U64, _U64, _U64, _P, _P, _P]
        L.eo_build_ref_pbwt.restype = None
        L.eo_process_target.restype = C.c_int
        L.eo_process_target.argtypes = [_P, _P, C.c_int, _P]
        L.eo_result_free.argtypes = [_P]
        L.eo_process_batch.restype = C.c_int
        L.eo_process_batch.argtypes = [_P, _U64, _P, C.c_int, _P, _P, _P, _P]
        _LIB = L
    return _LIB


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _REF
    if _REF is None:
        if not ref_available():
            raise RuntimeError(f"{REF_SO} missing (reference sources absent?)")
        R = C.CDLL(REF_SO)
        R.ref_build_pbwt.restype = C.c_int
        R.ref_build_pbwt.argtypes = [_P, _U64, _U64, _U64, _P, _P, _U64, _P, _P]
        R.ref_genotype_compare.restype = C.c_int
        R.ref_genotype_compare.argtypes = [_P, _U64, _P, _U64, _P, _U64, _P]
        R.ref_ingest_raw.restype = C.c_int
        R.ref_ingest_raw.argtypes = [_P, _U64, _U64, _U64, _P, _P]
        R.ref_build_ref_pbwt.restype = C.c_int
        R.ref_build_ref_pbwt.argtypes = [_P, _U64, _U64, _U64, _P, _P, _P]
        R.ref_phase_target.restype = C.c_int
        R.ref_phase_target.argtypes = [_P, _P, _P, _P]
        R.ref_rle_encode.restype = C.c_int64
        R.ref_rle_encode.argtypes = [_P, _U64, _P, _U64]
        R.ref_rle_decode.restype = C.c_int
        R.ref_rle_decode.argtypes = [_P, _U64, _P, _U64]
        R.ref_rle_decode_masked.restype = C.c_int
        R.ref_rle_decode_masked.argtypes = [_P, _U64, _P, _U64, _P, _P, _U64, _P]
        _REF = R
    return _REF


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


# ---------------------------------------------------------------------------
# small helpers
# ---------------------------------------------------------------------------
def unit_words(nbits: int) -> int:
    return ((nbits + 511) // 512) * 8


def pbwt_row_words(K: int) -> int:
    return 2 * ((K + 31) // 32)


def gpu_row_words(K: int) -> int:
    return 4 * (((K + 511) // 512) * 8)


class PanelView:
    """Keeps numpy arrays alive for an EoPanel."""

    def __init__(self, hap, site, cM, N_ad=None, N_corr=None, M=None, haploid_ref=None,
                 ignore_sample=-1):
        self.hap = np.ascontiguousarray(hap, dtype=np.uint64)
        self.site = np.ascontiguousarray(site, dtype=np.uint64)
        self.cM = np.ascontiguousarray(cM, dtype=np.float32)
        self.haploid = None if haploid_ref is None else np.ascontiguousarray(haploid_ref,
                                                                              dtype=np.uint8)
        N_ad = self.hap.shape[0] if N_ad is None else N_ad
        M = self.cM.shape[0] if M is None else M
        if N_corr is None:
            N_corr = N_ad
            if self.haploid is not None:
                N_corr -= int(self.haploid.sum())
            if ignore_sample >= 0:
                N_corr -= 1 if (self.haploid is not None and self.haploid[ignore_sample]) else 2
        self.c = EoPanel(N_ad, N_corr, M, _p(self.hap), self.hap.shape[1], _p(self.site),
                         self.site.shape[1], _p(self.cM), _p(self.haploid), ignore_sample)

    @property
    def N_corr(self):
        return self.c.N_corr


def make_target(gt, s0, W, K, cMmaxSplit=0.125):
    gt = np.ascontiguousarray(gt, dtype=np.uint64)
    t = EoTarget(_p(gt), int(s0), int(W), int(K), float(np.float32(cMmaxSplit)))
    t._keep = gt
    return t


def process_target(pv: PanelView, gt, s0, W, K, cMmaxSplit=0.125, keep=True,
                   rows=True) -> dict:
    """Runs a4..a9 for one target and returns every intermediate as numpy
    (rows=False: all but the condensed reference and the cPBWT rows)."""
    K = min(int(K), int(pv.N_corr))
    t = make_target(gt, s0, W, K, cMmaxSplit)
    r = EoResult()
    rc = lib().eo_process_target(C.byref(pv.c), C.byref(t), 1 if keep else 0, C.byref(r))
    if rc != 0:
        raise RuntimeError(f"eo_process_target failed ({rc})")
    out = dict(K=r.K, W=r.W, S=r.S, Mp=r.Mp, ncalls=r.ncalls, dig_pbwt=r.dig_pbwt,
               dig_count0=r.dig_count0)
    if keep:
        def arr(ptr_, n, ct, dt):
            if n == 0:
                return np.zeros(0, dtype=dt)
            return np.ctypeslib.as_array(C.cast(ptr_, C.POINTER(ct)), shape=(n,)).copy().view(dt)
        Nw = unit_words(pv.c.N_ad)
        Ww = unit_words(int(W))
        Kw = unit_words(K)
        out["besthaps"] = arr(r.besthaps, K, C.c_uint32, np.uint32)
        out["flags"] = arr(r.flags, Nw, C.c_uint64, np.uint64)
        out["monos"] = arr(r.monos, Ww, C.c_uint64, np.uint64)
        out["splits"] = arr(r.splits, Ww, C.c_uint64, np.uint64)
        out["splitsites"] = arr(r.splitsites, r.S, C.c_uint32, np.uint32)
        out["callLocs"] = arr(r.callLocs, r.ncalls, C.c_uint32, np.uint32)
        out["cMdiffs"] = arr(r.cMdiffs, r.S + 1 if r.S else 0, C.c_float, np.float32)
        out["count0"] = arr(r.count0, r.Mp, C.c_int32, np.int32)
        if rows:
            out["cref"] = arr(r.cref, r.Mp * Kw, C.c_uint64, np.uint64).reshape(r.Mp, Kw)
            out["pbwt"] = arr(r.pbwt, 2 * r.Mp * pbwt_row_words(K), C.c_uint32,
                              np.uint32).reshape(2 * r.Mp, pbwt_row_words(K))
        lib().eo_result_free(C.byref(r))
    return out


def process_batch(pv: PanelView, gts, s0s, Ws, Ks, cMmaxSplit=0.125, nthreads=0):
    n = len(gts)
    ts = (EoTarget * n)()
    keep = []
    for i in range(n):
        K = min(int(Ks[i]), int(pv.N_corr))
        t = make_target(gts[i], s0s[i], Ws[i], K, cMmaxSplit)
        keep.append(t._keep)
        ts[i] = t
    S = np.zeros(n, np.uint64)
    Mp = np.zeros(n, np.uint64)
    dp = np.zeros(n, np.uint64)
    dc = np.zeros(n, np.uint64)
    rc = lib().eo_process_batch(C.byref(pv.c), n, ts, nthreads, _p(S), _p(Mp), _p(dp), _p(dc))
    if rc != 0:
        raise RuntimeError("eo_process_batch failed")
    return dict(S=S, Mp=Mp, dig_pbwt=dp, dig_count0=dc)


def build_pbwt(cref, K, count0):
    cref = np.ascontiguousarray(cref, dtype=np.uint64)
    count0 = np.ascontiguousarray(count0, dtype=np.int32)
    Mp = cref.shape[0]
    out = np.zeros((2 * Mp, pbwt_row_words(K)), dtype=np.uint32)
    lib().eo_build_pbwt(_p(cref), cref.shape[1], K, Mp, _p(count0), _p(out))
    return out


def map_interval(pbwt, K, count0, reverse, m, l, r):
    pbwt = np.ascontiguousarray(pbwt, dtype=np.uint32)
    count0 = np.ascontiguousarray(count0, dtype=np.int32)
    o = np.zeros(4, np.int32)
    Mp = pbwt.shape[0] // 2
    lib().eo_map_interval(_p(pbwt), K, Mp, _p(count0), int(reverse), m, l, r, _p(o))
    return tuple(int(x) for x in o)


def digest(pbwt_rows, K, count0):
    """pbwt_rows: (2*Mp, stride) uint32 in any stride >= 2*ceil(K/32)."""
    pbwt_rows = np.ascontiguousarray(pbwt_rows, dtype=np.uint32)
    count0 = np.ascontiguousarray(count0, dtype=np.int32)
    dp = np.zeros(1, np.uint64)
    dc = np.zeros(1, np.uint64)
    Mp = pbwt_rows.shape[0] // 2
    lib().eo_digest(_p(pbwt_rows), pbwt_rows.shape[1], K, Mp, _p(count0), _p(dp), _p(dc))
    return int(dp[0]), int(dc[0])


def compare_with_preinit(gt, hap, hap_words):
    gt = np.ascontiguousarray(gt, dtype=np.uint64)
    hap = np.ascontiguousarray(hap, dtype=np.uint64)
    return int(lib().eo_compare_with_preinit(_p(gt), gt.shape[0], _p(hap), hap_words))


# ---------------------------------------------------------------------------
# reference harness
# ---------------------------------------------------------------------------
def ref_build_pbwt(cref, K, count0, queries=None):
    cref = np.ascontiguousarray(cref, dtype=np.uint64)
    count0 = np.ascontiguousarray(count0, dtype=np.int32)
    Mp = cref.shape[0]
    out = np.zeros((2 * Mp, pbwt_row_words(K)), dtype=np.uint32)
    q = np.zeros((0, 4), np.int32) if queries is None else np.ascontiguousarray(queries, np.int32)
    qo = np.zeros((max(len(q), 1), 4), np.int32)
    rc = ref().ref_build_pbwt(_p(cref), cref.shape[1], K, Mp, _p(count0), _p(out), len(q),
                              _p(q) if len(q) else None, _p(qo))
    if rc != 0:
        raise RuntimeError("ref_build_pbwt failed")
    return out, qo[:len(q)]


def ref_genotype_compare(codes, hap, hap_len):
    codes = np.ascontiguousarray(codes, dtype=np.uint8)
    hap = np.ascontiguousarray(hap, dtype=np.uint64)
    W = codes.shape[0]
    cap = 2 * unit_words(W) + 16
    words = np.zeros(cap, np.uint64)
    nerr = np.zeros(1, np.uint64)
    n = ref().ref_genotype_compare(_p(codes), W, _p(hap), hap_len, _p(words), cap, _p(nerr))
    if n < 0:
        raise RuntimeError("ref_genotype_compare failed")
    return words[:n], int(nerr[0])


def raw_rows(pbwt_rows, K):
    """The raw (FPGA handler) layout of compact rows: the data words of each
    row, padded with all-ones words to roundToMultiple(K,512)/32 words."""
    nw = (K + 31) // 32
    kr = (K + 511) // 512 * 16
    out = np.full((pbwt_rows.shape[0], kr), 0xFFFFFFFF, np.uint32)
    out[:, :nw] = pbwt_rows[:, 0:2 * nw:2]
    return out


def ref_ingest_raw(raw, K):
    """The reference's PBWT(uint32_t*, K, Kwords, M) on a raw buffer: (rows with
    inserted zero counts, count0)."""
    raw = np.ascontiguousarray(raw, np.uint32)
    M = raw.shape[0] // 2
    out = np.zeros((2 * M, 2 * raw.shape[1]), np.uint32)
    c0 = np.zeros(M, np.int32)
    if ref().ref_ingest_raw(_p(raw), K, raw.shape[1], M, _p(out), _p(c0)) != 0:
        raise RuntimeError("ref_ingest_raw failed")
    return out, c0


def ref_build_ref_pbwt(rows, K, absperm=True):
    """The reference's refPBWT over site rows (uint64, bit h = haplotype h):
    (rows of 2*ceil(K/32) uint32, count0, absolute permutations (M+1) x K)."""
    rows = np.ascontiguousarray(rows, np.uint64)
    M = rows.shape[0]
    out = np.zeros((M, 2 * ((K + 31) // 32)), np.uint32)
    c0 = np.zeros(M, np.int32)
    ap = np.zeros((M + 1, K), np.int32) if absperm else None
    if ref().ref_build_ref_pbwt(_p(rows), rows.shape[1], K, M, _p(out), _p(c0), _p(ap)) != 0:
        raise RuntimeError("ref_build_ref_pbwt failed")
    return out, c0, ap


def ref_phase_target(params, job, monos):
    """The reference's own beam search (src/Beam.cpp, src/BeamPath.cpp,
    src/HaplotypePath.cpp) driven as Target::phaseExtPBWT over the inputs of a
    phasing job (eagleimp_amd.phase.job); monos: the target's mono bits over
    the window.  Returns a PhaseResult."""
    monos = np.ascontiguousarray(monos, np.uint64)
    if ref().ref_phase_target(C.byref(params), C.byref(job.inp), _p(monos),
                              C.byref(job.out)) != 0:
        raise RuntimeError("ref_phase_target failed")
    return job.result()


def ref_rle_encode(words):
    """The reference's Qref run-length coder: bytes, or None when it gives up."""
    words = np.ascontiguousarray(words, np.uint64)
    out = np.zeros(words.size * 8 + 16, np.uint8)
    n = ref().ref_rle_encode(_p(words), words.size, _p(out), out.size)
    if n == -2:
        raise RuntimeError("ref_rle_encode: buffer too small")
    return None if n < 0 else out[:n].tobytes()


def ref_rle_decode(code: bytes, nwords: int):
    """The reference's Qref run-length decoder (no sample masking)."""
    buf = np.frombuffer(code, np.uint8).copy() if code else np.zeros(1, np.uint8)
    out = np.zeros(nwords, np.uint64)
    ref().ref_rle_decode(_p(buf), len(code), _p(out), nwords)
    return out


def ref_rle_decode_masked(code: bytes, nwords: int, masked, masked_haploid):
    """The reference's Qref run-length decoder with sample masking: (words,
    count of the masked samples' variant alleles)."""
    buf = np.frombuffer(code, np.uint8).copy() if code else np.zeros(1, np.uint8)
    ms = np.ascontiguousarray(masked, np.uint64)
    mh = np.ascontiguousarray(masked_haploid, np.uint8)
    out = np.zeros(max(nwords, 1), np.uint64)
    cnt = np.zeros(1, np.uint64)
    ref().ref_rle_decode_masked(_p(buf), len(code), _p(ms) if ms.size else None, ms.size,
                                _p(mh) if mh.size else None, _p(out), nwords, _p(cnt))
    return out[:nwords], int(cnt[0])


def build_ref_pbwt(rows, K, absperm=True):
    """Panel PBWT for imputation over site rows (uint64, bit h = haplotype h):
    (rows of 2*ceil(K/32) uint32, count0, absolute permutations (M+1) x K)."""
    rows = np.ascontiguousarray(rows, np.uint64)
    M = rows.shape[0]
    out = np.zeros((M, 2 * ((K + 31) // 32)), np.uint32)
    c0 = np.zeros(M, np.int32)
    ap = np.zeros((M + 1, K), np.int32) if absperm else None
    lib().eo_build_ref_pbwt(_p(rows), rows.shape[1], K, M, _p(out), _p(c0), _p(ap))
    return out, c0, ap
