"""ctypes binding of the engine's C ABI (include/eimp.h, libeimp.so).

Plumbing for tests, the benchmark and the serving loop; the work happens in
the HIP kernels behind the C ABI.  There is no fallback: if libeimp.so is
missing or no device is visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeimp.so")

EIMP_BATCH_DIGEST = 1
EIMP_BATCH_RAW = 2
EIMP_BATCH_COMPACT = 4
EIMP_STREAM_NO_GRAPH = 8
EIMP_STREAM_TIMED = 16
EIMP_BUSY = 1
KCLASSES = ["distance", "select", "monos", "splits", "cref", "pbwt", "copy"]

# every entry point declared in include/eimp.h
EXPORTS = [
    "eimp_last_error", "eimp_device_count", "eimp_engine_create", "eimp_engine_destroy",
    "eimp_engine_initialize", "eimp_engine_block_size", "eimp_engine_run_cref_pbwt",
    "eimp_engine_free", "eimp_engine_device_index", "eimp_engine_set_panel_extras",
    "eimp_engine_run_batch", "eimp_engine_copy_target", "eimp_engine_set_timing",
    "eimp_engine_get_timing", "eimp_engine_reset_timing", "eimp_engine_set_pbwt_variant",
    "eimp_pbwt_variant_count", "eimp_pbwt_variant_info", "eimp_stream_create",
    "eimp_stream_create_shared",
    "eimp_stream_submit", "eimp_stream_query", "eimp_stream_wait", "eimp_stream_copy_target",
    "eimp_stream_destroy", "eimp_engine_mem_info", "eimp_stream_kernel_ms",
    "eimp_engine_ref_pbwt", "eimp_phase_default_params", "eimp_phase_target", "eimp_phase_batch",
    "eimp_qref_open", "eimp_qref_close", "eimp_qref_metadata", "eimp_qref_variant",
    "eimp_qref_read_sites", "eimp_qref_set_mask", "eimp_qref_parse_sample_list",
    "eimp_qref_mask_sample_list", "eimp_qref_read_sites_af", "eimp_engine_copy_monos", "eimp_stream_copy_monos",
    "eimp_stream_digest_outputs", "eimp_engine_run_batch_noalloc",
]

_P = C.c_void_p
_SZ = C.c_size_t


class EimpTarget(C.Structure):
    _fields_ = [("gt", _P), ("s0", C.c_uint64), ("W", C.c_uint64), ("K", C.c_uint64),
                ("ignoreSample", C.c_int64)]


class EimpResult(C.Structure):
    _fields_ = [("S", C.c_uint64), ("Mp", C.c_uint64), ("ncalls", C.c_uint64), ("K", C.c_uint64),
                ("digest_pbwt", C.c_uint64), ("digest_count0", C.c_uint64)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.eimp_last_error.restype = C.c_char_p
        L.eimp_device_count.argtypes = [C.POINTER(C.c_int)]
        L.eimp_engine_create.argtypes = [C.c_int, C.POINTER(_P)]
        L.eimp_engine_destroy.argtypes = [_P]
        L.eimp_engine_initialize.argtypes = [_P, _SZ, _SZ, _SZ, _SZ, _SZ, _SZ, _P]
        L.eimp_engine_block_size.argtypes = [_P, C.POINTER(_SZ)]
        L.eimp_engine_run_cref_pbwt.argtypes = [_P, _SZ, _P, _P, _P, _P, _P, _P]
        L.eimp_engine_free.argtypes = [_P]
        L.eimp_engine_device_index.argtypes = [_P, C.POINTER(C.c_int)]
        L.eimp_engine_set_panel_extras.argtypes = [_P, _P, _P, _SZ, _P]
        L.eimp_engine_run_batch.argtypes = [_P, _SZ, _P, C.c_float, C.c_uint32, _P]
        L.eimp_engine_copy_target.argtypes = [_P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P]
        L.eimp_engine_set_timing.argtypes = [_P, C.c_int]
        L.eimp_engine_get_timing.argtypes = [_P, _P, _P]
        L.eimp_engine_reset_timing.argtypes = [_P]
        L.eimp_engine_set_pbwt_variant.argtypes = [_P, C.c_int]
        L.eimp_pbwt_variant_count.argtypes = []
        L.eimp_pbwt_variant_info.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                             C.POINTER(C.c_int)]
        L.eimp_stream_create.argtypes = [_P, _SZ, _SZ, _SZ, _SZ, C.c_uint32, C.POINTER(_P)]
        L.eimp_stream_create_shared.argtypes = [_P, _SZ, _SZ, _SZ, C.c_uint32, C.POINTER(_P)]
        L.eimp_stream_submit.argtypes = [_P, _SZ, _P, C.c_float]
        L.eimp_stream_query.argtypes = [_P]
        L.eimp_stream_wait.argtypes = [_P, _P, C.POINTER(_SZ), _P]
        L.eimp_stream_copy_target.argtypes = [_P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P]
        L.eimp_stream_destroy.argtypes = [_P]
        L.eimp_engine_mem_info.argtypes = [_P, C.POINTER(_SZ), C.POINTER(_SZ)]
        L.eimp_stream_kernel_ms.argtypes = [_P, _P]
        L.eimp_engine_ref_pbwt.argtypes = [_P, _SZ, _SZ, _P, _P, _P, _P]
        L.eimp_engine_copy_monos.argtypes = [_P, _SZ, _P]
        L.eimp_stream_copy_monos.argtypes = [_P, _SZ, _P]
        L.eimp_stream_digest_outputs.argtypes = [_P, _P]
        L.eimp_engine_run_batch_noalloc.argtypes = [_P, _SZ, _P, C.c_float, C.c_uint32, _SZ, _P,
                                                    _P, _P]
        _LIB = L
    return _LIB


class EimpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eimp error {code}: {msg}")
        self.code = code


def _check(rc):
    if rc != 0:
        raise EimpError(rc, lib().eimp_last_error().decode(errors="replace"))


def _p(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().eimp_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def unit_words(nbits: int) -> int:
    return ((nbits + 511) // 512) * 8


def pbwt_variants():
    out = []
    for v in range(lib().eimp_pbwt_variant_count()):
        nt, cpw, mk = C.c_int(), C.c_int(), C.c_int()
        _check(lib().eimp_pbwt_variant_info(v, C.byref(nt), C.byref(cpw), C.byref(mk)))
        out.append(dict(variant=v, threads=nt.value, chunks_per_wave=cpw.value, max_k=mk.value))
    return out


@dataclass
class BatchResult:
    S: np.ndarray
    Mp: np.ndarray
    ncalls: np.ndarray
    K: np.ndarray
    dig_pbwt: np.ndarray
    dig_count0: np.ndarray


class Engine:
    """One device, resident panel, batched pre-stage (see include/eimp.h)."""

    def __init__(self, device: int = 0):
        self._h = _P()
        _check(lib().eimp_engine_create(device, C.byref(self._h)))
        self.device = device
        self.K = self.M = self.Ncorr = self.Nad = None

    def close(self):
        if self._h:
            lib().eimp_engine_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- reference GPUEngine API
    def initialize(self, K, M, Ncorr, Nad, site: np.ndarray):
        site = np.ascontiguousarray(site, dtype=np.uint64)
        assert site.shape[0] == M
        self._site = site
        _check(lib().eimp_engine_initialize(self._h, K, unit_words(K), M, Ncorr, Nad,
                                            site.shape[1] * 8, _p(site)))
        self.K, self.M, self.Ncorr, self.Nad = K, M, Ncorr, Nad
        self.refTCap = site.shape[1] * 8

    def block_size(self) -> int:
        v = _SZ()
        _check(lib().eimp_engine_block_size(self._h, C.byref(v)))
        return v.value

    def free(self):
        _check(lib().eimp_engine_free(self._h))

    def set_panel_extras(self, cM, hap=None, haploid=None):
        cM = np.ascontiguousarray(cM, dtype=np.float32)
        hap = None if hap is None else np.ascontiguousarray(hap, dtype=np.uint64)
        hp = None if haploid is None else np.ascontiguousarray(haploid, dtype=np.uint8)
        _check(lib().eimp_engine_set_panel_extras(self._h, _p(cM), _p(hap),
                                                  0 if hap is None else hap.shape[1], _p(hp)))

    def run_cref_pbwt(self, gts, splits, bhflags, offsets):
        """Drop-in GPUEngine::runCrefPBWTKernel; returns (cPBWT list, count0 list)."""
        n = len(gts)
        gts = [np.ascontiguousarray(g, dtype=np.uint64) for g in gts]
        splits = [np.ascontiguousarray(s, dtype=np.uint64) for s in splits]
        bhflags = [np.ascontiguousarray(b, dtype=np.uint64) for b in bhflags]
        offs = np.ascontiguousarray(offsets, dtype=np.uint32)
        kw = unit_words(self.K)
        outs, c0s = [], []
        for t in range(n):
            Mp = int(offs[t + 1] - offs[t])
            outs.append(np.zeros(8 * Mp * kw, dtype=np.uint32))
            c0s.append(np.zeros(Mp, dtype=np.int32))
        A = C.c_void_p * n
        _check(lib().eimp_engine_run_cref_pbwt(
            self._h, n, A(*[g.ctypes.data for g in gts]), A(*[s.ctypes.data for s in splits]),
            A(*[b.ctypes.data for b in bhflags]), _p(offs), A(*[o.ctypes.data for o in outs]),
            A(*[c.ctypes.data for c in c0s])))
        return outs, c0s

    # -- batched device pre-stage
    def run_batch(self, gts, s0s, Ws, Ks, ignore=None, cMmaxSplit=0.125, digest=True,
                  compact=False, raw=False) -> BatchResult:
        """compact: rows of 2*ceil(K/32) (CPU layout); raw: data words only in
        rows of roundToMultiple(K,512)/32 (FPGA handler layout); default: the
        drop-in GPU layout."""
        n = len(gts)
        self._gts = [np.ascontiguousarray(g, dtype=np.uint64) for g in gts]
        ts = (EimpTarget * n)()
        for i in range(n):
            ts[i] = EimpTarget(self._gts[i].ctypes.data, int(s0s[i]), int(Ws[i]), int(Ks[i]),
                               -1 if ignore is None else int(ignore[i]))
        rs = (EimpResult * n)()
        flags = ((EIMP_BATCH_DIGEST if digest else 0) | (EIMP_BATCH_COMPACT if compact else 0)
                 | (EIMP_BATCH_RAW if raw else 0))
        _check(lib().eimp_engine_run_batch(self._h, n, ts, C.c_float(cMmaxSplit), flags, rs))
        arr = np.frombuffer(rs, dtype=np.uint64).reshape(n, 6)
        self._last = arr.copy()
        self._last_layout = layout_name(compact, raw)
        self._last_W = [int(w) for w in Ws]
        return BatchResult(S=arr[:, 0].copy(), Mp=arr[:, 1].copy(), ncalls=arr[:, 2].copy(),
                           K=arr[:, 3].copy(), dig_pbwt=arr[:, 4].copy(),
                           dig_count0=arr[:, 5].copy())

    def run_batch_noalloc(self, gt_dev, s0s, Ws, Ks, max_Mp, rows_dev, count0_dev,
                          ignore=None, cMmaxSplit=0.125, digest=True, compact=False,
                          raw=False) -> BatchResult:
        """eimp_engine_run_batch_noalloc: device addresses (ints) of each
        target's genotype words, output rows (2*max_Mp*rowstride uint32) and
        count0 (max_Mp int32); nothing is allocated by the call."""
        n = len(gt_dev)
        ts = (EimpTarget * n)()
        for i in range(n):
            ts[i] = EimpTarget(int(gt_dev[i]), int(s0s[i]), int(Ws[i]), int(Ks[i]),
                               -1 if ignore is None else int(ignore[i]))
        rs = (EimpResult * n)()
        A = C.c_void_p * n
        flags = ((EIMP_BATCH_DIGEST if digest else 0) | (EIMP_BATCH_COMPACT if compact else 0)
                 | (EIMP_BATCH_RAW if raw else 0))
        _check(lib().eimp_engine_run_batch_noalloc(
            self._h, n, ts, C.c_float(cMmaxSplit), flags, int(max_Mp),
            A(*[int(p) for p in rows_dev]), A(*[int(p) for p in count0_dev]), rs))
        arr = np.frombuffer(rs, dtype=np.uint64).reshape(n, 6)
        self._last = arr.copy()
        self._last_layout = layout_name(compact, raw)
        self._last_W = [int(w) for w in Ws]
        return BatchResult(S=arr[:, 0].copy(), Mp=arr[:, 1].copy(), ncalls=arr[:, 2].copy(),
                           K=arr[:, 3].copy(), dig_pbwt=arr[:, 4].copy(),
                           dig_count0=arr[:, 5].copy())

    def copy_target(self, t, cref=False, rows=True) -> dict:
        """Outputs of target t of the last batch; rows=False leaves out the
        cPBWT rows (the largest part)."""
        S, Mp, _, K = (int(x) for x in self._last[t, :4])
        return _copy_out(lib().eimp_engine_copy_target, self._h, t, S, Mp, K, self._last_W[t],
                         self._last_layout, self.Nad, cref, rows)

    def copy_monos(self, t) -> np.ndarray:
        """Mono bits of target t of the last batch (window-relative words)."""
        out = np.zeros(unit_words(self._last_W[t]), np.uint64)
        _check(lib().eimp_engine_copy_monos(self._h, t, _p(out)))
        return out

    def ref_pbwt(self, m0=0, nsites=None, absperm=True):
        """Panel PBWT for imputation over sites [m0, m0+nsites): (rows, count0,
        absperm or None, device ms)."""
        nsites = self.M - m0 if nsites is None else nsites
        N = self.Nad
        rows = np.zeros((nsites, 2 * ((N + 31) // 32)), np.uint32)
        c0 = np.zeros(nsites, np.int32)
        ap = np.zeros((nsites + 1, N), np.int32) if absperm else None
        ms = np.zeros(1, np.float64)
        _check(lib().eimp_engine_ref_pbwt(self._h, m0, nsites, _p(rows), _p(c0), _p(ap), _p(ms)))
        return rows, c0, ap, float(ms[0])

    def mem_info(self):
        """(free, total) device bytes."""
        f, t = _SZ(), _SZ()
        _check(lib().eimp_engine_mem_info(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def stream(self, max_targets, max_window, max_K, pool_bytes, digest=True, compact=False,
               graph=True, timed=False, raw=False, share=None) -> "Stream":
        return Stream(self, max_targets, max_window, max_K, pool_bytes, digest, compact, graph,
                      timed, raw, share)

    # -- measurement
    def set_timing(self, on=True):
        _check(lib().eimp_engine_set_timing(self._h, 1 if on else 0))

    def reset_timing(self):
        _check(lib().eimp_engine_reset_timing(self._h))

    def timing(self) -> dict:
        ms = np.zeros(len(KCLASSES), np.float64)
        n = np.zeros(len(KCLASSES), np.uint64)
        _check(lib().eimp_engine_get_timing(self._h, _p(ms), _p(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KCLASSES)}

    def set_pbwt_variant(self, v: int):
        _check(lib().eimp_engine_set_pbwt_variant(self._h, v))


def layout_name(compact=False, raw=False) -> str:
    if compact and raw:
        raise ValueError("compact and raw layouts exclude each other")
    return "compact" if compact else ("raw" if raw else "gpu")


def row_stride(K: int, layout: str) -> int:
    """uint32 words per stored PBWT row in an output layout."""
    if layout == "compact":
        return 2 * ((K + 31) // 32)
    if layout == "raw":
        return (K + 511) // 512 * 16
    return 4 * unit_words(K)


def _copy_out(fn, h, t, S, Mp, K, W, layout, Nad, cref, rows=True) -> dict:
    rs = row_stride(K, layout)
    out = dict(
        besthaps=np.zeros(K, np.uint32),
        bhflags=np.zeros(unit_words(Nad), np.uint64),
        splitsites=np.zeros(max(S, 1), np.uint32),
        splits=np.zeros(unit_words(W), np.uint64),
        cMdiffs=np.zeros(S + 1 if S else 1, np.float32),
        count0=np.zeros(Mp, np.int32),
        pbwt=np.zeros(2 * Mp * rs if rows else 0, np.uint32),
    )
    cr = np.zeros(Mp * unit_words(K), np.uint64) if cref else None
    _check(fn(h, t, _p(out["besthaps"]), _p(out["bhflags"]), _p(out["splitsites"]),
              _p(out["splits"]), _p(out["cMdiffs"]), _p(out["count0"]),
              _p(out["pbwt"]) if rows else None, _p(cr)))
    out["splitsites"] = out["splitsites"][:S]
    out["cMdiffs"] = out["cMdiffs"][:S + 1] if S else out["cMdiffs"][:0]
    out["pbwt"] = out["pbwt"].reshape(2 * Mp if rows else 0, rs)
    if cref:
        out["cref"] = cr.reshape(Mp, unit_words(K))
    return out


class Stream:
    """Asynchronous device-driven batches on their own HIP stream (eimp_stream_*).

    submit() returns at once; wait() collects the per-target summaries.  The
    launch sequence is a hipGraph captured at creation (graph=False launches
    the same kernels one by one; timed=True also records events between the
    kernel classes, read with kernel_ms())."""

    def __init__(self, eng: Engine, max_targets, max_window, max_K, pool_bytes, digest=True,
                 compact=False, graph=True, timed=False, raw=False, share=None):
        """share: another Stream whose output pool this one uses (pool_bytes
        is then ignored); their builds take turns, each preparation runs
        beside the other's build."""
        self.eng = eng
        self._h = _P()
        self.layout = layout_name(compact, raw)
        self.digest = digest
        flags = ((EIMP_BATCH_DIGEST if digest else 0) | (EIMP_BATCH_COMPACT if compact else 0)
                 | (EIMP_BATCH_RAW if raw else 0)
                 | (0 if graph else EIMP_STREAM_NO_GRAPH) | (EIMP_STREAM_TIMED if timed else 0))
        if share is not None:
            _check(lib().eimp_stream_create_shared(share._h, max_targets, max_window, max_K,
                                                   flags, C.byref(self._h)))
        else:
            _check(lib().eimp_stream_create(eng._h, max_targets, max_window, max_K, pool_bytes,
                                            flags, C.byref(self._h)))
        self.max_targets = max_targets
        self.timed = timed
        self._n = 0
        self._W = []

    def close(self):
        if self._h:
            lib().eimp_stream_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, gts, s0s, Ws, Ks, ignore=None, cMmaxSplit=0.125):
        n = len(gts)
        gts = [np.ascontiguousarray(g, dtype=np.uint64) for g in gts]
        ts = (EimpTarget * max(n, 1))()
        for i in range(n):
            ts[i] = EimpTarget(gts[i].ctypes.data, int(s0s[i]), int(Ws[i]), int(Ks[i]),
                               -1 if ignore is None else int(ignore[i]))
        _check(lib().eimp_stream_submit(self._h, n, ts, C.c_float(cMmaxSplit)))
        self._n = n
        self._W = [int(w) for w in Ws]

    def done(self) -> bool:
        rc = lib().eimp_stream_query(self._h)
        if rc == EIMP_BUSY:
            return False
        _check(rc)
        return True

    def wait(self):
        """Returns (BatchResult, ndeferred, (kernel_ms, submission_ms))."""
        n = self._n
        rs = (EimpResult * max(n, 1))()
        nd = _SZ()
        tm = np.zeros(2, np.float64)
        _check(lib().eimp_stream_wait(self._h, rs, C.byref(nd), _p(tm)))
        arr = np.frombuffer(rs, dtype=np.uint64).reshape(-1, 6)[:n]
        self._last = arr.copy()
        res = BatchResult(S=arr[:, 0].copy(), Mp=arr[:, 1].copy(), ncalls=arr[:, 2].copy(),
                          K=arr[:, 3].copy(), dig_pbwt=arr[:, 4].copy(),
                          dig_count0=arr[:, 5].copy())
        return res, nd.value, (float(tm[0]), float(tm[1]))

    def kernel_ms(self) -> dict:
        ms = np.zeros(len(KCLASSES), np.float64)
        _check(lib().eimp_stream_kernel_ms(self._h, _p(ms)))
        return {k: float(ms[i]) for i, k in enumerate(KCLASSES)}

    def digest_outputs(self):
        """(pbwt digests, count0 digests) recomputed from the rows the last
        collected submission left in the pool (0 for deferred targets)."""
        d = np.zeros((max(self._n, 1), 2), np.uint64)
        _check(lib().eimp_stream_digest_outputs(self._h, _p(d)))
        return d[:self._n, 0].copy(), d[:self._n, 1].copy()

    def copy_monos(self, t) -> np.ndarray:
        out = np.zeros(unit_words(self._W[t]), np.uint64)
        _check(lib().eimp_stream_copy_monos(self._h, t, _p(out)))
        return out

    def copy_target(self, t, cref=False, rows=True) -> dict:
        S, Mp, _, K = (int(x) for x in self._last[t, :4])
        return _copy_out(lib().eimp_stream_copy_target, self._h, t, S, Mp, K, self._W[t],
                         self.layout, self.eng.Nad, cref, rows)
