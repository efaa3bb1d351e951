"""Qref reference panels (the reference's binary panel format) read into the
engine's resident site-major layout: ctypes binding of eimp_qref_* in
libeimp.so (include/eimp.h).

    with Qref(path) as q:
        rows = q.site_rows()                  # Mref x hap_words uint64
        eng.initialize(K, q.Mref, q.Ncorr, q.Nad, rows)
        eng.set_panel_extras(cM, haploid=q.haploid_flags())
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import _check, _p, lib

_P = C.c_void_p


class QrefInfo(C.Structure):
    _fields_ = [("v_major", C.c_int32), ("v_minor", C.c_int32), ("chrom", C.c_int32),
                ("has_haploid_flags", C.c_int32), ("Nref", C.c_uint64), ("NrefHap", C.c_uint64),
                ("Mref", C.c_uint64), ("MrefMA", C.c_uint64), ("hap_words", C.c_uint64)]


def _bind():
    L = lib()
    if not getattr(L, "_qref_bound", False):
        L.eimp_qref_open.argtypes = [C.c_char_p, C.POINTER(_P), C.POINTER(QrefInfo)]
        L.eimp_qref_close.argtypes = [_P]
        L.eimp_qref_metadata.argtypes = [_P, _P, _P, _P, _P]
        L.eimp_qref_variant.argtypes = [_P, C.c_size_t, C.POINTER(C.c_char_p),
                                        C.POINTER(C.c_char_p), C.POINTER(C.c_char_p)]
        L.eimp_qref_read_sites.argtypes = [_P, C.c_size_t, C.c_size_t, _P, C.c_size_t, C.c_int]
        L.eimp_qref_set_mask.argtypes = [_P, _P, C.c_size_t, C.POINTER(QrefInfo)]
        L.eimp_qref_parse_sample_list.argtypes = [C.c_char_p, _P, C.c_size_t,
                                                  C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        L.eimp_qref_mask_sample_list.argtypes = [_P, C.c_char_p, C.POINTER(QrefInfo)]
        L.eimp_qref_read_sites_af.argtypes = [_P, C.c_size_t, C.c_size_t, _P, C.c_size_t, _P, _P,
                                              C.c_int]
        L._qref_bound = True
    return L


class Qref:
    def __init__(self, path: str):
        self._h = _P()
        self.info = QrefInfo()
        _check(_bind().eimp_qref_open(str(path).encode(), C.byref(self._h), C.byref(self.info)))

    def close(self):
        if self._h:
            _bind().eimp_qref_close(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def Nref(self) -> int:
        return int(self.info.Nref)

    @property
    def Mref(self) -> int:
        return int(self.info.Mref)

    @property
    def Nad(self) -> int:
        """Haplotypes as the engine counts them (haploid samples keep two)."""
        return 2 * int(self.info.Nref)

    @property
    def Ncorr(self) -> int:
        """Haplotypes without the second haplotype of haploid samples."""
        return 2 * int(self.info.Nref) - int(self.info.NrefHap)

    def metadata(self) -> dict:
        M, N = self.Mref, self.Nref
        pos = np.zeros(M, np.int64)
        af = np.zeros(M, np.float32)
        hap = np.zeros(N, np.uint8)
        ma = np.zeros(max((M + 63) // 64, 1), np.uint64)
        _check(_bind().eimp_qref_metadata(self._h, _p(pos), _p(af), _p(hap), _p(ma)))
        ma_bits = np.unpackbits(ma.view(np.uint8), bitorder="little")[:M].astype(bool)
        return dict(positions=pos, af=af, haploid=hap, multiallelic=ma_bits)

    def haploid_flags(self) -> np.ndarray:
        return self.metadata()["haploid"]

    def variant(self, m: int):
        r, a, i = C.c_char_p(), C.c_char_p(), C.c_char_p()
        _check(_bind().eimp_qref_variant(self._h, m, C.byref(r), C.byref(a), C.byref(i)))
        return r.value.decode(), a.value.decode(), i.value.decode()

    def site_rows(self, m0: int = 0, n: int | None = None, row_words: int | None = None,
                  threads: int = 0) -> np.ndarray:
        """Haplotypes of variants [m0, m0+n) as site-major rows (bit h = allele
        of haplotype h); row_words defaults to the stored width (a multiple of
        8 words, as eimp_engine_initialize requires)."""
        n = self.Mref - m0 if n is None else n
        rw = int(self.info.hap_words) if row_words is None else row_words
        rows = np.zeros((n, rw), np.uint64)
        _check(_bind().eimp_qref_read_sites(self._h, m0, n, _p(rows), rw, threads))
        return rows

    def mask(self, samples) -> None:
        """Masks samples (strictly ascending indices of the file; empty clears
        the mask): rows, counts and haploid flags then cover the kept ones."""
        s = np.ascontiguousarray(samples, np.uint64)
        _check(_bind().eimp_qref_set_mask(self._h, _p(s) if s.size else None, s.size,
                                          C.byref(self.info)))

    def mask_sample_list(self, path: str) -> None:
        """Masks the withdrawals of a sample list (VCFData::parseSampleList)."""
        _check(_bind().eimp_qref_mask_sample_list(self._h, str(path).encode(),
                                                  C.byref(self.info)))

    def allele_freqs(self, m0: int = 0, n: int | None = None, threads: int = 0):
        """(allele frequency after masking, masked variant alleles) of
        variants [m0, m0+n)."""
        n = self.Mref - m0 if n is None else n
        af = np.zeros(max(n, 1), np.float32)
        cnt = np.zeros(max(n, 1), np.uint64)
        _check(_bind().eimp_qref_read_sites_af(self._h, m0, n, None, 0, _p(af), _p(cnt), threads))
        return af[:n], cnt[:n]


def parse_sample_list(path: str):
    """(withdrawal sample indices, samples listed) of a sample list."""
    L = _bind()
    nm, ns = C.c_size_t(), C.c_size_t()
    _check(L.eimp_qref_parse_sample_list(str(path).encode(), None, 0, C.byref(nm), C.byref(ns)))
    out = np.zeros(max(nm.value, 1), np.uint64)
    _check(L.eimp_qref_parse_sample_list(str(path).encode(), _p(out), nm.value, C.byref(nm),
                                         C.byref(ns)))
    return out[:nm.value], ns.value
