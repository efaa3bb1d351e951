"""Beam-search phasing over a target's compressed PBWT: the consumer of the
engine's output, on the host (eimp_phase_target / eimp_phase_batch in
libeimp.so, include/eimp.h).

The reference's GPU handler wraps each target's cPBWT as PBWT(cpbwt, count0,
K, 4*Kwords, Mp) and calls Target::phaseExtPBWT (src/GPUHandler.cpp:259-266,
src/Target.cpp:291-777); `phase_target` takes the same inputs (the engine's
copy_target output plus the target's genotypes) and returns the phased
haplotypes, dosages and call confidences.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import _check, _p, lib

_P = C.c_void_p


class PhaseParams(C.Structure):
    _fields_ = [("history", C.c_uint32), ("history_fast", C.c_uint32),
                ("beam_width", C.c_uint32), ("beam_width_fast", C.c_uint32),
                ("delta", C.c_uint32), ("delta_fast", C.c_uint32),
                ("hist", C.c_float), ("p_err", C.c_float), ("p_limit", C.c_float),
                ("expect_ibd_cM", C.c_float), ("min_fix_size", C.c_float),
                ("max_fix_size", C.c_float), ("min_fix_thresh", C.c_float),
                ("max_fix_thresh", C.c_float), ("pre_phasing", C.c_int32),
                ("rev_phasing", C.c_int32), ("last_iter", C.c_int32),
                ("imp_missing", C.c_int32)]


class PhaseInput(C.Structure):
    _fields_ = [("pbwt", _P), ("count0", _P), ("K", C.c_uint32), ("row_stride", C.c_uint32),
                ("Mp", C.c_uint32), ("W", C.c_uint32), ("gt", _P), ("splitsites", _P),
                ("cMdiffs", _P), ("af", _P), ("start_phase", C.c_int32),
                ("raw_rows", C.c_int32)]


class PhaseOutput(C.Structure):
    _fields_ = [("mat", _P), ("pat", _P), ("dosage_mat", _P), ("dosage_pat", _P),
                ("call_conf", _P), ("total_conf", C.c_float), ("ncalls", C.c_uint32),
                ("ncallsites", C.c_uint64), ("nmissing", C.c_int32), ("nmonhets", C.c_int32)]


def _bind():
    L = lib()
    if not getattr(L, "_phase_bound", False):
        L.eimp_phase_default_params.argtypes = [_P]
        L.eimp_phase_default_params.restype = None
        L.eimp_phase_target.argtypes = [_P, _P, _P]
        L.eimp_phase_batch.argtypes = [_P, C.c_size_t, _P, _P, C.c_int]
        L._phase_bound = True
    return L


def default_params(**overrides) -> PhaseParams:
    """The reference's defaults (src/Target.cpp:41-51, src/Args.cpp:276-278),
    with any field overridden by keyword."""
    p = PhaseParams()
    _bind().eimp_phase_default_params(C.byref(p))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise AttributeError(f"no phasing parameter {k!r}")
        setattr(p, k, v)
    return p


@dataclass
class PhaseResult:
    mat: np.ndarray         # W bools: allele of the maternal haplotype
    pat: np.ndarray         # W bools
    dosage_mat: np.ndarray  # W float32
    dosage_pat: np.ndarray  # W float32
    call_conf: np.ndarray   # ncalls float32 (last iteration)
    total_conf: float
    ncalls: int
    ncallsites: int
    nmissing: int
    nmonhets: int

    def same(self, o: "PhaseResult") -> bool:
        """Bit-for-bit equality (floats compared as bits)."""
        return (np.array_equal(self.mat, o.mat) and np.array_equal(self.pat, o.pat)
                and np.array_equal(self.dosage_mat.view(np.uint32), o.dosage_mat.view(np.uint32))
                and np.array_equal(self.dosage_pat.view(np.uint32), o.dosage_pat.view(np.uint32))
                and np.array_equal(self.call_conf.view(np.uint32), o.call_conf.view(np.uint32))
                and np.float32(self.total_conf).view(np.uint32)
                == np.float32(o.total_conf).view(np.uint32)
                and (self.ncalls, self.ncallsites, self.nmissing, self.nmonhets)
                == (o.ncalls, o.ncallsites, o.nmissing, o.nmonhets))


class _Job:
    """One target's ctypes input and output with the arrays they point to."""

    def __init__(self, pbwt, count0, K, gt, W, splitsites, cMdiffs, af=None, start_phase=0,
                 raw=False):
        pbwt = np.ascontiguousarray(pbwt, dtype=np.uint32)
        if pbwt.ndim != 2 or pbwt.shape[0] % 2:
            raise ValueError("pbwt must be 2*Mp rows")
        self.count0 = None if count0 is None else np.ascontiguousarray(count0, dtype=np.int32)
        self.gt = np.ascontiguousarray(gt, dtype=np.uint64)
        self.splitsites = np.ascontiguousarray(splitsites, dtype=np.uint32)
        self.cMdiffs = np.ascontiguousarray(cMdiffs, dtype=np.float32)
        self.af = None if af is None else np.ascontiguousarray(af, dtype=np.float32)
        self.pbwt = pbwt
        W = int(W)
        Mp = pbwt.shape[0] // 2
        ncalls_max = max(1, (Mp - 1) // 2)
        self.inp = PhaseInput(_p(pbwt), _p(self.count0), int(K), pbwt.shape[1], Mp, W,
                              _p(self.gt), _p(self.splitsites), _p(self.cMdiffs), _p(self.af),
                              int(start_phase), 1 if raw else 0)
        words = (W + 63) // 64
        self.mat = np.zeros(max(words, 1), np.uint64)
        self.pat = np.zeros(max(words, 1), np.uint64)
        self.dm = np.zeros(max(W, 1), np.float32)
        self.dp = np.zeros(max(W, 1), np.float32)
        self.cc = np.zeros(ncalls_max, np.float32)
        self.W = W
        self.out = PhaseOutput(_p(self.mat), _p(self.pat), _p(self.dm), _p(self.dp), _p(self.cc),
                               0.0, 0, 0, 0, 0)

    def result(self) -> PhaseResult:
        o = self.out
        W = self.W
        bits = lambda w: np.unpackbits(w.view(np.uint8), bitorder="little")[:W].astype(bool)  # noqa: E731
        return PhaseResult(bits(self.mat), bits(self.pat), self.dm[:W].copy(), self.dp[:W].copy(),
                           self.cc[:o.ncalls].copy(), float(o.total_conf), int(o.ncalls),
                           int(o.ncallsites), int(o.nmissing), int(o.nmonhets))


def job(pbwt, count0, K, gt, W, splitsites, cMdiffs, af=None, start_phase=0, raw=False) -> _Job:
    return _Job(pbwt, count0, K, gt, W, splitsites, cMdiffs, af, start_phase, raw)


def phase_target(params: PhaseParams, pbwt, count0, K, gt, W, splitsites, cMdiffs, af=None,
                 start_phase=0, raw=False) -> PhaseResult:
    """raw: pbwt rows hold data words only (the FPGA handler's raw layout)."""
    j = _Job(pbwt, count0, K, gt, W, splitsites, cMdiffs, af, start_phase, raw)
    _check(_bind().eimp_phase_target(C.byref(params), C.byref(j.inp), C.byref(j.out)))
    return j.result()


def phase_batch(params: PhaseParams, jobs: list, threads: int = 0) -> list:
    n = len(jobs)
    ins = (PhaseInput * n)(*[j.inp for j in jobs])
    outs = (PhaseOutput * n)(*[j.out for j in jobs])
    _check(_bind().eimp_phase_batch(C.byref(params), n, ins, outs, int(threads)))
    for j, o in zip(jobs, outs):
        j.out = o
    return [j.result() for j in jobs]


def engine_job(src, t: int, gt, W, af=None, start_phase=0) -> _Job:
    """Phasing job for target t of the last batch of an Engine or Stream (its
    outputs are copied from the device; the target's genotypes are the caller's)."""
    got = src.copy_target(t)
    K = int(got["besthaps"].shape[0])
    raw = getattr(src, "layout", getattr(src, "_last_layout", "gpu")) == "raw"
    return _Job(got["pbwt"], got["count0"], K, gt, W, got["splitsites"], got["cMdiffs"], af,
                start_phase, raw)
