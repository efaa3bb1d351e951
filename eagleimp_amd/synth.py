"""Deterministic synthetic workloads (panels, targets) for the phasing pre-stage.

The heavy lifting is in ``csrc/synth.c`` (``libeimp_synth.so``): founder
mosaics with Beta-distributed allele frequencies, Poisson switch points along
a linear genetic map and random allele flips.  All streams derive from
``(seed, tag, index)``, so the same bytes come out on every host and for any
thread count.

Named workloads (``WORKLOADS``) restate the benchmark configurations: panel
size, founders, MAF shape, targets, per-target windows, K and missingness.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libeimp_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run __graft_entry__.build() first")
        lib = C.CDLL(path)
        P = C.c_void_p
        lib.synth_panel.argtypes = [P, C.c_uint64, P, C.c_uint64, P, C.c_uint64, P, P, P, C.c_int]
        lib.synth_panel.restype = C.c_int
        lib.synth_transpose.argtypes = [P, C.c_uint64, C.c_uint64, C.c_uint64, P, C.c_uint64, C.c_int]
        lib.synth_transpose.restype = C.c_int
        lib.synth_target_gt.argtypes = [P, P, C.c_uint64, C.c_uint64, C.c_double, C.c_uint64,
                                        C.c_uint64, P]
        lib.synth_target_gt.restype = C.c_int
        lib.synth_target_params.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64,
                                            C.c_uint64, P, C.c_uint64, C.c_double, C.c_double,
                                            P, P, P, P]
        lib.synth_target_params.restype = C.c_int
        lib.synth_arrivals.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint64,
                                       C.c_uint64, C.c_uint64, P, P, P]
        lib.synth_arrivals.restype = C.c_int
        _LIB = lib
    return _LIB


class _Params(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("N", C.c_uint64), ("M", C.c_uint64), ("F", C.c_uint64),
                ("beta_a", C.c_double), ("maf_lo", C.c_double), ("maf_hi", C.c_double),
                ("bp_per_site", C.c_int64), ("switch_per_cM", C.c_double),
                ("flip_rate", C.c_double)]


def unit_words(nbits: int) -> int:
    """uint64 words of a bit vector padded to 512-bit units (src/Datatypes.h:70)."""
    return ((nbits + 511) // 512) * 8


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None else None


@dataclass
class Panel:
    """A reference panel in both of EagleImp's layouts.

    hap:  (N, hap_stride) uint64, bit m of row h = allele of haplotype h at site m
          (``reference``, src/VCFData.h:203)
    site: (M, site_stride) uint64, bit h of row m = allele of haplotype h at site m
          (``referenceT``, src/VCFData.h:205)
    cM:   (M,) float32 genetic positions
    """
    N: int
    M: int
    hap: np.ndarray
    site: np.ndarray
    cM: np.ndarray
    tgt_hap: np.ndarray  # (2T, hap_stride) held-out haplotypes for targets

    @property
    def hap_stride(self) -> int:
        return self.hap.shape[1]

    @property
    def site_stride(self) -> int:
        return self.site.shape[1]


@dataclass
class Targets:
    s0: np.ndarray     # (T,) uint64 window start
    W: np.ndarray      # (T,) uint64 window length
    K: np.ndarray      # (T,) uint32 conditioning haplotypes requested
    miss: np.ndarray   # (T,) float64 missing rate
    gt: list = field(default_factory=list)  # per target: (2*unit_words(W),) uint64

    def __len__(self):
        return len(self.s0)


def make_panel(seed: int, N: int, M: int, F: int, T: int, beta_a: float = 0.5,
               maf_lo: float = 0.005, maf_hi: float = 0.5, bp_per_site: int = 1000,
               switch_per_cM: float = 1.0, flip_rate: float = 1e-3,
               nthreads: int = 0, with_site: bool = True) -> Panel:
    p = _Params(seed, N, M, F, beta_a, maf_lo, maf_hi, bp_per_site, switch_per_cM, flip_rate)
    hs, ss = unit_words(M), unit_words(N)
    hap = np.zeros((N, hs), dtype=np.uint64)
    site = np.zeros((M, ss), dtype=np.uint64) if with_site else None
    tgt = np.zeros((max(2 * T, 1), hs), dtype=np.uint64)
    cM = np.zeros(M, dtype=np.float32)
    bp = np.zeros(M, dtype=np.int32)
    rc = _lib().synth_panel(C.byref(p), T, ptr(hap), hs, ptr(site), ss, ptr(tgt), ptr(cM),
                            ptr(bp), nthreads)
    if rc != 0:
        raise RuntimeError(f"synth_panel failed ({rc})")
    return Panel(N=N, M=M, hap=hap, site=site, cM=cM, tgt_hap=tgt)


def make_targets(panel: Panel, seed: int, T: int, *, K=None, kset=None, wmin=None, wmax=None,
                 miss=0.01, miss_hi=None) -> Targets:
    """Per-target windows/K/missingness and their genotype vectors.

    Full-window targets: ``wmin = wmax = M`` (the default).  ``K`` fixes K for
    all targets, ``kset`` draws it uniformly from a set.
    """
    M = panel.M
    wmin = M if wmin is None else wmin
    wmax = M if wmax is None else wmax
    ks = np.asarray(kset if kset is not None else [K if K is not None else 0], dtype=np.uint32)
    s0 = np.zeros(T, dtype=np.uint64)
    W = np.zeros(T, dtype=np.uint64)
    Ks = np.zeros(T, dtype=np.uint32)
    ms = np.zeros(T, dtype=np.float64)
    mhi = miss if miss_hi is None else miss_hi
    _lib().synth_target_params(seed, T, M, wmin, wmax, ptr(ks), len(ks), miss, mhi,
                               ptr(s0), ptr(W), ptr(Ks), ptr(ms))
    gts = []
    hs = panel.hap_stride
    for t in range(T):
        w = int(W[t])
        gt = np.zeros(2 * unit_words(w), dtype=np.uint64)
        h1 = panel.tgt_hap[2 * t]
        h2 = panel.tgt_hap[2 * t + 1]
        assert h1.shape[0] == hs
        _lib().synth_target_gt(ptr(h1), ptr(h2), int(s0[t]), w, float(ms[t]), seed, t, ptr(gt))
        gts.append(gt)
    return Targets(s0=s0, W=W, K=Ks, miss=ms, gt=gts)


def make_arrivals(seed: int, n: int, rate: float, npool: int, bmin: int = 1, bmax: int = 256):
    """Serving workload: (arrival seconds, batch sizes, first pool index) of n
    request batches, a Poisson process of `rate` batches/s with log-uniform
    sizes on [bmin, bmax]."""
    t = np.zeros(n, np.float64)
    size = np.zeros(n, np.uint32)
    first = np.zeros(n, np.uint32)
    rc = _lib().synth_arrivals(seed, n, float(rate), bmin, bmax, npool, ptr(t), ptr(size), ptr(first))
    if rc != 0:
        raise ValueError("bad arrival parameters")
    return t, size, first


# Benchmark workloads.  Site spacing is 1 site/kb with a linear 1 cM/Mb map
# for every panel; the seed of each workload is its number.
WORKLOADS = {
    1: dict(name="cpu-2kx10k-K1000", N=2000, M=10000, F=50, beta_a=0.5, T=20,
            K=1000, miss=0.01),
    2: dict(name="plain-10kx100k-K10000", N=10000, M=100000, F=200, beta_a=0.5, T=500,
            K=10000, miss=0.01),
    3: dict(name="batched-20kx200k-window-K2k..10k", N=20000, M=200000, F=200, beta_a=0.5,
            T=5000, kset=[2000, 4000, 6000, 8000, 10000], wmin=20000, wmax=200000,
            miss=0.0, miss_hi=0.05),
    4: dict(name="large-200kx500k-K10000", N=200000, M=500000, F=2000, beta_a=0.3,
            T=10000, K=10000, miss=0.01),
    # serving: the 20k x 200k panel (panel seed 3) resident, a pool of targets
    # with the windows / K / missingness of the batched workload, request
    # batches of 1-256 targets arriving at 20-200 batches/s
    5: dict(name="serving-20kx200k-poisson", panel_seed=3, N=20000, M=200000, F=200,
            beta_a=0.5, T=1024, kset=[2000, 4000, 6000, 8000, 10000], wmin=20000,
            wmax=200000, miss=0.0, miss_hi=0.05, rates=(20.0, 200.0), bmin=1, bmax=256),
}


def make_workload(cfg: int, T: int | None = None, nthreads: int = 0):
    c = dict(WORKLOADS[cfg])
    T = c["T"] if T is None else T
    panel = make_panel(c.get("panel_seed", cfg), c["N"], c["M"], c["F"], T, beta_a=c["beta_a"],
                       nthreads=nthreads)
    tg = make_targets(panel, cfg, T, K=c.get("K"), kset=c.get("kset"), wmin=c.get("wmin"),
                      wmax=c.get("wmax"), miss=c["miss"], miss_hi=c.get("miss_hi"))
    return c, panel, tg
