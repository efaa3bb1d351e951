"""eimp_engine_run_batch_noalloc (no device allocation: scratch in a
file-scope device array, inputs and outputs in the caller's device memory)
against eimp_engine_run_batch on the same targets: two calls in a row, the
second smaller, every result, row, count0 and small intermediate equal.
Device buffers come from the HIP runtime libeimp.so itself links."""
import ctypes as C

import numpy as np
import pytest

from eagleimp_amd import synth
from eagleimp_amd.engine import EimpError, lib, row_stride

from conftest import say
from parity import panel_engine

pytestmark = pytest.mark.gpu


def _hip():
    lib()  # libeimp.so first: it loads the HIP runtime this test must share
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class DevMem:
    """Device buffers for one test (freed at close)."""

    def __init__(self):
        self.h = _hip()
        self.ptrs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(int(nbytes), 256)) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def download(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        self.h.hipDeviceSynchronize()
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def _compare(eng, tg, idx, layout, mem, tag):
    """run_batch, then run_batch_noalloc on the same targets; equal outputs."""
    gts = [tg.gt[i] for i in idx]
    s0s, Ws, Ks = tg.s0[idx], tg.W[idx], tg.K[idx]
    kw = dict(compact=layout == "compact", raw=layout == "raw")
    say(tag, "targets", idx, "K", [int(k) for k in Ks], "W", [int(w) for w in Ws], layout)
    ref = eng.run_batch(gts, s0s, Ws, Ks, digest=True, **kw)
    want = [eng.copy_target(j) for j in range(len(idx))]
    max_Mp = int(ref.Mp.max()) + 3
    gt_dev = [mem.upload(g) for g in gts]
    Kc = [int(k) for k in ref.K]
    rows_dev = [mem.alloc(2 * max_Mp * row_stride(K, layout) * 4) for K in Kc]
    c0_dev = [mem.alloc(max_Mp * 4) for _ in idx]
    got = eng.run_batch_noalloc(gt_dev, s0s, Ws, Ks, max_Mp, rows_dev, c0_dev, digest=True, **kw)
    for f in ("S", "Mp", "ncalls", "K", "dig_pbwt", "dig_count0"):
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f"{tag}: {f}"
    for j in range(len(idx)):
        Mp, K = int(got.Mp[j]), Kc[j]
        rs = row_stride(K, layout)
        rows = mem.download(rows_dev[j], (2 * Mp, rs), np.uint32)
        c0 = mem.download(c0_dev[j], Mp, np.int32)
        assert np.array_equal(rows, want[j]["pbwt"]), f"{tag} target {j}: rows"
        assert np.array_equal(c0, want[j]["count0"]), f"{tag} target {j}: count0"
        small = eng.copy_target(j, rows=False)  # intermediates left in the scratch
        for f in ("besthaps", "bhflags", "splitsites", "splits", "count0"):
            assert np.array_equal(small[f], want[j][f]), f"{tag} target {j}: {f}"
        assert np.array_equal(small["cMdiffs"].view(np.uint32), want[j]["cMdiffs"].view(np.uint32))
    return max_Mp, gt_dev, rows_dev, c0_dev


@pytest.mark.parametrize("layout", ["gpu", "compact"])
def test_noalloc_equals_run_batch(built, layout):
    panel = synth.make_panel(31, N=320, M=5000, F=20, T=16)
    tg = synth.make_targets(panel, 31, 12, kset=[320, 100, 64, 257, 319], wmin=300, wmax=5000,
                            miss=0.0, miss_hi=0.05)
    eng = panel_engine(panel, give_hap=True)
    mem = DevMem()
    try:
        _compare(eng, tg, list(range(10)), layout, mem, "first call")
        # a second call, smaller: the scratch keeps the first call's contents
        max_Mp, gt_dev, rows_dev, c0_dev = _compare(eng, tg, [10, 11, 3], layout, mem,
                                                    "second call")
        # an output capacity below a target's Mp is refused, nothing allocated
        with pytest.raises(EimpError):
            eng.run_batch_noalloc(gt_dev, tg.s0[[10, 11, 3]], tg.W[[10, 11, 3]],
                                  tg.K[[10, 11, 3]], 1, rows_dev, c0_dev)
    finally:
        mem.close()
        eng.close()


def test_noalloc_needs_resident_hap_rows(built):
    """Subset targets need the hap-major rows resident: without them (no
    hapData and no earlier allocating batch) the call refuses instead of
    allocating them."""
    panel = synth.make_panel(32, N=200, M=2000, F=10, T=2)
    tg = synth.make_targets(panel, 32, 2, K=50, miss=0.01)
    eng = panel_engine(panel, give_hap=False)
    mem = DevMem()
    try:
        gt_dev = [mem.upload(g) for g in tg.gt]
        rows = [mem.alloc(2 * 4001 * row_stride(50, "gpu") * 4) for _ in range(2)]
        c0 = [mem.alloc(4001 * 4) for _ in range(2)]
        with pytest.raises(EimpError):
            eng.run_batch_noalloc(gt_dev, tg.s0, tg.W, tg.K, 4001, rows, c0)
    finally:
        mem.close()
        eng.close()
