"""The synthetic generator is deterministic and its layouts agree."""
import numpy as np

from eagleimp_amd import synth


def test_deterministic_across_thread_counts(built):
    a = synth.make_panel(9, N=130, M=1000, F=7, T=3, nthreads=1)
    b = synth.make_panel(9, N=130, M=1000, F=7, T=3, nthreads=4)
    assert np.array_equal(a.hap, b.hap) and np.array_equal(a.site, b.site)
    assert np.array_equal(a.tgt_hap, b.tgt_hap) and np.array_equal(a.cM, b.cM)
    c = synth.make_panel(10, N=130, M=1000, F=7, T=3)
    assert not np.array_equal(a.hap, c.hap)


def test_layouts_are_transposes(built):
    p = synth.make_panel(2, N=200, M=777, F=9, T=1)
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)
    S = np.unpackbits(p.site.view(np.uint8), bitorder="little").reshape(p.M, -1)
    assert np.array_equal(H[:, :p.M].T, S[:, :p.N])
    assert not H[:, p.M:].any() and not S[:, p.N:].any()  # zero padding


def test_linear_map_and_spacing(built):
    p = synth.make_panel(2, N=10, M=500, F=3, T=1)
    assert p.cM.dtype == np.float32
    assert np.all(np.diff(p.cM) > 0)
    assert abs(float(p.cM[-1]) - 0.5) < 1e-4  # 500 sites at 1 kb, 1 cM/Mb


def test_mosaic_structure(built):
    """Panel haplotypes copy founders: neighbouring sites agree far more often
    than independent alleles would."""
    p = synth.make_panel(4, N=400, M=3000, F=8, T=1)
    H = np.unpackbits(p.hap.view(np.uint8), bitorder="little").reshape(p.N, -1)[:, :p.M]
    # number of distinct haplotypes over a short stretch is bounded by founders (+ flips)
    stretch = H[:, 1000:1040]
    distinct = len({r.tobytes() for r in stretch})
    assert distinct < 60


def test_target_params(built):
    p = synth.make_panel(3, N=50, M=5000, F=4, T=40)
    tg = synth.make_targets(p, 3, 40, kset=[5, 10], wmin=100, wmax=4000, miss=0.0, miss_hi=0.05)
    assert np.all(tg.W >= 100) and np.all(tg.W <= 4000)
    assert np.all(tg.s0 + tg.W <= p.M)
    assert set(tg.K.tolist()) <= {5, 10}
    assert np.all((tg.miss >= 0) & (tg.miss <= 0.05))
