"""K beyond the LDS-resident PBWT kernels (whole-panel K on panels wider
than 32,768 haplotypes, as the reference allows with K = 0 or a large K,
src/Args.cpp:432-433, src/PBWTPhaser.cpp:683): the order of such targets
lives in global memory.  Device vs the oracle, every layout, mixed with
targets that take the LDS-resident kernels in the same batch."""
import numpy as np
import pytest

import oracle
from eagleimp_amd import synth
from eagleimp_amd.engine import unit_words

from conftest import say
from parity import check_target, panel_engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wide():
    return synth.make_panel(41, N=33500, M=500, F=64, T=6)


@pytest.mark.parametrize("layout", ["gpu", "compact", "raw"])
def test_whole_panel_K_beyond_lds(built, wide, layout):
    panel = wide
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    # K = N (whole panel), K just below N (top-K), a window, and a small K in
    # the same batch (LDS-resident kernel)
    cases = [(0, 500, 33500), (0, 500, 33499), (37, 401, 33000), (0, 500, 5000), (100, 333, 40000)]
    gts = []
    for i, (s0, W, K) in enumerate(cases):
        gt = np.zeros(2 * unit_words(W), np.uint64)
        synth._lib().synth_target_gt(synth.ptr(panel.tgt_hap[2 * i]),
                                     synth.ptr(panel.tgt_hap[2 * i + 1]), s0, W, 0.01, 41, i,
                                     synth.ptr(gt))
        gts.append(gt)
    s0s = [c[0] for c in cases]
    Ws = [c[1] for c in cases]
    Ks = [c[2] for c in cases]
    say("large K", layout, "cases", cases)
    res = eng.run_batch(gts, s0s, Ws, Ks, digest=True, compact=layout == "compact",
                        raw=layout == "raw")
    for t in range(len(cases)):
        r = oracle.process_target(pv, gts[t], s0s[t], Ws[t], Ks[t])
        what = f"case {cases[t]}"
        assert int(res.Mp[t]) == r["Mp"] and int(res.K[t]) == r["K"], what
        got = eng.copy_target(t)
        check_target(r, got, int(r["K"]), layout, what, monos=eng.copy_monos(t))
        assert int(res.dig_pbwt[t]) == r["dig_pbwt"], f"{what}: pbwt digest"
        assert int(res.dig_count0[t]) == r["dig_count0"], f"{what}: count0 digest"
    eng.close()


def test_dropin_whole_panel_K_beyond_lds(built, wide):
    """The reference's runCrefPBWTKernel with K = N = 33,500: host prepPhase
    (oracle), device cref + PBWT, the drop-in layout."""
    panel = wide
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    eng = panel_engine(panel)
    K = panel.N
    eng.initialize(K, panel.M, panel.N, panel.N, panel.site)
    eng.set_panel_extras(panel.cM)
    tg = synth.make_targets(panel, 41, 2, K=K, miss=0.01)
    refs = [oracle.process_target(pv, tg.gt[t], 0, panel.M, K) for t in range(2)]
    offs = np.zeros(3, np.uint32)
    offs[1:] = np.cumsum([r["Mp"] for r in refs])
    gtfull = []
    for t in range(2):
        g = np.zeros(2 * unit_words(panel.M), np.uint64)
        g[:len(tg.gt[t])] = tg.gt[t]
        gtfull.append(g)
    say("drop-in large K", K)
    outs, c0s = eng.run_cref_pbwt(gtfull, [r["splits"] for r in refs], [r["flags"] for r in refs],
                                  offs)
    KW = oracle.pbwt_row_words(K)
    for t in range(2):
        rows = outs[t].reshape(2 * refs[t]["Mp"], -1)
        assert np.array_equal(rows[:, :KW], refs[t]["pbwt"]), f"target {t}: rows"
        assert np.array_equal(c0s[t], refs[t]["count0"]), f"target {t}: count0"
    eng.close()
