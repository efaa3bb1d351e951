"""A panel loaded from a Qref file feeds the resident-panel engine: targets
give the same outputs as with the panel handed over directly, and as the
oracle computes them."""
import numpy as np
import pytest

import oracle
import qref_writer as qw
from eagleimp_amd import synth
from eagleimp_amd.engine import Engine
from eagleimp_amd.qref import Qref

from conftest import say
from parity import panel_engine

pytestmark = pytest.mark.gpu


def test_qref_panel_feeds_engine(built, tmp_path):
    panel = synth.make_panel(41, N=512, M=2500, F=16, T=4)
    tg = synth.make_targets(panel, 41, 4, kset=[64, 256, 500], wmin=1200, wmax=2500, miss=0.01)
    M = panel.M
    path = str(tmp_path / "panel.qref")
    qw.write_qref(path, panel.site, panel.N // 2, np.arange(M) * 100, np.zeros(M, np.float32),
                  [("A", "G")] * M, [""] * M)
    say("qref panel", path, "N", panel.N, "M", M)
    with Qref(path) as q:
        rows = q.site_rows()
        eng = Engine(0)
        eng.initialize(2, q.Mref, q.Ncorr, q.Nad, rows)
        eng.set_panel_extras(panel.cM)
    ref = panel_engine(panel)
    a = eng.run_batch(tg.gt, tg.s0, tg.W, tg.K)
    b = ref.run_batch(tg.gt, tg.s0, tg.W, tg.K)
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    for t in range(4):
        assert int(a.Mp[t]) == int(b.Mp[t])
        assert int(a.dig_pbwt[t]) == int(b.dig_pbwt[t])
        assert int(a.dig_count0[t]) == int(b.dig_count0[t])
        r = oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t], keep=False)
        assert int(a.dig_pbwt[t]) == r["dig_pbwt"] and int(a.Mp[t]) == r["Mp"]
    eng.close()
    ref.close()


def test_masked_qref_panel_feeds_engine(built, tmp_path):
    """Withdrawn samples masked from a chromosome-X Qref panel with haploid
    samples: the engine fed the masked rows and flags gives the oracle's
    outputs on the panel without those samples."""
    from oracle import qref_mask
    panel = synth.make_panel(43, N=600, M=2200, F=16, T=3)
    tg = synth.make_targets(panel, 43, 3, kset=[64, 300], wmin=1000, wmax=2200, miss=0.01)
    M, Ns = panel.M, panel.N // 2
    hap = (np.arange(Ns) % 9 == 4).astype(np.uint8)
    S = np.unpackbits(panel.site.view(np.uint8), axis=1, bitorder="little")
    for s in np.flatnonzero(hap):  # haploid samples are stored homozygous
        S[:, 2 * s + 1] = S[:, 2 * s]
    site = np.packbits(S, axis=1, bitorder="little").view(np.uint64)
    path = str(tmp_path / "x.qref")
    qw.write_qref(path, site, Ns, np.arange(M) * 100, np.zeros(M, np.float32),
                  [("A", "G")] * M, [""] * M, chrom=23, haploid=hap)
    ms = np.array([0, 4, 13, 31, 32, 150, 151, 152, 299], np.uint64)
    say("masked qref panel", path, "samples", Ns, "masked", ms.size, "M", M)
    with Qref(path) as q:
        q.mask(ms)
        rows = q.site_rows()
        flags = q.haploid_flags()
        eng = Engine(0)
        eng.initialize(2, q.Mref, q.Ncorr, q.Nad, rows)
        eng.set_panel_extras(panel.cM, haploid=flags)
        Nad, Ncorr = q.Nad, q.Ncorr
    assert np.array_equal(flags, qref_mask.kept_haploid(hap, ms))
    keep = np.ones(2 * Ns, bool)
    for s in ms:
        keep[2 * int(s)] = keep[2 * int(s) + 1] = False
    Sk = S[:, :2 * Ns][:, keep]
    assert Sk.shape[1] == Nad
    hw, sw = synth.unit_words(M), synth.unit_words(Nad)
    Hf = np.zeros((Nad, hw * 64), np.uint8)
    Hf[:, :M] = Sk.T
    Sf = np.zeros((M, sw * 64), np.uint8)
    Sf[:, :Nad] = Sk
    pv = oracle.PanelView(np.packbits(Hf, axis=1, bitorder="little").view(np.uint64),
                          np.packbits(Sf, axis=1, bitorder="little").view(np.uint64), panel.cM,
                          haploid_ref=flags)
    assert pv.N_corr == Ncorr
    a = eng.run_batch(tg.gt, tg.s0, tg.W, tg.K)
    for t in range(3):
        r = oracle.process_target(pv, tg.gt[t], tg.s0[t], tg.W[t], tg.K[t], keep=False)
        assert int(a.Mp[t]) == r["Mp"]
        assert int(a.dig_pbwt[t]) == r["dig_pbwt"] and int(a.dig_count0[t]) == r["dig_count0"]
    eng.close()
