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
