"""Serving loop on the device: requests arriving as a Poisson process overlap
in three stream contexts; every request's results equal the same request run
alone, and a sample of targets equals the CPU oracle."""
import numpy as np
import pytest

import oracle
from eagleimp_amd import serve, synth

from conftest import say
from parity import panel_engine

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.mark.parametrize("pool_mb", [256, 8])
def test_alone_equals_mix(built, pool_mb):
    """pool_mb=8 holds only a few targets per context: the mix also runs
    through deferral and re-admission."""
    panel = synth.make_panel(13, N=400, M=6000, F=24, T=48)
    pool = synth.make_targets(panel, 13, 48, kset=[100, 250, 400], wmin=500, wmax=6000,
                              miss=0.0, miss_hi=0.05)
    eng = panel_engine(panel)
    srv = serve.Server(eng, pool, 6000, 400, contexts=3, max_batch=16, pool_bytes=pool_mb * MB)
    t, size, first = synth.make_arrivals(7, 40, 400.0, 48, 1, 24)
    mix = serve.make_requests(t, size, first, 48)
    say("serving mix", len(mix), "requests, pool MB", pool_mb)
    wall = serve.run_requests(srv, mix)
    s = serve.summarize(mix, wall)
    say("mix", s, "deferred", srv.deferred, "submissions", srv.submissions)
    assert s["requests"] == 40
    for r in mix:
        alone = serve.make_requests([0.0], [len(r.targets)], [r.targets[0]], 48)
        serve.run_requests(srv, alone, realtime=False)
        assert alone[0].results == r.results, f"request {r.rid}"
    srv.close()
    # a sample against the oracle
    pv = oracle.PanelView(panel.hap, panel.site, panel.cM)
    byt = {}
    for r in mix:
        for p, i in enumerate(r.targets):
            byt[i] = r.results[p]
    for i in sorted(byt)[::7]:
        ref = oracle.process_target(pv, pool.gt[i], pool.s0[i], pool.W[i], pool.K[i], keep=False)
        S, Mp, ncalls, K, dig, dig0 = byt[i]
        assert (S, Mp, ncalls, dig, dig0) == (ref["S"], ref["Mp"], ref["ncalls"], ref["dig_pbwt"],
                                              ref["dig_count0"]), f"target {i}"
    eng.close()
