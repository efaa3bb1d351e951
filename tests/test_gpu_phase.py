"""Phasing over the device's output: the engine's cPBWTs (drop-in layout, as
the reference's GPU handler hands them to Target::phaseExtPBWT,
src/GPUHandler.cpp:259-260) phase each target exactly as the oracle's CPU
cPBWTs do, and as the reference's own beam search does where its harness is
built."""
import numpy as np
import pytest

import oracle
import phase_cases as pc
from eagleimp_amd import phase

from conftest import say
from parity import panel_engine

pytestmark = pytest.mark.gpu

CASES = [("default", 0), ("roh_hom_splits_prephase", 0), ("long_roh_deep_history", 2),
         ("tiny_beam_small_K", 1)]


@pytest.mark.parametrize("name,seed", CASES)
def test_phase_device_cpbwt(built, name, seed):
    c = pc.make_case(name, seed)
    r = c["r"]
    say("phase over device cPBWT", name, seed, "K", r["K"], "S", r["S"])
    eng = panel_engine(c["panel"])
    res = eng.run_batch([c["gt"]], [0], [c["W"]], [r["K"]], cMmaxSplit=c["split"])
    assert int(res.Mp[0]) == r["Mp"]
    P = pc.params_for(c)
    dev = phase.phase_batch(P, [phase.engine_job(eng, 0, c["gt"], c["W"], af=c["af"],
                                                 start_phase=c["start"])])[0]
    cpu = phase.phase_target(P, r["pbwt"], r["count0"], r["K"], c["gt"], c["W"],
                             r["splitsites"], r["cMdiffs"], af=c["af"], start_phase=c["start"])
    assert dev.same(cpu)
    if oracle.ref_available():
        assert dev.same(oracle.ref_phase_target(P, pc.job_for(c), r["monos"]))
    eng.close()


@pytest.mark.parametrize("raw", [False, True])
def test_phase_stream_outputs(built, raw):
    """Targets of one stream submission (device-driven path; the raw layout
    too, as the FPGA handler's consumer takes it) phase as on the CPU."""
    c0 = pc.make_case("default", 1)
    panel = c0["panel"]
    # three targets over the same panel: the case's own and two windows of it
    W = c0["W"]
    gts = [c0["gt"], pc.encode_gt(c0["codes"][200:1500]), pc.encode_gt(c0["codes"][700:])]
    s0 = [0, 200, 700]
    Ws = [W, 1300, W - 700]
    Ks = [200, 100, 64]
    say("phase over stream outputs", len(gts), "targets")
    eng = panel_engine(panel)
    st = eng.stream(4, W, 200, 1 << 30, raw=raw)
    st.submit(gts, s0, Ws, Ks)
    res, nd, _ = st.wait()
    assert nd == 0
    P = phase.default_params(pre_phasing=1)
    jobs = [phase.engine_job(st, t, gts[t], Ws[t]) for t in range(3)]
    got = phase.phase_batch(P, jobs, threads=3)
    for t in range(3):
        r = oracle.process_target(c0["pv"], gts[t], s0[t], Ws[t], Ks[t])
        cpu = phase.phase_target(P, r["pbwt"], r["count0"], r["K"], gts[t], Ws[t],
                                 r["splitsites"], r["cMdiffs"])
        assert got[t].same(cpu), t
    st.close()
    eng.close()
