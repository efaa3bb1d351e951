"""Shared checks of the GPU parity tests (device outputs vs the oracle)."""
import numpy as np

import oracle
from eagleimp_amd.engine import Engine


def panel_engine(panel, Ncorr=None, haploid=None, give_hap=False):
    eng = Engine(0)
    Nad = panel.N
    Ncorr = Nad if Ncorr is None else Ncorr
    eng.initialize(min(Ncorr, 2), panel.M, Ncorr, Nad, panel.site)
    eng.set_panel_extras(panel.cM, hap=panel.hap if give_hap else None, haploid=haploid)
    return eng


def check_target(ref, got, K, layout, what, monos=None):
    """layout: "gpu", "compact", "raw" (True / False stand for compact / gpu).
    Rows are compared when both sides carry them; monos: the device's mono
    words (copy_monos), compared with the oracle's findMonos."""
    if layout is True:
        layout = "compact"
    elif layout is False:
        layout = "gpu"
    assert np.array_equal(got["besthaps"], ref["besthaps"]), f"{what}: besthaps"
    assert np.array_equal(got["bhflags"], ref["flags"]), f"{what}: best-hap flags"
    if monos is not None:
        assert np.array_equal(monos, ref["monos"]), f"{what}: monos"
    assert np.array_equal(got["splitsites"], ref["splitsites"]), f"{what}: split sites"
    assert np.array_equal(got["splits"], ref["splits"]), f"{what}: split bitmap"
    assert np.array_equal(got["cMdiffs"].view(np.uint32), ref["cMdiffs"].view(np.uint32)), \
        f"{what}: cM diffs"
    assert np.array_equal(got["count0"], ref["count0"]), f"{what}: count0"
    if "cref" in got and "cref" in ref:
        assert np.array_equal(got["cref"], ref["cref"]), f"{what}: condensed reference"
    if "pbwt" not in ref or got["pbwt"].shape[0] == 0:
        return
    KW = oracle.pbwt_row_words(K)
    if layout == "compact":
        assert np.array_equal(got["pbwt"], ref["pbwt"]), f"{what}: cPBWT (compact)"
    elif layout == "raw":
        assert np.array_equal(got["pbwt"], oracle.raw_rows(ref["pbwt"], K)), f"{what}: cPBWT (raw)"
    else:
        assert np.array_equal(got["pbwt"][:, :KW], ref["pbwt"]), f"{what}: cPBWT logical words"
        assert not got["pbwt"][:, KW:].any(), f"{what}: cPBWT padding pairs not zero"
