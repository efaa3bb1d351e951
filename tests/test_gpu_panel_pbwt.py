"""Panel PBWT for imputation (eimp_engine_ref_pbwt, the reference's
PBWT(refT, true)) on the device vs the oracle: forward rows, count0 and every
absolute order, over whole panels and sub-ranges of sites."""
import numpy as np
import pytest

import oracle
from eagleimp_amd import synth
from eagleimp_amd.engine import EimpError

from conftest import say
from parity import panel_engine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,M,m0,n", [(300, 4000, 0, 4000), (300, 4000, 1234, 777),
                                      (64, 500, 3, 97), (1000, 900, 0, 900),
                                      (24000, 40, 5, 33),
                                      # wider than one workgroup's LDS: order in global memory
                                      (24600, 60, 7, 41), (200000, 240, 17, 150)])
def test_panel_pbwt_matches_oracle(built, N, M, m0, n):
    panel = synth.make_panel(31, N=N, M=M, F=16 if N < 100000 else 2000, T=1)
    eng = panel_engine(panel)
    say("panel pbwt N", N, "M", M, "sites", m0, n)
    rows, c0, ap, ms = eng.ref_pbwt(m0, n)
    er, ec, ea = oracle.build_ref_pbwt(panel.site[m0:m0 + n], N)
    assert np.array_equal(rows, er)
    assert np.array_equal(c0, ec)
    assert np.array_equal(ap, ea)
    assert ms > 0
    rows2, c02, ap2, _ = eng.ref_pbwt(m0, n, absperm=False)
    assert ap2 is None and np.array_equal(rows2, er) and np.array_equal(c02, ec)
    eng.close()


def test_panel_pbwt_limits(built):
    panel = synth.make_panel(32, N=300, M=100, F=8, T=1)
    eng = panel_engine(panel)
    with pytest.raises(EimpError):
        eng.ref_pbwt(90, 20)  # past the panel's last site
    with pytest.raises(EimpError):
        eng.ref_pbwt(0, 0)  # an empty site range
    eng.close()
