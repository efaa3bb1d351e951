"""The C++ drop-in GPUEngine (include/eimp_gpu_engine.hpp) driven the way the
reference's GPUHandler drives its engine (tests/dropin/gpuhandler_driver.cpp,
built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import say

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "_build",
                      "gpuhandler_driver")


@pytest.mark.parametrize("N,M,F,T,K,seed", [(256, 3000, 12, 6, 256, 1), (400, 2500, 20, 5, 150, 2),
                                            (300, 1000, 8, 300, 64, 3)])
def test_cpp_dropin_matches_oracle(N, M, F, T, K, seed):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} missing: run __graft_entry__.build()")
    say("cpp drop-in", N, M, F, T, K, seed)
    r = subprocess.run([DRIVER, str(N), str(M), str(F), str(T), str(K), str(seed)],
                       capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
