"""The C++ drop-ins driven the way the reference drives them (built by
__graft_entry__.build()): the GPUEngine (include/eimp_gpu_engine.hpp) as
GPUHandler::preProcessGPUOnly drives it with host prepPhase
(tests/dropin/gpuhandler_driver.cpp), and the GPUHandler
(include/eimp_gpu_handler.hpp) with device prepPhase as PBWTPhaser::phaseGPU
drives it (tests/dropin/gpuhandler_device_driver.cpp)."""
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


DEVICE_DRIVER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "_build",
                             "gpuhandler_device_driver")


@pytest.mark.parametrize("N,M,F,T,K,seed,consumers,haploid_every", [
    (400, 3000, 16, 40, 200, 4, 4, 0),      # subset K, top-K on the device
    (300, 2500, 12, 300, 300, 5, 8, 7),     # two blocks; haploid samples, K = Ncorr (take all)
    (512, 2000, 20, 12, 512, 6, 2, 0),      # whole panel (identity)
])
def test_cpp_gpuhandler_device_prephase(N, M, F, T, K, seed, consumers, haploid_every):
    """GPUHandler drop-in (include/eimp_gpu_handler.hpp): producer with device
    prepPhase, cPBWT queue items carrying count0, consumer threads phasing;
    every item equals the oracle and every phased target equals phasing of
    the oracle's rows."""
    if not os.path.exists(DEVICE_DRIVER):
        pytest.fail(f"{DEVICE_DRIVER} missing: run __graft_entry__.build()")
    say("cpp GPUHandler device prephase", N, M, F, T, K, seed, consumers, haploid_every)
    r = subprocess.run([DEVICE_DRIVER, str(N), str(M), str(F), str(T), str(K), str(seed),
                        str(consumers), str(haploid_every)],
                       capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("OK"), r.stdout + r.stderr
