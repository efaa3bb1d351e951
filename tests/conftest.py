import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libeimp.so's C ABI)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def built():
    """The oracle and the generator must be built (build() does it)."""
    import oracle
    from eagleimp_amd import synth
    for p in (oracle.ORACLE_SO, os.path.join(os.path.dirname(synth.__file__), "libeimp_synth.so")):
        if not os.path.exists(p):
            pytest.fail(f"{p} missing: run __graft_entry__.build() first")
    return True


def say(*a):
    """Progress line naming the case, flushed before a GPU step."""
    print("[case]", *a, flush=True)
