import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")
    config.addinivalue_line("markers", "slow: long-running (full-size workloads)")


@pytest.fixture(scope="session")
def kf():
    from openfusion_amd import kinfu
    kinfu.lib()  # fails loudly if libkinfu.so is missing
    return kinfu


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def synth():
    from openfusion_amd import synth
    synth.slib()
    return synth


@pytest.fixture(scope="session")
def gpu(kf):
    n = kf.device_count()
    if n < 1:
        raise RuntimeError("GPU test on a machine without a visible GPU")
    return n
