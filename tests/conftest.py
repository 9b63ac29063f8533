import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def ref():
    from oracle.oracle import Reference, have_ref

    if not have_ref():
        pytest.skip("spica's compiled sources (oracle/_ref) are only built where /root/reference exists")
    return Reference()


@pytest.fixture(scope="session")
def gpu():
    """The HIP library on cuda:0; fails (never skips) when no GPU is visible."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from spica_amd import render

    render.init(0)
    return render
