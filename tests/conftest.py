import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "ct-xafs_analysis_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU-oracle runs")


@pytest.fixture(scope="session")
def ctx():
    """A library context on cuda:0.  Fails (does not skip) without a GPU: the
    gpu-marked tests exist to exercise the HIP path."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("gpu-marked test run without a visible GPU")
    from ctrecon import Context
    c = Context(0)
    yield c
    c.close()
