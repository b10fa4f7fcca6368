import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("spion-emol_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def rel_err(got, ref):
    """max-abs error over max-abs reference (the reference tests' measure)."""
    import numpy as np
    got = np.asarray(got)
    ref = np.asarray(ref)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """Device sessions: torch, which ships its own HIP runtime, initialises
    it before libsehf's runtime is first used, the order of a host program
    that imports torch first (smoke(), the autograd wrapper)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.cuda.init()
    yield
