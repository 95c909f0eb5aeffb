import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


# PyTorch-ROCm ships its own HIP runtime under the same soname
# (libamdhip64.so.7).  When torch initialises it first, libkinc.so binds to
# that copy and one runtime serves both; when libkinc.so loads first (from
# /opt/rocm), a later torch import brings a second runtime that finds no GPU.
# The GPU tests that use torch tensors therefore need torch's GPU set up
# before any test loads the library.
try:
    import torch as _torch

    if _torch.cuda.is_available():
        _torch.cuda.init()
except Exception:  # no torch, or no GPU: nothing to set up
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def native():
    from kinc_amd import _native
    _native.lib()
    return _native
