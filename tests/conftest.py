import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "liboracle.so")
    src = os.path.join(ROOT, "oracle", "mfx_oracle.c")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       stdout=subprocess.DEVNULL)


_ensure_oracle()


@pytest.fixture(scope="session")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "reference_serial.npz"))


@pytest.fixture(scope="session")
def gpu():
    """Fails (does not skip) when no GPU is visible: a GPU test that cannot
    reach the device must not pass silently."""
    import torch
    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    from benchmarks_amd import _lib
    _lib.lib()  # raises if libmfx.so is missing
    return torch.device("cuda")
