"""Shared pytest set-up: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle checks, host logic and the library-export checks on the
CPU; `-m gpu` runs the parity tests of the HIP path against the oracle."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP kernels)")


def _gpu_present() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    """Skips nothing: a gpu-marked test on a machine without a GPU must fail loudly."""
    import cybermeshgenerator_amd as cm
    assert cm.hip_available(), "no HIP device visible to libcmg_hip"
    return cm


def pytest_report_header(config):
    return f"gpu present: {_gpu_present()}"
