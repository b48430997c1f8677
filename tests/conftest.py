import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ref_mod(oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("reference CPU build (oracle/_ref) not present")
    oracle_mod.ref()
    return oracle_mod


@pytest.fixture(scope="session")
def uh():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import ultrahigh_ann_amd

    if not ultrahigh_ann_amd.device_available():
        pytest.fail("torch sees a GPU but the HIP engine reports none")
    return ultrahigh_ann_amd


def rng(seed=0):
    return np.random.default_rng(seed)


def bf16_round(a):
    """binary32 -> bf16 (round to nearest even), returned as binary64."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return b.astype(np.uint32).view(np.float32).astype(np.float64)
