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


def check_against_direct(uh, oracle_mod, refs, q, strategy):
    """The bar for the Gram strategies of the exact L2 index: every result is
    the binary32 direct path's result (bit-identical distance) except where
    the binary64 near-tie re-rank chose another row, whose binary64 distance
    is then not larger; every returned distance is the binary32
    reference-order distance of its row; agreement with the binary64 argmin
    is at least the direct path's. Returns (mixed statistics, agreement of
    the strategy, agreement of the direct path)."""
    import torch

    b = q.shape[0]
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(strategy)
    ws = index.make_query_workspace(b)
    d_idx, d_dist = index.query_batch(q, b, ws, uh.STRATEGY_DIRECT)
    m_idx, m_dist = index.query_batch(q, b, ws, strategy)
    stats = ws.mixed_stats()
    stats["certified"] = ws.shortlist_stats()["certified"]
    dm_idx, dm_dist = index.query_device_batch(torch.from_numpy(q).cuda(), b, ws, strategy)
    np.testing.assert_array_equal(dm_idx, m_idx)
    np.testing.assert_array_equal(dm_dist, m_dist)
    k2 = oracle_mod.k2_distances(refs, q)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(k2)
    np.testing.assert_array_equal(d_idx, exp_idx)
    table = oracle_mod.fp64_table(False, refs, q)
    fp64_idx = table.argmin(1)
    rows = np.arange(b)
    diff = m_idx != d_idx
    np.testing.assert_array_equal(m_dist[~diff].view(np.uint32), d_dist[~diff].view(np.uint32))
    np.testing.assert_array_equal(m_dist.view(np.uint32), k2[rows, m_idx.astype(np.int64)].view(np.uint32))
    assert np.all(table[rows[diff], m_idx[diff].astype(np.int64)] <= table[rows[diff], d_idx[diff].astype(np.int64)])
    agree_mixed = np.mean(m_idx == fp64_idx)
    agree_fp32 = np.mean(d_idx == fp64_idx)
    assert agree_mixed >= agree_fp32
    return stats, agree_mixed, agree_fp32


I8_CODE = 32512


def int8_model(x):
    """The GRAM_INT8 row conversion in numpy (binary32 steps as the kernel
    takes them): returns (a, b, s, n) with int64 planes a, b and codes n
    (x^ = s n / 256 = s (a + b / 256)) and the binary32 scales s."""
    x = np.ascontiguousarray(x, np.float32)
    mx = np.abs(x).max(axis=1)
    pos = mx > 0
    inv = np.zeros_like(mx)
    inv[pos] = np.float32(I8_CODE) / mx[pos]
    s = np.ones_like(mx)
    s[pos] = mx[pos] / np.float32(127.0)
    n = np.rint((x * inv[:, None]).astype(np.float32)).astype(np.int64)
    n = np.clip(n, -I8_CODE, I8_CODE)
    a = (n + 128) >> 8
    b = n - 256 * a
    return a, b, s, n
