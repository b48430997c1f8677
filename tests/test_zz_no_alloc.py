"""The allocation-free direct query (uhann_scan_query_device_noalloc), called
through the C ABI, against uhann_scan_query_device_async(DIRECT) on the same
handle: same rows and binary32 distances bit for bit, on two calls in a row
(the second with fewer queries, so stale keys of the first call would show),
for the tiled scan (L2, d = 3000, strided queries), the pair-per-thread scan
(weighted L1, d = 200) and the thread-per-query scan (L2, d = 64). Sizes
beyond the array's capacity are rejected before any device work."""

import ctypes

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def _noalloc(uh, index, dq, ld, count, oi, od):
    from ultrahigh_ann_amd import _native

    return _native.lib().uhann_scan_query_device_noalloc(
        index._h, ctypes.c_void_p(dq.data_ptr()), ld, count, ctypes.c_void_p(oi.data_ptr()),
        ctypes.c_void_p(od.data_ptr()), None)


@pytest.mark.parametrize("metric,d,weighted", [("l2", 3000, False), ("l1", 200, True), ("l2", 64, False)])
def test_noalloc_matches_workspace_query(uh, metric, d, weighted):
    import torch

    g = rng(d + (7 if weighted else 0))
    r = 900
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    if metric == "l2":
        index = uh.HipDenseL2ScanIndex(refs)
    else:
        w = (g.random(d, dtype=np.float32) * 3 + 0.1).astype(F32) if weighted else None
        index = uh.HipDenseL1ScanIndex(refs, w)
    ld = d + 5  # strided rows: the call must honour ld_queries
    for count in (700, 37):
        q = torch.zeros((count, ld), dtype=torch.float32, device="cuda")
        q[:, :d] = torch.from_numpy((g.random((count, d), dtype=np.float32) * 4).astype(F32)).cuda()
        ws = index.make_query_workspace(count)
        ref_i = torch.zeros(count, dtype=torch.int64, device="cuda")
        ref_d = torch.zeros(count, dtype=torch.float32, device="cuda")
        index.query_device_async(q, count, ws, ref_i, ref_d, strategy=uh.STRATEGY_DIRECT, ld_queries=ld)
        oi = torch.full((count,), -1, dtype=torch.int64, device="cuda")
        od = torch.full((count,), -1.0, dtype=torch.float32, device="cuda")
        assert _noalloc(uh, index, q, ld, count, oi, od) == 0
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy(), ref_i.cpu().numpy())
        np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), ref_d.cpu().numpy().view(np.uint32))


def test_noalloc_rejects_oversized_batches(uh):
    import torch

    index = uh.HipDenseL2ScanIndex(np.ones((4, 8), F32))
    q = torch.zeros((1, 8), dtype=torch.float32, device="cuda")
    oi = torch.zeros(1, dtype=torch.int64, device="cuda")
    od = torch.zeros(1, dtype=torch.float32, device="cuda")
    from ultrahigh_ann_amd import _native

    rc = _native.lib().uhann_scan_query_device_noalloc(index._h, ctypes.c_void_p(q.data_ptr()), 8, 65537,
                                                       ctypes.c_void_p(oi.data_ptr()), ctypes.c_void_p(od.data_ptr()),
                                                       None)
    assert rc == 2  # UHANN_ERR_LENGTH
    assert b"65536" in _native.lib().uhann_last_error()
