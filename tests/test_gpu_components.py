"""The stand-alone hierarchical components (uhann_projection_*, uhann_median_*,
the handles behind the reference's detail::CudaFloatProjection and
detail::CudaMedianL1ScanIndex) and median-L1 at projection dimensions above
the register networks.

* projection: bit-identical to the oracle's K7 replay (cuda_float_projection.cu:
  234-263) for several p and k, including k = 0 (all +0);
* median scan: argmins bit-identical to the oracle's K9 replay
  (cuda_median_l1_scan_index.cu:202-267) for p = 6, 64, 129, 256, 1024;
* the hierarchical L1 index at p = 129, 256, 1024 (the LDS sort) is
  bit-identical to K7 + K9 on every query, like p <= 128;
* errors: a non-finite projection -> OVERFLOW, a p whose padded sort exceeds
  the device's per-block shared memory and a foreign workspace ->
  INVALID_ARGUMENT (the reference's exception types)."""

import ctypes

import numpy as np
import pytest

from conftest import rng
from test_gpu_indexes import _hier_expect

pytestmark = pytest.mark.gpu
F32 = np.float32
OK, INVALID, OVERFLOW = 0, 1, 4


def _lib():
    from ultrahigh_ann_amd import _native

    return _native.lib()


def _f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _device_to_host(ptr, shape):
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    out = np.empty(shape, F32)
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return out


def _projection(mat):
    h = ctypes.c_void_p()
    rc = _lib().uhann_projection_create(_f32p(mat), mat.shape[0], mat.shape[1], 0, ctypes.byref(h))
    return rc, h


def _median(proj_refs):
    h = ctypes.c_void_p()
    rc = _lib().uhann_median_create(_f32p(proj_refs), proj_refs.shape[0], proj_refs.shape[1], 0, ctypes.byref(h))
    return rc, h


def _workspace(create, h, cap):
    ws = ctypes.c_void_p()
    assert create(h, cap, ctypes.byref(ws)) == OK
    return ws


@pytest.mark.parametrize("p,k,b", [(16, 300, 9), (31, 513, 20), (64, 917, 33), (5, 0, 4)])
def test_projection_matches_k7(uh, oracle_mod, p, k, b):
    g = rng(p * 7 + k)
    mat = (g.standard_normal((p, k)) * 2).astype(F32)
    x = (g.random((b, k), dtype=np.float32) * 3).astype(F32)
    rc, h = _projection(mat)
    assert rc == OK
    ws = _workspace(_lib().uhann_projection_workspace_create, h, b)
    out = ctypes.c_void_p()
    assert _lib().uhann_projection_project(h, ws, _f32p(x), b, ctypes.byref(out)) == OK
    got = _device_to_host(out.value, (b, p))
    if k:
        want, bad = oracle_mod.k7_project(x, mat)
        assert not bad
    else:
        want = np.zeros((b, p), F32)
    np.testing.assert_array_equal(got.view(np.uint32), np.asarray(want, F32).view(np.uint32))
    # a foreign workspace is rejected before any work
    rc2, h2 = _projection(mat)
    assert _lib().uhann_projection_project(h2, ws, _f32p(x), b, ctypes.byref(out)) == INVALID
    _lib().uhann_workspace_destroy(ws)
    _lib().uhann_projection_destroy(h)
    _lib().uhann_projection_destroy(h2)


def test_projection_overflow(uh):
    mat = np.full((2, 3), 3e38, F32)
    rc, h = _projection(mat)
    assert rc == OK
    ws = _workspace(_lib().uhann_projection_workspace_create, h, 1)
    out = ctypes.c_void_p()
    assert _lib().uhann_projection_project(h, ws, _f32p(np.full((1, 3), 3e38, F32)), 1, ctypes.byref(out)) == OVERFLOW
    _lib().uhann_workspace_destroy(ws)
    _lib().uhann_projection_destroy(h)
    bad = np.array([[1.0, np.inf]], F32)
    assert _projection(bad)[0] == OVERFLOW


@pytest.mark.parametrize("p", [6, 64, 129, 256, 1024])
def test_median_scan_matches_k9(uh, oracle_mod, p):
    import torch

    g = rng(p)
    r, b = 150, 12
    refs = (g.standard_normal((r, p)) * 2).astype(F32)
    q = (refs[g.integers(0, r, b)] + g.standard_normal((b, p)) * 0.5).astype(F32)
    refs[-1] = refs[3]  # duplicate row: the lower index wins
    q[0] = refs[3]
    rc, h = _median(refs)
    assert rc == OK
    ws = _workspace(_lib().uhann_median_workspace_create, h, b)
    dq = torch.from_numpy(q).cuda()
    out = np.zeros(b, np.uint64)
    assert _lib().uhann_median_query_device_batch(
        h, ws, ctypes.c_void_p(dq.data_ptr()), b, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) == OK
    dist, bad = oracle_mod.k9_median_l1(refs, q)
    assert not bad
    exp_idx, _ = oracle_mod.fp32_argmin(dist)
    np.testing.assert_array_equal(out, exp_idx)
    assert out[0] == 3
    _lib().uhann_workspace_destroy(ws)
    _lib().uhann_median_destroy(h)


def test_median_limits(uh):
    assert _median(np.ones((2, 1 << 20), F32))[0] == INVALID  # 4 MiB sort: beyond any LDS
    assert _median(np.zeros((0, 4), F32))[0] == INVALID
    assert _median(np.array([[np.nan, 1.0]], F32))[0] == OVERFLOW


@pytest.mark.parametrize("p", [129, 256, 1024])
def test_hierarchical_l1_wide_projection(uh, oracle_mod, p):
    r, d, b, T = 120, 700, 24, 32
    g = rng(p + 1)
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(F32)
    q = (refs[g.integers(0, r, b)] + g.random((b, d), dtype=np.float32) * 0.5).astype(F32)
    probs = oracle_mod.importance_probabilities(True, refs)
    index = uh.HipHierarchicalL1AnnIndex(refs, probs, T, p, uh.Rng(11))
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws)
    built, (e_idx, e_dist) = _hier_expect(oracle_mod, True, refs, q, probs, T, p, 11)
    np.testing.assert_array_equal(index.sample["source_column"], built["cols"])
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))


def test_serving_rejects_wide_median(uh):
    refs = np.random.default_rng(0).random((10, 50)).astype(F32)
    probs = np.full(50, 0.5)
    with uh.Server(hbm_budget_bytes=1 << 30) as srv:
        with pytest.raises(uh.UhannInvalidArgument):
            srv.register_hierarchical(refs, probs, 4, 129, uh.Rng(1), uh.METRIC_L1)
