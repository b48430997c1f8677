"""MI355X query engine for the exact, flat and hierarchical nearest-neighbour
indexes of martingherold/ultrahigh_ann.

This Python module is a thin host mirror of the engine's C ABI
(include/uhann.h), shaped like the reference's C++ index classes so tests and
benchmarks read like the reference's own tests:

* ``HipDenseL2ScanIndex`` / ``HipDenseL1ScanIndex`` -- the dense scan engines
  (reference ``CudaDenseL2ScanIndex``, ``CudaDenseL1ScanIndex``,
  include/ultrahigh_ann/core/cuda_dense_l{2,1}_scan_index.hpp).
* ``HipExactL2Index`` / ``HipExactL1Index`` -- exact facades
  (include/ultrahigh_ann/exact/l{2,1}/cuda_exact_l{2,1}_index.hpp).

Every query runs the hand-written HIP kernels; nothing here computes a
distance on the CPU.
"""

from __future__ import annotations

import ctypes
from ctypes import c_int, c_uint64, c_void_p

import numpy as np

from . import _native
from ._native import (
    UhannDomainError,
    UhannError,
    UhannInvalidArgument,
    UhannLengthError,
    UhannLogicError,
    UhannOutOfRange,
    UhannOverflowError,
    check,
    f32ptr,
    u64ptr,
)

from .indexes import (  # noqa: E402
    HipExactIndex,
    HipFlatL1AnnIndex,
    HipFlatL2AnnIndex,
    HipHierarchicalL1AnnIndex,
    HipHierarchicalL2AnnIndex,
    HipSampledCoordinateL1AnnIndex,
    HipSampledCoordinateL2AnnIndex,
    HipUniformL1AnnIndex,
    HipUniformL2AnnIndex,
    Rng,
    build_coordinate_sample,
    build_projection,
    importance_probabilities,
    importance_probabilities_device,
    ProbabilityFile,
    load_importance_probabilities,
    make_sample,
    probability_source_matches,
    sampling_mass,
    save_importance_probabilities,
    sha256_file,
    uniform_probabilities,
)
from .serving import INDEX_FAILED, INDEX_PENDING, INDEX_RESIDENT, Server  # noqa: E402
from .diagnostics import ExactDistanceTable, evaluate_selected_distances  # noqa: E402

METRIC_L2 = 0
METRIC_L1 = 1
STRATEGY_DIRECT = 0
STRATEGY_GRAM_BF16 = 1
STRATEGY_PROJECTION_BF16 = 2
STRATEGY_GRAM_FP8 = 3
STRATEGY_GRAM_INT8 = 4

__all__ = [
    "METRIC_L1",
    "METRIC_L2",
    "STRATEGY_DIRECT",
    "STRATEGY_GRAM_BF16",
    "STRATEGY_PROJECTION_BF16",
    "STRATEGY_GRAM_FP8",
    "STRATEGY_GRAM_INT8",
    "int8_rows",
    "HipDenseL1ScanIndex",
    "HipExactIndex",
    "HipFlatL1AnnIndex",
    "HipFlatL2AnnIndex",
    "HipHierarchicalL1AnnIndex",
    "HipHierarchicalL2AnnIndex",
    "HipSampledCoordinateL1AnnIndex",
    "HipSampledCoordinateL2AnnIndex",
    "HipUniformL1AnnIndex",
    "HipUniformL2AnnIndex",
    "Rng",
    "sampling_mass",
    "uniform_probabilities",
    "ProbabilityFile",
    "load_importance_probabilities",
    "probability_source_matches",
    "save_importance_probabilities",
    "sha256_file",
    "build_coordinate_sample",
    "build_projection",
    "importance_probabilities",
    "importance_probabilities_device",
    "make_sample",
    "Server",
    "ExactDistanceTable",
    "evaluate_selected_distances",
    "INDEX_FAILED",
    "INDEX_PENDING",
    "INDEX_RESIDENT",
    "HipDenseL2ScanIndex",
    "HipExactL1Index",
    "HipExactL2Index",
    "UhannDomainError",
    "UhannError",
    "UhannInvalidArgument",
    "UhannLengthError",
    "UhannLogicError",
    "UhannOutOfRange",
    "UhannOverflowError",
    "device_available",
    "device_name",
]


def device_available() -> bool:
    return bool(_native.lib().uhann_device_available())


def device_name(device: int = 0) -> str:
    buf = ctypes.create_string_buffer(256)
    check(_native.lib().uhann_device_name(device, buf, 256))
    return buf.value.decode()


def _f32_matrix(values, rows=None, cols=None) -> np.ndarray:
    arr = np.ascontiguousarray(values, dtype=np.float32)
    if rows is not None:
        arr = arr.reshape(rows, cols)
    return arr


def _torch_stream_ptr(stream):
    if stream is None:
        return None
    return c_void_p(int(stream.cuda_stream) if hasattr(stream, "cuda_stream") else int(stream))


class QueryWorkspace:
    """Per-caller device buffers (reference ``QueryWorkspace``)."""

    def __init__(self, handle: c_void_p, owner):
        self._h = handle
        self.owner = owner

    @property
    def handle(self):
        return self._h

    def maximum_batch_size(self) -> int:
        mb = c_uint64()
        check(_native.lib().uhann_workspace_info(self._h, ctypes.byref(mb), None))
        return mb.value

    def payload_bytes(self) -> int:
        pb = c_uint64()
        check(_native.lib().uhann_workspace_info(self._h, None, ctypes.byref(pb)))
        return pb.value

    def set_timing(self, enable: bool = True):
        check(_native.lib().uhann_workspace_set_timing(self._h, int(enable)))

    def stage_times(self) -> list:
        ms = np.zeros(4, np.float32)
        check(_native.lib().uhann_workspace_stage_times(self._h, f32ptr(ms), 4))
        return [float(x) for x in ms]

    def shortlist_stats(self) -> dict:
        """Certified queries of the last GRAM_FP8 or GRAM_INT8 batch."""
        c = c_uint64()
        check(_native.lib().uhann_workspace_shortlist_stats(self._h, ctypes.byref(c)))
        return {"certified": c.value}

    def mixed_stats(self) -> dict:
        c, f, e = c_uint64(), c_uint64(), c_uint64()
        check(_native.lib().uhann_workspace_mixed_stats(self._h, ctypes.byref(c), ctypes.byref(f),
                                                          ctypes.byref(e)))
        return {"candidates": c.value, "fp64_reranked": f.value, "escalated": e.value}

    def close(self):
        if self._h:
            _native.lib().uhann_workspace_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DenseScan:
    _metric = METRIC_L2

    def __init__(self, representatives, weights=None, device: int = 0):
        refs = _f32_matrix(representatives)
        if refs.ndim != 2:
            raise UhannInvalidArgument("representatives must be a 2-D matrix")
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32)
        if w is not None and w.shape != (refs.shape[1],):
            raise UhannInvalidArgument("coordinate weights do not match dimension")
        h = c_void_p()
        check(
            _native.lib().uhann_scan_create(
                self._metric, f32ptr(refs), refs.shape[0], refs.shape[1], f32ptr(w), device,
                ctypes.byref(h),
            )
        )
        self._h = h
        self.rows, self.dim = refs.shape

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().uhann_scan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def payload_bytes(self) -> int:
        pb = c_uint64()
        check(_native.lib().uhann_scan_info(self._h, None, None, ctypes.byref(pb), None))
        return pb.value

    def device(self) -> int:
        d = c_int()
        check(_native.lib().uhann_scan_info(self._h, None, None, None, ctypes.byref(d)))
        return d.value

    def autotune(self, device_queries, max_batch: int, workspace: QueryWorkspace, ld_queries=None):
        """Time every direct-scan kernel per batch bucket; returns the table."""
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        table = np.zeros((24, 6), np.int32)
        n = c_int()
        check(_native.lib().uhann_scan_autotune(
            self._h, workspace.handle, c_void_p(ptr), self.dim if ld_queries is None else ld_queries, max_batch,
            table.ctypes.data_as(ctypes.POINTER(c_int)), 24, ctypes.byref(n)))
        return [{"max_batch": 1 << int(r[0]), "rpt": int(r[1]), "tq": int(r[2]), "tr": int(r[3]),
                 "group_q": int(r[4]), "us": int(r[5])} for r in table[: n.value]]

    def enable_mixed(self, strategy=STRATEGY_GRAM_BF16):
        """Prepare the BF16 reference copy for the mixed-precision strategy."""
        check(_native.lib().uhann_scan_enable_mixed(self._h, strategy))

    def set_fp8_shortlist(self, rows: int):
        """Rows the GRAM_FP8 strategy re-ranks per query (1 .. 16, default 16)."""
        check(_native.lib().uhann_scan_set_fp8_shortlist(self._h, rows))

    def make_query_workspace(self, maximum_batch_size: int) -> QueryWorkspace:
        h = c_void_p()
        check(_native.lib().uhann_scan_workspace_create(self._h, maximum_batch_size, ctypes.byref(h)))
        return QueryWorkspace(h, self)

    def query_batch(self, queries, count: int, workspace: QueryWorkspace, strategy=STRATEGY_DIRECT):
        """Host queries (count x dim). Returns (indices uint64, distances float32)."""
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1)
        if q.size != count * self.dim:
            raise UhannInvalidArgument("query batch dimensions do not match index dimension")
        idx = np.zeros(count, dtype=np.uint64)
        dist = np.zeros(count, dtype=np.float32)
        check(
            _native.lib().uhann_scan_query_batch(
                self._h, workspace.handle, f32ptr(q), count, u64ptr(idx), f32ptr(dist), strategy
            )
        )
        return idx, dist

    def query(self, query, workspace: QueryWorkspace, strategy=STRATEGY_DIRECT) -> int:
        idx, _ = self.query_batch(query, 1, workspace, strategy)
        return int(idx[0])

    def query_device_batch(self, device_queries, count: int, workspace: QueryWorkspace,
                           strategy=STRATEGY_DIRECT):
        """Device queries (a torch tensor or raw pointer, row stride dim)."""
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        idx = np.zeros(count, dtype=np.uint64)
        dist = np.zeros(count, dtype=np.float32)
        check(
            _native.lib().uhann_scan_query_device_batch(
                self._h, workspace.handle, c_void_p(ptr), count, u64ptr(idx), f32ptr(dist), strategy
            )
        )
        return idx, dist

    def query_device_async(self, device_queries, count: int, workspace: QueryWorkspace,
                           out_idx=None, out_dist=None, strategy=STRATEGY_DIRECT, stream=None,
                           ld_queries=None):
        """Device-to-device, asynchronous on `stream` (torch stream or raw handle)."""
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        ld = self.dim if ld_queries is None else ld_queries
        oi = c_void_p(out_idx.data_ptr()) if out_idx is not None else None
        od = c_void_p(out_dist.data_ptr()) if out_dist is not None else None
        check(
            _native.lib().uhann_scan_query_device_async(
                self._h, workspace.handle, c_void_p(ptr), ld, count, oi, od, strategy,
                _torch_stream_ptr(stream),
            )
        )


    def query_device_tile(self, device_queries, count: int, workspace: QueryWorkspace, out_idx,
                          out_dist, tile, stream=None, ld_queries=None):
        """query_device_async with an explicit (rpt, tq, tr, group_q) tile."""
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        ld = self.dim if ld_queries is None else ld_queries
        rpt, tq, tr, gq = tile
        check(
            _native.lib().uhann_scan_query_device_tile(
                self._h, workspace.handle, c_void_p(ptr), ld, count,
                c_void_p(out_idx.data_ptr()) if out_idx is not None else None,
                c_void_p(out_dist.data_ptr()) if out_dist is not None else None,
                rpt, tq, tr, gq, _torch_stream_ptr(stream),
            )
        )


    def gram_estimates(self, device_queries, count: int, workspace: QueryWorkspace, out, tile: int = -1,
                       stream=None, ld_queries=None):
        """Gram estimates ||q||^2 + ||v||^2 - 2 <bf16(q), bf16(v)> into the
        device tensor out (count x rows, row stride out.stride(0)). tile -1
        runs the queries' configuration (default block shape, the tiles of a
        partial last round split along K); 0..7 force a block shape with every
        tile accumulated whole."""
        ld = ld_queries if ld_queries is not None else self.dim
        check(
            _native.lib().uhann_scan_gram_estimates(
                self._h, workspace.handle, c_void_p(device_queries.data_ptr()), ld, count,
                c_void_p(out.data_ptr()), out.stride(0), tile, _torch_stream_ptr(stream),
            )
        )


    def distance_table(self, device_queries, count: int, out, stream=None, ld_queries=None):
        """Exact binary64 diagnostic distances (L2: sqrt of the squared
        distance, L1: the L1 distance; the reference CPU order) of count
        device queries against every reference into the float64 device
        tensor out (count x rows, row stride out.stride(0))."""
        import torch

        if out.dtype != torch.float64:
            raise UhannInvalidArgument("the distance table is float64")
        ld = ld_queries if ld_queries is not None else self.dim
        check(
            _native.lib().uhann_scan_distance_table(
                self._h, c_void_p(device_queries.data_ptr()), ld, count, c_void_p(out.data_ptr()),
                out.stride(0), _torch_stream_ptr(stream),
            )
        )

    def fp8_estimates(self, device_queries, count: int, workspace: QueryWorkspace, out, stream=None,
                      ld_queries=None):
        """FP8 estimates ||q||^2 + ||v||^2 - 2 s_q s_v <q8, v8> into the device
        tensor out (count x rows, row stride out.stride(0))."""
        ld = ld_queries if ld_queries is not None else self.dim
        check(
            _native.lib().uhann_scan_fp8_estimates(
                self._h, workspace.handle, c_void_p(device_queries.data_ptr()), ld, count,
                c_void_p(out.data_ptr()), out.stride(0), _torch_stream_ptr(stream),
            )
        )

    def int8_estimates(self, device_queries, count: int, workspace: QueryWorkspace, out, tile: int = -1,
                       stream=None, ld_queries=None):
        """GRAM_INT8 estimates ||q||^2 + ||v||^2 - 2 s_q s_v (H + X / 256)
        into the device tensor out (count x rows, row stride out.stride(0));
        tile -1 runs the default block shape with its K split, 0..5 force a
        shape (all give the same bits: the integer sums are exact)."""
        ld = ld_queries if ld_queries is not None else self.dim
        check(
            _native.lib().uhann_scan_int8_estimates(
                self._h, workspace.handle, c_void_p(device_queries.data_ptr()), ld, count,
                c_void_p(out.data_ptr()), out.stride(0), tile, _torch_stream_ptr(stream),
            )
        )


def int8_rows(device_rows, count: int, dim: int, stream=None, ld_rows=None):
    """The GRAM_INT8 conversion of count device rows: (bytes, stats) device
    tensors, bytes count x 128 ceil(dim / 64) uint8 (per 64 columns the high
    plane, then the low plane), stats count x 32 uint8 records (float64
    ||x||^2; float32 s, ||x - x^||, ||x^||, s ||b|| / 256, ||x||, pad)."""
    import torch

    ld = ld_rows if ld_rows is not None else dim
    width = max(128, (dim + 63) // 64 * 128)
    out = torch.zeros((max(count, 1), width), dtype=torch.uint8, device=device_rows.device)
    stats = torch.zeros((max(count, 1), 32), dtype=torch.uint8, device=device_rows.device)
    check(
        _native.lib().uhann_int8_rows(
            c_void_p(device_rows.data_ptr()), ld, count, dim, c_void_p(out.data_ptr()), c_void_p(stats.data_ptr()),
            _torch_stream_ptr(stream),
        )
    )
    return out[:count], stats[:count]


class HipDenseL2ScanIndex(_DenseScan):
    """Exhaustive binary32 squared-L2 scan (reference CudaDenseL2ScanIndex)."""

    _metric = METRIC_L2

    def __init__(self, representatives, device: int = 0):
        super().__init__(representatives, None, device)


class HipDenseL1ScanIndex(_DenseScan):
    """Exhaustive binary32 (optionally weighted) L1 scan (reference CudaDenseL1ScanIndex)."""

    _metric = METRIC_L1


class HipExactL2Index(HipDenseL2ScanIndex):
    """Exact L2 facade (reference CudaExactL2Index, cuda_exact_l2_index.cpp:13-60)."""


class HipExactL1Index(HipDenseL1ScanIndex):
    """Exact L1 facade (reference CudaExactL1Index)."""

    def __init__(self, representatives, device: int = 0):
        super().__init__(representatives, None, device)
