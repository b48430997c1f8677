"""Serving loop over the engine's C ABI (uhann_server_*, include/uhann.h).

A ``Server`` keeps indexes resident on one device under an HBM budget and
answers requests (one batch of queries for one index) by continuous
batching: a native worker packs queued requests of resident indexes into
ticks of grouped persistent kernels, one captured hipGraph per tick shape.
Index construction draws from the caller's ``Rng`` exactly as the one-index
facades do (``HipFlatL2AnnIndex`` etc.), so a served index answers with the
bits of the corresponding facade.

No counterpart in the reference, which answers one facade ``query_batch`` at
a time (nearest_neighbor_benchmark.cpp:386-413).
"""

from __future__ import annotations

import ctypes
from ctypes import c_int, c_uint64, c_void_p

import numpy as np

from . import _native
from ._native import IndexSpec, ServerConfig, ServingStats, TraceEntry, UhannInvalidArgument, check
from .indexes import (
    KIND_EXACT,
    KIND_FLAT,
    KIND_HIERARCHICAL,
    METRIC_L2,
    SAMPLE_DTYPE,
    STRATEGY_DIRECT,
    Rng,
    _flat_sample,
    _refs,
    build_projection,
)

INDEX_PENDING, INDEX_RESIDENT, INDEX_FAILED = 0, 1, 2


class Request:
    """One submitted request; ``wait`` returns (indices, distances)."""

    def __init__(self, server, rid, count, out_idx, out_dist, keep):
        self.server, self.id, self.count = server, rid, count
        self.out_idx, self.out_dist = out_idx, out_dist
        self._keep = keep  # query buffers stay alive until completion
        self._status = None

    def wait(self, timeout=None):
        if self._status is None:
            done, st = c_int(), c_int()
            check(_native.lib().uhann_server_wait(self.server._h, self.id, -1.0 if timeout is None else timeout,
                                                  ctypes.byref(done), ctypes.byref(st)))
            if not done.value:
                raise TimeoutError(f"request {self.id} did not complete within {timeout} s")
            self._status = st.value
            self._message = _native.lib().uhann_last_error().decode(errors="replace")
            self._keep = None
        if self._status != 0:
            raise _native._STATUS.get(self._status, _native.UhannError)(self._message)
        return self.out_idx, self.out_dist


def _queries(queries, dim):
    """(pointer, row stride, on_device, count, keep-alive) of a host array or a torch device tensor."""
    if hasattr(queries, "data_ptr") and getattr(queries, "is_cuda", False):
        t = queries if queries.dim() == 2 else queries.reshape(-1, dim)
        if t.dtype != __import__("torch").float32 or t.stride(-1) != 1:
            raise UhannInvalidArgument("device queries must be float32 rows with unit column stride")
        return t.data_ptr(), t.stride(0), 1, t.shape[0], t
    q = np.ascontiguousarray(queries, dtype=np.float32)
    if q.ndim == 1:
        q = q.reshape(-1, dim) if dim else q.reshape(1, 0)
    return q.ctypes.data, q.shape[1], 0, q.shape[0], q


class Server:
    """uhann_server: resident indexes, request queue and tick worker."""

    def __init__(self, device=0, hbm_budget_bytes=64 << 30, max_tick_queries=8192, max_tick_requests=512,
                 scratch_floats=64 << 20, staging_floats=16 << 20, use_graph=True, pairs_tq=4):
        cfg = ServerConfig(device, hbm_budget_bytes, max_tick_queries, max_tick_requests, scratch_floats,
                           staging_floats, int(bool(use_graph)), pairs_tq)
        h = c_void_p()
        check(_native.lib().uhann_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._dims = {}
        self._keep = {}

    # ---- indexes -------------------------------------------------------------
    def _register(self, kind, metric, refs, sample=None, projection=None, p=0, strategy=STRATEGY_DIRECT):
        s = np.ascontiguousarray(sample if sample is not None else np.zeros(0, SAMPLE_DTYPE), dtype=SAMPLE_DTYPE)
        proj = None if projection is None else np.ascontiguousarray(projection, dtype=np.float64)
        spec = IndexSpec(kind, metric, strategy, refs.ctypes.data, refs.shape[0], refs.shape[1],
                         s.ctypes.data if s.size else None, s.size,
                         proj.ctypes.data if proj is not None and proj.size else None, p)
        rid = c_uint64()
        check(_native.lib().uhann_server_register(self._h, ctypes.byref(spec), ctypes.byref(rid)))
        self._dims[rid.value] = refs.shape[1]
        self._keep[rid.value] = refs  # borrowed by the server until the index is resident
        return rid.value

    def register_exact(self, refs, metric=METRIC_L2):
        return self._register(KIND_EXACT, metric, _refs(refs))

    def register_flat(self, refs, probabilities, repetitions, rng: Rng, metric=METRIC_L2, sample=None):
        """Draws the sample like HipFlatL{1,2}AnnIndex(refs, probabilities, T, rng)."""
        refs = _refs(refs)
        if sample is None:
            sample = _flat_sample(refs, probabilities, repetitions, rng)
        return self._register(KIND_FLAT, metric, refs, sample)

    def register_hierarchical(self, refs, probabilities, repetitions, projection_dimension, rng: Rng,
                              metric=METRIC_L2, strategy=STRATEGY_DIRECT, sample=None):
        """Draws sample and projection like HipHierarchicalL{1,2}AnnIndex(refs, probabilities, T, p, rng)."""
        refs = _refs(refs)
        if projection_dimension <= 0:
            raise UhannInvalidArgument("projection_dimension has to be a positive integer")
        if sample is None:
            sample = _flat_sample(refs, probabilities, repetitions, rng)
        s = np.ascontiguousarray(sample, dtype=SAMPLE_DTYPE)
        if s.size == 0 and refs.shape[0] > 1:
            raise _native.UhannError("importance sampling selected no coordinates")
        proj = build_projection(s, projection_dimension, metric, rng)
        return self._register(KIND_HIERARCHICAL, metric, refs, s, proj, projection_dimension, strategy)

    def state(self, index_id):
        st, b = c_int(), c_uint64()
        check(_native.lib().uhann_server_index_state(self._h, index_id, ctypes.byref(st), ctypes.byref(b)))
        return st.value, b.value

    def remove(self, index_id):
        check(_native.lib().uhann_server_remove(self._h, index_id))
        self._keep.pop(index_id, None)

    # ---- requests ------------------------------------------------------------
    def submit(self, index_id, queries, count=None) -> Request:
        dim = self._dims.get(index_id, 0)
        ptr, ld, on_dev, n, keep = _queries(queries, dim)
        if count is not None:
            n = count
        out_idx = np.zeros(n, np.uint64)
        out_dist = np.zeros(n, np.float32)
        rid = c_uint64()
        check(_native.lib().uhann_server_submit(self._h, index_id, ptr, on_dev, ld, n, out_idx.ctypes.data,
                                                out_dist.ctypes.data, ctypes.byref(rid)))
        return Request(self, rid.value, n, out_idx, out_dist, keep)

    def run_trace(self, arrivals, index_ids, queries, counts):
        """Open-loop replay: request i (``counts[i]`` rows of ``queries[i]``) for
        ``index_ids[i]`` is submitted ``arrivals[i]`` seconds after the start.
        Returns the Request objects."""
        n = len(arrivals)
        entries = (TraceEntry * max(n, 1))()
        reqs = []
        for i in range(n):
            ptr, ld, on_dev, rows, keep = _queries(queries[i], self._dims.get(int(index_ids[i]), 0))
            c = int(counts[i]) if counts is not None else rows
            out_idx = np.zeros(c, np.uint64)
            out_dist = np.zeros(c, np.float32)
            entries[i] = TraceEntry(float(arrivals[i]), int(index_ids[i]), ptr, ld, c, on_dev, out_idx.ctypes.data,
                                    out_dist.ctypes.data)
            reqs.append(Request(self, 0, c, out_idx, out_dist, keep))
        ids = np.zeros(max(n, 1), np.uint64)
        check(_native.lib().uhann_server_run_trace(self._h, entries, n, ids.ctypes.data_as(
            ctypes.POINTER(c_uint64))))
        for r, i in zip(reqs, ids[:n]):
            r.id = int(i)
        return reqs

    def drain(self):
        check(_native.lib().uhann_server_drain(self._h))

    def stats(self) -> dict:
        st = ServingStats()
        check(_native.lib().uhann_server_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ServingStats._fields_}

    def reset_stats(self):
        check(_native.lib().uhann_server_reset_stats(self._h))

    def autotune(self, device_queries, rows_per_index=32) -> int:
        tq = c_int()
        check(_native.lib().uhann_server_autotune(self._h, device_queries.data_ptr(), device_queries.stride(0),
                                                  rows_per_index, ctypes.byref(tq)))
        return tq.value

    def close(self):
        if getattr(self, "_h", None):
            _native.lib().uhann_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
