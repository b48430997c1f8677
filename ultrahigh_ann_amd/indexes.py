"""Host mirror of the reference's flat and hierarchical index facades over
the engine's C ABI (include/uhann.h).

Index construction draws from a std::mt19937_64 held by the engine
(``Rng``) in the reference's call order, so ``HipFlatL2AnnIndex(x, p, T,
Rng(42))`` samples exactly the coordinates the reference's
``CudaFlatL2AnnIndex(x, p, T, engine)`` with ``std::mt19937_64 engine(42)``
samples (hnsvw25/l2/flat/cuda_l2_ann_index.cpp:13-37).
"""

from __future__ import annotations

import ctypes
import os
from ctypes import c_int, c_uint64, c_void_p

import numpy as np

from . import _native
from ._native import UhannInvalidArgument, check, f32ptr, f64ptr, u64ptr

METRIC_L2 = 0
METRIC_L1 = 1
STRATEGY_DIRECT = 0
STRATEGY_GRAM_BF16 = 1
STRATEGY_PROJECTION_BF16 = 2
STRATEGY_GRAM_FP8 = 3
STRATEGY_GRAM_INT8 = 4
KIND_EXACT, KIND_FLAT, KIND_HIERARCHICAL = 0, 1, 2

SAMPLE_DTYPE = np.dtype([("source_column", "<u8"), ("multiplicity", "<u8"), ("inverse_probability", "<f8")])


class Rng:
    """std::mt19937_64 behind the C ABI (the engine every reference constructor takes)."""

    def __init__(self, seed: int):
        h = c_void_p()
        check(_native.lib().uhann_rng_create(seed, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def __eq__(self, other):
        eq = c_int()
        check(_native.lib().uhann_rng_equal(self._h, other._h, ctypes.byref(eq)))
        return bool(eq.value)

    def next(self) -> int:
        v = c_uint64()
        check(_native.lib().uhann_rng_next(self._h, ctypes.byref(v)))
        return v.value

    def __del__(self):
        try:
            if self._h:
                _native.lib().uhann_rng_destroy(self._h)
                self._h = None
        except Exception:
            pass


def make_sample(cols, mult, invp) -> np.ndarray:
    s = np.zeros(len(cols), dtype=SAMPLE_DTYPE)
    s["source_column"] = cols
    s["multiplicity"] = mult
    s["inverse_probability"] = invp
    return s


def build_coordinate_sample(probabilities, repetitions: int, rng: Rng) -> np.ndarray:
    """Reference build_coordinate_sample (coordinate_sampling.cpp:14-45)."""
    p = np.ascontiguousarray(probabilities, dtype=np.float64)
    out = np.zeros(max(1, p.size), dtype=SAMPLE_DTYPE)
    n = c_uint64()
    check(_native.lib().uhann_build_coordinate_sample(rng.handle, f64ptr(p), p.size, repetitions,
                                                      out.ctypes.data_as(c_void_p), out.size, ctypes.byref(n)))
    return out[: n.value].copy()


def build_projection(sample: np.ndarray, projection_dimension: int, metric: int, rng: Rng) -> np.ndarray:
    """JL (L2) or Cauchy (L1) projection drawn after the sample; p x |I| binary64."""
    s = np.ascontiguousarray(sample, dtype=SAMPLE_DTYPE)
    out = np.zeros((projection_dimension, s.size), np.float64)
    check(_native.lib().uhann_build_projection(rng.handle, metric, s.ctypes.data_as(c_void_p), s.size,
                                               projection_dimension, f64ptr(out)))
    return out


def sampling_mass(probabilities) -> float:
    """Reference compute_sampling_mass (uniform_probabilities.cpp:12-26)."""
    p = np.ascontiguousarray(probabilities, dtype=np.float64)
    out = ctypes.c_double()
    check(_native.lib().uhann_sampling_mass(f64ptr(p), p.size, ctypes.byref(out)))
    return out.value


def uniform_probabilities(dimension: int, mass: float) -> np.ndarray:
    """Reference make_mass_matched_uniform_probabilities
    (uniform_probabilities.cpp:28-45)."""
    out = np.zeros(max(0, int(dimension)), np.float64)
    check(_native.lib().uhann_uniform_probabilities(int(dimension), float(mass), f64ptr(out)))
    return out


def _path(path) -> bytes:
    return os.fsencode(os.fspath(path))


def sha256_file(path) -> bytes:
    """SHA-256 of a file's bytes (reference io::sha256_file)."""
    out = ctypes.create_string_buffer(32)
    check(_native.lib().uhann_sha256_file(_path(path), out))
    return out.raw


class ProbabilityFile:
    """Contents of an importance-probability file (reference
    io::ImportanceProbabilityFile)."""

    def __init__(self, metric: int, representative_count: int, sha256: bytes, probabilities: np.ndarray):
        self.metric = metric
        self.representative_count = representative_count
        self.sha256 = sha256
        self.probabilities = probabilities


def save_importance_probabilities(path, metric: int, representative_count: int, sha256: bytes,
                                  probabilities) -> None:
    """Reference io::save_importance_probabilities (UAP v2 file)."""
    p = np.ascontiguousarray(probabilities, dtype=np.float64)
    digest = bytes(sha256)
    if len(digest) != 32:
        raise UhannInvalidArgument("the representative digest must be 32 bytes")
    check(_native.lib().uhann_save_probabilities(_path(path), metric, int(representative_count),
                                                 ctypes.create_string_buffer(digest, 32), f64ptr(p), p.size))


def load_importance_probabilities(path) -> ProbabilityFile:
    """Reference io::load_importance_probabilities."""
    metric, count, dim = c_int(), c_uint64(), c_uint64()
    digest = ctypes.create_string_buffer(32)
    lib = _native.lib()
    check(lib.uhann_load_probabilities(_path(path), ctypes.byref(metric), ctypes.byref(count), digest, None, 0,
                                       ctypes.byref(dim)))
    out = np.zeros(dim.value, np.float64)
    check(lib.uhann_load_probabilities(_path(path), ctypes.byref(metric), ctypes.byref(count), digest,
                                       f64ptr(out), out.size, ctypes.byref(dim)))
    return ProbabilityFile(metric.value, count.value, digest.raw, out)


def probability_source_matches(f: ProbabilityFile, metric: int, representative_count: int,
                               coordinate_count: int, sha256: bytes) -> bool:
    """Reference io::importance_probability_source_matches."""
    return (f.metric == metric and f.representative_count == representative_count
            and f.probabilities.size == coordinate_count and f.sha256 == bytes(sha256))


def importance_probabilities(metric: int, representatives) -> np.ndarray:
    refs = np.ascontiguousarray(representatives, dtype=np.float32)
    out = np.zeros(refs.shape[1], np.float64)
    check(_native.lib().uhann_importance_probabilities(metric, f32ptr(refs), refs.shape[0], refs.shape[1],
                                                       f64ptr(out)))
    return out


def importance_probabilities_device(metric: int, representatives, device: int = 0) -> np.ndarray:
    """GPU importance probabilities, bit-identical to importance_probabilities()."""
    refs = np.ascontiguousarray(representatives, dtype=np.float32)
    out = np.zeros(refs.shape[1], np.float64)
    check(_native.lib().uhann_importance_probabilities_device(metric, f32ptr(refs), refs.shape[0], refs.shape[1],
                                                              device, f64ptr(out)))
    return out


class IndexWorkspace:
    def __init__(self, handle, owner):
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

    def status(self) -> None:
        """Raise the data error left by asynchronous batches, if any."""
        check(_native.lib().uhann_workspace_status(self._h))

    def set_timing(self, enable: bool) -> None:
        """HIP-event stage timing of later calls on this workspace."""
        check(_native.lib().uhann_workspace_set_timing(self._h, int(bool(enable))))

    def stage_times(self) -> list:
        """Stage times (ms) of the last timed call: the direct scan kernel alone
        in [0]; for the mixed strategies conversion, Gram, selection + re-rank,
        escalation."""
        ms = (ctypes.c_float * 4)()
        check(_native.lib().uhann_workspace_stage_times(self._h, ms, 4))
        return list(ms)

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

    def __del__(self):
        try:
            if self._h:
                _native.lib().uhann_workspace_destroy(self._h)
                self._h = None
        except Exception:
            pass


class _Index:
    def _adopt(self, handle, dim):
        self._h = handle
        self.dim = dim

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                _native.lib().uhann_index_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def space_usage(self) -> dict:
        kind, metric = c_int(), c_int()
        rows, dim, uq, mult, payload = (c_uint64() for _ in range(5))
        check(_native.lib().uhann_index_info(self._h, ctypes.byref(kind), ctypes.byref(metric), ctypes.byref(rows),
                                             ctypes.byref(dim), ctypes.byref(uq), ctypes.byref(mult),
                                             ctypes.byref(payload)))
        return {"kind": kind.value, "metric": metric.value, "rows": rows.value, "dim": dim.value,
                "unique_query_coordinates": uq.value, "sampled_multiplicity": mult.value,
                "index_payload_bytes": payload.value}

    def enable_mixed(self, strategy=STRATEGY_GRAM_BF16):
        check(_native.lib().uhann_index_enable_mixed(self._h, strategy))

    def autotune(self, device_queries, max_batch: int, workspace: IndexWorkspace, ld_queries=None):
        """Tune the direct scan per batch bucket on max_batch device queries."""
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        table = np.zeros((24, 6), np.int32)
        n = c_int()
        check(_native.lib().uhann_index_autotune(
            self._h, workspace.handle, c_void_p(ptr), self.dim if ld_queries is None else ld_queries, max_batch,
            table.ctypes.data_as(ctypes.POINTER(c_int)), 24, ctypes.byref(n)))
        return [{"max_batch": 1 << int(r[0]), "rpt": int(r[1]), "tq": int(r[2]), "tr": int(r[3]),
                 "group_q": int(r[4]), "us": int(r[5])} for r in table[: n.value]]

    def make_query_workspace(self, maximum_batch_size: int) -> IndexWorkspace:
        h = c_void_p()
        check(_native.lib().uhann_index_workspace_create(self._h, maximum_batch_size, ctypes.byref(h)))
        return IndexWorkspace(h, self)

    def query_batch(self, queries, count: int, workspace: IndexWorkspace, strategy=STRATEGY_DIRECT):
        q = np.ascontiguousarray(queries, dtype=np.float32).reshape(-1)
        if q.size != count * self.dim:
            raise UhannInvalidArgument("query batch dimensions do not match original dimension")
        idx = np.zeros(count, np.uint64)
        dist = np.zeros(count, np.float32)
        check(_native.lib().uhann_index_query_batch(self._h, workspace.handle, f32ptr(q), count, u64ptr(idx),
                                                    f32ptr(dist), strategy))
        return idx, dist

    def query(self, query, workspace: IndexWorkspace, strategy=STRATEGY_DIRECT) -> int:
        return int(self.query_batch(query, 1, workspace, strategy)[0][0])

    def query_device_batch(self, device_queries, count: int, workspace: IndexWorkspace, strategy=STRATEGY_DIRECT):
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        idx = np.zeros(count, np.uint64)
        dist = np.zeros(count, np.float32)
        check(_native.lib().uhann_index_query_device_batch(self._h, workspace.handle, c_void_p(ptr), count,
                                                           u64ptr(idx), f32ptr(dist), strategy))
        return idx, dist

    def query_device_async(self, device_queries, count: int, workspace: IndexWorkspace, out_idx, out_dist,
                           strategy=STRATEGY_DIRECT, stream=None, ld_queries=None):
        ptr = device_queries.data_ptr() if hasattr(device_queries, "data_ptr") else int(device_queries)
        s = None if stream is None else c_void_p(int(getattr(stream, "cuda_stream", stream)))
        check(_native.lib().uhann_index_query_device_async(
            self._h, workspace.handle, c_void_p(ptr), self.dim if ld_queries is None else ld_queries, count,
            c_void_p(out_idx.data_ptr()) if out_idx is not None else None,
            c_void_p(out_dist.data_ptr()) if out_dist is not None else None, strategy, s))


def _refs(input_matrix):
    refs = np.ascontiguousarray(input_matrix, dtype=np.float32)
    if refs.ndim != 2:
        raise UhannInvalidArgument("input must be a 2-D matrix")
    return refs


class HipExactIndex(_Index):
    """Exact index on the index ABI (uhann_exact_create)."""

    def __init__(self, input_matrix, metric=METRIC_L2, device: int = 0):
        refs = _refs(input_matrix)
        h = c_void_p()
        check(_native.lib().uhann_exact_create(metric, f32ptr(refs), refs.shape[0], refs.shape[1], device,
                                               ctypes.byref(h)))
        self._adopt(h, refs.shape[1])


class _SampledCoordinate(_Index):
    _metric = METRIC_L2

    def __init__(self, input_matrix, sample, device: int = 0):
        refs = _refs(input_matrix)
        s = np.ascontiguousarray(sample, dtype=SAMPLE_DTYPE)
        self.sample = s
        h = c_void_p()
        check(_native.lib().uhann_flat_create(self._metric, f32ptr(refs), refs.shape[0], refs.shape[1],
                                              s.ctypes.data_as(c_void_p), s.size, device, ctypes.byref(h)))
        self._adopt(h, refs.shape[1])


class HipSampledCoordinateL2AnnIndex(_SampledCoordinate):
    """Reference CudaSampledCoordinateL2AnnIndex(input, sample, device)."""

    _metric = METRIC_L2


class HipSampledCoordinateL1AnnIndex(_SampledCoordinate):
    """Reference CudaSampledCoordinateL1AnnIndex(input, sample, device)."""

    _metric = METRIC_L1


def _flat_sample(refs, probabilities, repetitions, rng):
    p = np.ascontiguousarray(probabilities, dtype=np.float64)
    if p.size != refs.shape[1]:
        raise UhannInvalidArgument("sampling probability dimension does not match input")
    return build_coordinate_sample(p, repetitions, rng)


class HipFlatL2AnnIndex(HipSampledCoordinateL2AnnIndex):
    """Reference CudaFlatL2AnnIndex(input, probabilities, T, engine, device)."""

    def __init__(self, input_matrix, probabilities, repetitions: int, rng: Rng, device: int = 0):
        refs = _refs(input_matrix)
        super().__init__(refs, _flat_sample(refs, probabilities, repetitions, rng), device)


class HipFlatL1AnnIndex(HipSampledCoordinateL1AnnIndex):
    """Reference CudaFlatL1AnnIndex(input, probabilities, T, engine, device)."""

    def __init__(self, input_matrix, probabilities, repetitions: int, rng: Rng, device: int = 0):
        refs = _refs(input_matrix)
        super().__init__(refs, _flat_sample(refs, probabilities, repetitions, rng), device)


class HipUniformL2AnnIndex(HipSampledCoordinateL2AnnIndex):
    """Reference UniformL2AnnIndex(input, sampling_mass, T, engine)
    (uniform_l2_ann_index.cpp:11-21): the mass-matched uniform probabilities'
    coordinate sample, scanned on the GPU as the reference's benchmark does
    (nearest_neighbor_benchmark.cpp:755-768)."""

    def __init__(self, input_matrix, sampling_mass: float, repetitions: int, rng: Rng, device: int = 0):
        refs = _refs(input_matrix)
        probs = uniform_probabilities(refs.shape[1], sampling_mass)
        super().__init__(refs, build_coordinate_sample(probs, repetitions, rng), device)


class HipUniformL1AnnIndex(HipSampledCoordinateL1AnnIndex):
    """Reference UniformL1AnnIndex(input, sampling_mass, T, engine)
    (uniform_l1_ann_index.cpp:11-23), scanned on the GPU."""

    def __init__(self, input_matrix, sampling_mass: float, repetitions: int, rng: Rng, device: int = 0):
        refs = _refs(input_matrix)
        probs = uniform_probabilities(refs.shape[1], sampling_mass)
        super().__init__(refs, build_coordinate_sample(probs, repetitions, rng), device)


class _Hierarchical(_Index):
    _metric = METRIC_L2

    def __init__(self, input_matrix, probabilities, repetitions: int, projection_dimension: int, rng: Rng,
                 device: int = 0, sample=None):
        refs = _refs(input_matrix)
        # validate_index_parameters runs before any randomness is drawn
        # (cuda_hierarchical_l2_ann_index.cpp:49-59, 61-73).
        if refs.shape[0] == 0:
            raise UhannInvalidArgument("the hierarchical index expects at least one representative")
        if projection_dimension <= 0:
            raise UhannInvalidArgument("projection_dimension has to be a positive integer")
        if sample is None:
            sample = _flat_sample(refs, probabilities, repetitions, rng)
        s = np.ascontiguousarray(sample, dtype=SAMPLE_DTYPE)
        if s.size == 0 and refs.shape[0] > 1:
            raise _native.UhannError("importance sampling selected no coordinates")
        proj = build_projection(s, projection_dimension, self._metric, rng)
        self.sample, self.projection = s, proj
        h = c_void_p()
        check(_native.lib().uhann_hierarchical_create(self._metric, f32ptr(refs), refs.shape[0], refs.shape[1],
                                                      s.ctypes.data_as(c_void_p), s.size,
                                                      f64ptr(np.ascontiguousarray(proj)), projection_dimension,
                                                      device, ctypes.byref(h)))
        self._adopt(h, refs.shape[1])


class HipHierarchicalL2AnnIndex(_Hierarchical):
    """Reference CudaHierarchicalL2AnnIndex(input, probabilities, T, p, engine, device)."""

    _metric = METRIC_L2


class HipHierarchicalL1AnnIndex(_Hierarchical):
    """Reference CudaHierarchicalL1AnnIndex(input, probabilities, T, p, engine, device)."""

    _metric = METRIC_L1
