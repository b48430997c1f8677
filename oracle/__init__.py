"""TEST INFRASTRUCTURE -- the CPU oracle, not part of the engine.

ctypes bindings of oracle/liboracle.so (the C++ restatement of the
reference's query-path arithmetic in uhann_oracle.cpp) and, when it has been
built, oracle/_ref/libuhann_ref.so (the reference's own CPU sources compiled
here, see oracle/Makefile). Only tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py import this package, and only as the checker.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_uint64

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libuhann_ref.so")
REF_ROOT = "/root/reference"

_u64p = POINTER(c_uint64)
_f32p = POINTER(c_float)
_f64p = POINTER(c_double)
_i32p = POINTER(c_int)


def build(with_ref: bool | None = None) -> None:
    """Compile the oracle (and the reference build when its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if with_ref is None:
        with_ref = os.path.isdir(REF_ROOT)
    if with_ref:
        subprocess.run(["make", "-s", "-C", HERE, "-j8", "ref"], check=True)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build(with_ref=False)
        L = ctypes.CDLL(ORACLE_SO)
        sig = {
            "orc_l2_fp64_scan": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _u64p, _f64p],
            "orc_l2_fp64_table": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _f64p],
            "orc_l1_fp64_scan": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _u64p, _f64p],
            "orc_l1_fp64_table": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _f64p],
            "orc_weighted_l1_fp64_scan": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, _u64p, _u64p,
                                          _f64p, c_uint64, c_int, _u64p, _f64p],
            "orc_k2_l2_distances": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _f32p],
            "orc_k5_l1_distances": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, _f32p, c_int, c_int,
                                    _f32p],
            "orc_fp32_argmin": [_f32p, c_uint64, c_uint64, _u64p, _f32p],
            "orc_k7_project": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, _f32p, _i32p],
            "orc_k9_median_l1": [_f32p, c_uint64, c_uint64, _f32p, c_uint64, _f32p, _i32p],
            "orc_build_coordinate_sample": [_f64p, c_uint64, c_uint64, c_uint64, _u64p, _u64p, _f64p],
            "orc_embedding_factors": [_u64p, _f64p, c_uint64, _f32p, _f32p],
            "orc_build_hierarchical": [c_int, _f32p, c_uint64, c_uint64, _f64p, c_uint64, c_uint64,
                                       c_uint64, _u64p, _u64p, _f64p, _f64p, _f32p, _f64p, _f32p],
            "orc_hierarchical_query_fp64": [c_int, _f64p, c_uint64, c_uint64, _f64p, c_uint64, _u64p,
                                            _f32p, c_uint64, c_uint64, c_int, _u64p, _f64p],
            "orc_importance_probabilities": [c_int, _f32p, c_uint64, c_uint64, _f64p],
            "orc_max_threads": [],
        }
        ret = {"orc_build_coordinate_sample": c_int64, "orc_build_hierarchical": c_int64,
               "orc_max_threads": c_int}
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ret.get(name, None)
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference's own CPU code (oracle/_ref), or raise if not built."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_SO)
        L = ctypes.CDLL(REF_SO)
        sig = {
            "ref_exact_l2": ([_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_int, _u64p], c_int),
            "ref_exact_l1": ([_f32p, c_uint64, c_uint64, _f32p, c_uint64, _u64p], c_int),
            "ref_exact_l2_timed": ([_f32p, c_uint64, c_uint64, _f32p, c_uint64, c_uint64, c_int, _u64p, _f64p],
                                   c_int),
            "ref_flat_l2_timed": ([_f32p, c_uint64, c_uint64, _f64p, c_uint64, c_uint64, _f32p, c_uint64, c_uint64,
                                   c_int, _u64p, _f64p], c_int),
            "ref_build_coordinate_sample": ([_f64p, c_uint64, c_uint64, c_uint64, _u64p, _u64p, _f64p],
                                            c_int64),
            "ref_flat": ([c_int, _f32p, c_uint64, c_uint64, _f64p, c_uint64, c_uint64, _f32p, c_uint64,
                          _u64p], c_int),
            "ref_hierarchical": ([c_int, _f32p, c_uint64, c_uint64, _f64p, c_uint64, c_uint64, c_uint64,
                                  _f32p, c_uint64, _u64p], c_int),
            "ref_hierarchical_fp32_build": ([c_int, _f32p, c_uint64, c_uint64, _f64p, c_uint64, c_uint64,
                                             c_uint64, _u64p, _u64p, _f64p, _f32p, _f32p], c_int64),
            "ref_importance_probabilities": ([c_int, _f32p, c_uint64, c_uint64, _f64p], c_int),
            "ref_sampling_mass": ([_f64p, c_uint64, _f64p], c_int),
            "ref_distance_diagnostics": ([c_int, _f32p, c_uint64, c_uint64, _f32p, c_uint64, _u64p, _u64p, _f64p,
                                          _f64p, _f64p], c_int),
            "ref_sha256_file": ([ctypes.c_char_p, ctypes.c_void_p], c_int),
            "ref_save_probabilities": ([ctypes.c_char_p, c_int, c_uint64, ctypes.c_void_p, _f64p, c_uint64], c_int),
            "ref_load_probabilities": ([ctypes.c_char_p, POINTER(c_int), _u64p, ctypes.c_void_p, _f64p, c_uint64,
                                        _u64p], c_int),
            "ref_uniform_probabilities": ([c_uint64, ctypes.c_double, _f64p], c_int),
            "ref_uniform": ([c_int, _f32p, c_uint64, c_uint64, ctypes.c_double, c_uint64, c_uint64, _f32p,
                             c_uint64, _u64p, _u64p], c_int),
        }
        for name, (args, r) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = r
        _ref = L
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def threads() -> int:
    return int(lib().orc_max_threads())


# ------------------------------------------------------------ FP64 CPU index


def l2_fp64_scan(refs, queries, nthreads=0):
    refs, queries = _f32(refs), _f32(queries)
    r, d = refs.shape
    b = queries.shape[0]
    idx = np.zeros(b, np.uint64)
    dist = np.zeros(b, np.float64)
    lib().orc_l2_fp64_scan(_p(refs, _f32p), r, d, _p(queries, _f32p), b, nthreads or threads(),
                           _p(idx, _u64p), _p(dist, _f64p))
    return idx, dist


def l1_fp64_scan(refs, queries, nthreads=0):
    refs, queries = _f32(refs), _f32(queries)
    r, d = refs.shape
    b = queries.shape[0]
    idx = np.zeros(b, np.uint64)
    dist = np.zeros(b, np.float64)
    lib().orc_l1_fp64_scan(_p(refs, _f32p), r, d, _p(queries, _f32p), b, nthreads or threads(),
                           _p(idx, _u64p), _p(dist, _f64p))
    return idx, dist


def fp64_table(metric_l1, refs, queries, nthreads=0):
    refs, queries = _f32(refs), _f32(queries)
    r, d = refs.shape
    b = queries.shape[0]
    out = np.zeros((b, r), np.float64)
    fn = lib().orc_l1_fp64_table if metric_l1 else lib().orc_l2_fp64_table
    fn(_p(refs, _f32p), r, d, _p(queries, _f32p), b, nthreads or threads(), _p(out, _f64p))
    return out


def weighted_l1_fp64_scan(sampled_refs, queries, cols, mult, invp, nthreads=0):
    s, queries = _f32(sampled_refs), _f32(queries)
    r, k = s.shape
    b, d = queries.shape
    cols = np.ascontiguousarray(cols, np.uint64)
    mult = np.ascontiguousarray(mult, np.uint64)
    invp = _f64(invp)
    idx = np.zeros(b, np.uint64)
    dist = np.zeros(b, np.float64)
    lib().orc_weighted_l1_fp64_scan(_p(s, _f32p), r, k, _p(queries, _f32p), d, _p(cols, _u64p),
                                    _p(mult, _u64p), _p(invp, _f64p), b, nthreads or threads(),
                                    _p(idx, _u64p), _p(dist, _f64p))
    return idx, dist


# ------------------------------------------------------- FP32 CUDA emulators


def k2_distances(refs, queries, nthreads=0):
    refs, queries = _f32(refs), _f32(queries)
    r, d = refs.shape
    b = queries.shape[0]
    out = np.zeros((b, r), np.float32)
    lib().orc_k2_l2_distances(_p(refs, _f32p), r, d, _p(queries, _f32p), b, nthreads or threads(),
                              _p(out, _f32p))
    return out


def k5_distances(refs, queries, weights=None, weight_mode=0, nthreads=0):
    refs, queries = _f32(refs), _f32(queries)
    r, d = refs.shape
    b = queries.shape[0]
    w = None if weights is None else _f32(weights)
    out = np.zeros((b, r), np.float32)
    lib().orc_k5_l1_distances(_p(refs, _f32p), r, d, _p(queries, _f32p), b, _p(w, _f32p),
                              weight_mode, nthreads or threads(), _p(out, _f32p))
    return out


def fp32_argmin(dist):
    dist = _f32(dist)
    b, r = dist.shape
    idx = np.zeros(b, np.uint64)
    best = np.zeros(b, np.float32)
    lib().orc_fp32_argmin(_p(dist, _f32p), b, r, _p(idx, _u64p), _p(best, _f32p))
    return idx, best


def k7_project(x, proj):
    x, proj = _f32(x), _f32(proj)
    b, k = x.shape
    p = proj.shape[0]
    out = np.zeros((b, p), np.float32)
    bad = c_int(0)
    lib().orc_k7_project(_p(x, _f32p), b, k, _p(proj, _f32p), p, _p(out, _f32p), ctypes.byref(bad))
    return out, bool(bad.value)


def k9_median_l1(refs_proj, queries_proj):
    rp, qp = _f32(refs_proj), _f32(queries_proj)
    r, p = rp.shape
    b = qp.shape[0]
    out = np.zeros((b, r), np.float32)
    bad = c_int(0)
    lib().orc_k9_median_l1(_p(rp, _f32p), r, p, _p(qp, _f32p), b, _p(out, _f32p), ctypes.byref(bad))
    return out, bool(bad.value)


# ------------------------------------------------------- index construction


def build_coordinate_sample(probs, reps, seed):
    probs = _f64(probs)
    d = probs.shape[0]
    cols = np.zeros(d, np.uint64)
    mult = np.zeros(d, np.uint64)
    invp = np.zeros(d, np.float64)
    n = lib().orc_build_coordinate_sample(_p(probs, _f64p), d, reps, seed, _p(cols, _u64p),
                                          _p(mult, _u64p), _p(invp, _f64p))
    if n < 0:
        raise ValueError("invalid sampling probability")
    return cols[:n].copy(), mult[:n].copy(), invp[:n].copy()


def embedding_factors(mult, invp):
    mult = np.ascontiguousarray(mult, np.uint64)
    invp = _f64(invp)
    k = mult.shape[0]
    scales = np.zeros(k, np.float32)
    weights = np.zeros(k, np.float32)
    lib().orc_embedding_factors(_p(mult, _u64p), _p(invp, _f64p), k, _p(scales, _f32p),
                                _p(weights, _f32p))
    return scales, weights


def build_hierarchical(metric_l1, refs, probs, reps, p, seed):
    """Returns dict(cols, mult, invp, p64, p32, proj_refs64, proj_refs32)."""
    refs = _f32(refs)
    probs = _f64(probs)
    r, d = refs.shape
    cols = np.zeros(d, np.uint64)
    mult = np.zeros(d, np.uint64)
    invp = np.zeros(d, np.float64)
    p64 = np.zeros(p * d, np.float64)
    p32 = np.zeros(p * d, np.float32)
    pr64 = np.zeros((r, p), np.float64)
    pr32 = np.zeros((r, p), np.float32)
    k = lib().orc_build_hierarchical(int(metric_l1), _p(refs, _f32p), r, d, _p(probs, _f64p), reps, p,
                                     seed, _p(cols, _u64p), _p(mult, _u64p), _p(invp, _f64p),
                                     _p(p64, _f64p), _p(p32, _f32p), _p(pr64, _f64p), _p(pr32, _f32p))
    if k < 0:
        raise ValueError(f"hierarchical build failed ({k})")
    return {
        "cols": cols[:k].copy(), "mult": mult[:k].copy(), "invp": invp[:k].copy(),
        "p64": p64[: p * k].reshape(p, k).copy(), "p32": p32[: p * k].reshape(p, k).copy(),
        "proj_refs64": pr64, "proj_refs32": pr32,
    }


def hierarchical_query_fp64(metric_l1, built, queries, nthreads=0):
    queries = _f32(queries)
    b, d = queries.shape
    pr = _f64(built["proj_refs64"])
    r, p = pr.shape
    p64 = _f64(built["p64"])
    cols = np.ascontiguousarray(built["cols"], np.uint64)
    idx = np.zeros(b, np.uint64)
    dist = np.zeros(b, np.float64)
    lib().orc_hierarchical_query_fp64(int(metric_l1), _p(pr, _f64p), r, p, _p(p64, _f64p),
                                      cols.shape[0], _p(cols, _u64p), _p(queries, _f32p), d, b,
                                      nthreads or threads(), _p(idx, _u64p), _p(dist, _f64p))
    return idx, dist


def importance_probabilities(metric_l1, refs):
    refs = _f32(refs)
    r, d = refs.shape
    out = np.zeros(d, np.float64)
    lib().orc_importance_probabilities(int(metric_l1), _p(refs, _f32p), r, d, _p(out, _f64p))
    return out
