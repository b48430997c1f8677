"""Benchmark setups and schema-6 reports on the MI355X engine.

Runs a benchmark setup in the reference's format (ultrahigh_ann_benchmark_setup_v2,
src/benchmarks/benchmark_setup.cpp) on the HIP engine and writes the
reference's JSON report (schema version 6, nearest_neighbor_benchmark.cpp:
1302-1531) and per-trial CSV (:1546-1640), so the reference's reporting and
figure scripts read MI355X results like their own.

What maps to what:
  * the dataset directory holds the reference's NPY files (reference_vectors.npy
    and queries.npy as little-endian binary32 C-order matrices, optional
    reference_labels.npy / query_labels.npy as uint16 vectors;
    src/datasets/representative_query_dataset.cpp:196-330,
    src/benchmarks/benchmark_dataset.cpp:14-80), each identified by its SHA-256;
  * probabilities: policy "load" reads a UAP v2 file bound to the dataset;
    every other policy computes the binary64 importance probabilities on the
    GPU (bit-identical to the reference's sequential CPU policy, so the same
    samples; the reference's gpu_fp32 policies compute binary32 values);
  * runs with backend=cuda run on this engine (the drop-in for the reference's
    CUDA path): exact / flat / uniform / hierarchical, L1 or L2. Strategy
    "direct" is the binary32 path in the reference's order; "gemm" is the
    certified INT8 fixed-point Gram with exact re-rank for exact and
    flat/uniform L2 and the BF16 MFMA projection for hierarchical L2
    (hierarchical L1 stays binary32). backend=cpu runs are the reference's CPU path, which this
    engine does not ship: a setup with one is rejected before any work;
  * measurements: per batch size (clamped to the query count) the warmups and
    timed trials of a full pass over the queries in batches, host queries in,
    host indices out, as the reference's run_cuda_batched
    (nearest_neighbor_benchmark.cpp:468-495, 386-413);
  * diagnostics: none, selected_distances or full_distance_table, from the
    GPU binary64 distance table (diagnostics.py).
"""

from __future__ import annotations

import csv
import datetime as _dt
import io
import json
import os
import platform
import subprocess
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import diagnostics as _diag
from ._native import UhannError, UhannInvalidArgument

FORMAT_HEADER = "ultrahigh_ann_benchmark_setup_v2"
SCHEMA_VERSION = 6
PROJECT_VERSION = "0.2.0"
_INDEXES = ("exact", "flat", "uniform", "hierarchical")
_POLICIES = ("sequential", "cpu_parallel", "gpu_fp32", "gpu_cublas_fp32", "load")
_DIAGNOSTICS = ("none", "selected_distances", "full_distance_table")
_STRATEGIES = ("sequential", "parallel_queries", "parallel_reference_vectors", "automatic", "direct", "gemm")
_DIRECTIVES = {"dataset", "json_output", "csv_output", "reference", "distance", "max_queries", "warmups", "trials",
               "batch_sizes", "device", "diagnostics", "probability_policy", "probabilities",
               "reference_vectors_file", "reference_labels_file", "queries_file", "query_labels_file"}


class SetupError(UhannError):
    """A malformed benchmark setup (the reference's parse errors)."""


@dataclass
class Run:
    name: str
    index: str
    backend: str = "cpu"
    strategy: str = "sequential"
    batch_sizes: list = field(default_factory=lambda: [1])
    warmups: int = 0
    trials: int = 1
    repetitions: int = 0
    seed: int = 0
    projection_dimension: int = 0


@dataclass
class Setup:
    path: Path
    dataset_directory: Path
    json_output: Path
    csv_output: Path
    reference: str
    reference_vectors_path: Path
    reference_labels_path: Path
    queries_path: Path
    query_labels_path: Path
    distance: str = "l2"
    probability_policy: str = "sequential"
    probabilities_path: Path | None = None
    diagnostics: str = "selected_distances"
    maximum_queries: int = 0
    warmups: int = 0
    trials: int = 1
    batch_sizes: list = field(default_factory=lambda: [1])
    device: int = 0
    runs: list = field(default_factory=list)


def _err(path, line, message) -> SetupError:
    return SetupError(f"{path}:{line}: {message}")


def _integer(text, path, line, name, allow_zero):
    try:
        if not text or not text.isdigit():
            raise ValueError
        value = int(text)
    except ValueError:
        raise _err(path, line, f"{name} must be a {'nonnegative' if allow_zero else 'positive'} integer") from None
    if not allow_zero and value == 0:
        raise _err(path, line, f"{name} must be a positive integer")
    return value


def _batch_sizes(text, path, line):
    values = [_integer(part.strip(), path, line, "batch_sizes", False) for part in text.split(",")]
    if not values or len(values) != len(set(values)):
        raise _err(path, line, "batch_sizes must be unique positive integers")
    return values


def _valid_name(name: str) -> bool:
    return bool(name) and all(c.isascii() and (c.isalnum() or c in "_-.") for c in name)


def parse_setup(path) -> Setup:
    """load_benchmark_setup (benchmark_setup.cpp:458-690): the header line, tab
    separated directives, then `run <name> <index> key=value ...` lines."""
    path = Path(path).absolute()
    try:
        text = path.read_text(encoding="utf-8")
    except OSError as error:
        raise SetupError(f"cannot open benchmark setup: {path}") from error
    header = False
    directives: dict = {}
    raw_runs = []
    for number, raw in enumerate(text.splitlines(), start=1):
        line = raw.strip()
        if not line or line.startswith("#"):
            continue
        if not header:
            if line != FORMAT_HEADER:
                raise _err(path, number, f"expected {FORMAT_HEADER}")
            header = True
            continue
        fields = [f.strip() for f in raw.split("\t")]
        if fields[0] == "run":
            raw_runs.append((number, fields))
            continue
        if len(fields) != 2 or not fields[1]:
            raise _err(path, number, f"{fields[0]} requires one value")
        if fields[0] in directives:
            raise _err(path, number, f"duplicate directive: {fields[0]}")
        directives[fields[0]] = (fields[1], number)
    if not header:
        raise SetupError(f"benchmark setup is empty: {path}")
    for key, (_, number) in directives.items():
        if key not in _DIRECTIVES:
            raise _err(path, number, f"unknown directive: {key}")
    for required in ("dataset", "json_output", "csv_output", "reference"):
        if required not in directives:
            raise SetupError(f"benchmark setup does not define {required}: {path}")

    def value(key, default=None):
        return directives[key][0] if key in directives else default

    def line(key):
        return directives[key][1]

    def resolve(p):
        q = Path(p)
        return Path(os.path.normpath(q if q.is_absolute() else path.parent / q))

    dataset = resolve(value("dataset"))

    def dataset_file(key, default):
        q = Path(value(key, default))
        return Path(os.path.normpath(q if q.is_absolute() else dataset / q))

    s = Setup(path=path, dataset_directory=dataset, json_output=resolve(value("json_output")),
              csv_output=resolve(value("csv_output")), reference=value("reference"),
              reference_vectors_path=dataset_file("reference_vectors_file", "reference_vectors.npy"),
              reference_labels_path=dataset_file("reference_labels_file", "reference_labels.npy"),
              queries_path=dataset_file("queries_file", "queries.npy"),
              query_labels_path=dataset_file("query_labels_file", "query_labels.npy"))
    if "distance" in directives:
        if value("distance") not in ("l1", "l2"):
            raise _err(path, line("distance"), "distance must be l1 or l2")
        s.distance = value("distance")
    if "max_queries" in directives:
        s.maximum_queries = _integer(value("max_queries"), path, line("max_queries"), "max_queries", True)
    if "warmups" in directives:
        s.warmups = _integer(value("warmups"), path, line("warmups"), "warmups", True)
    if "trials" in directives:
        s.trials = _integer(value("trials"), path, line("trials"), "trials", False)
    if "batch_sizes" in directives:
        s.batch_sizes = _batch_sizes(value("batch_sizes"), path, line("batch_sizes"))
    if "device" in directives:
        s.device = _integer(value("device"), path, line("device"), "device", True)
    if "diagnostics" in directives:
        if value("diagnostics") not in _DIAGNOSTICS:
            raise _err(path, line("diagnostics"), "diagnostics must be none, selected_distances or full_distance_table")
        s.diagnostics = value("diagnostics")
    if "probability_policy" in directives:
        if value("probability_policy") not in _POLICIES:
            raise _err(path, line("probability_policy"), "unknown probability_policy")
        s.probability_policy = value("probability_policy")
    if "probabilities" in directives:
        s.probabilities_path = resolve(value("probabilities"))
        if "probability_policy" not in directives:
            s.probability_policy = "load"
    if s.probability_policy == "load" and s.probabilities_path is None:
        raise SetupError("probability_policy=load requires probabilities")
    if s.distance == "l1" and s.probability_policy == "gpu_cublas_fp32":
        raise _err(path, line("probability_policy"), "gpu_cublas_fp32 is unavailable for L1 probability computation")

    names = set()
    for number, fields in raw_runs:
        run = _parse_run(fields, s, path, number)
        if run.name in names:
            raise _err(path, number, f"duplicate run name: {run.name}")
        names.add(run.name)
        s.runs.append(run)
    if not s.runs:
        raise SetupError("benchmark setup does not define any runs")
    ref = [r for r in s.runs if r.name == s.reference]
    if not ref:
        raise SetupError("reference does not name a configured run")
    if ref[0].index != "exact":
        raise SetupError("reference run must use the exact index")
    return s


def _parse_run(fields, s: Setup, path, number) -> Run:
    """parse_run (benchmark_setup.cpp:249-380)."""
    if len(fields) < 3:
        raise _err(path, number, "run requires a name and index kind")
    if not _valid_name(fields[1]):
        raise _err(path, number, "run name may contain only letters, digits, '.', '_', and '-'")
    if fields[2] not in _INDEXES:
        raise _err(path, number, "index must be exact, flat, uniform, or hierarchical")
    run = Run(name=fields[1], index=fields[2], batch_sizes=list(s.batch_sizes), warmups=s.warmups, trials=s.trials)
    params = {}
    for f in fields[3:]:
        key, sep, val = f.partition("=")
        if not sep or not key or not val:
            raise _err(path, number, "run parameters must be key=value")
        if key in params:
            raise _err(path, number, f"duplicate run parameter: {key}")
        params[key] = val
    allowed = {"backend", "strategy", "batch_sizes", "warmups", "trials", "repetitions", "seed", "projection_dimension"}
    for key in params:
        if key not in allowed:
            raise _err(path, number, f"unsupported run parameter: {key}")
    if "backend" in params:
        if params["backend"] not in ("cpu", "cuda"):
            raise _err(path, number, "backend must be cpu or cuda")
        run.backend = params["backend"]
    run.strategy = ("direct" if s.distance == "l1" else "gemm") if run.backend == "cuda" else "sequential"
    if "strategy" in params:
        if params["strategy"] not in _STRATEGIES:
            raise _err(path, number, "unknown strategy")
        run.strategy = params["strategy"]
    if "batch_sizes" in params:
        run.batch_sizes = _batch_sizes(params["batch_sizes"], path, number)
    if "warmups" in params:
        run.warmups = _integer(params["warmups"], path, number, "warmups", True)
    if "trials" in params:
        run.trials = _integer(params["trials"], path, number, "trials", False)
    approximate = run.index != "exact"
    if approximate and not {"repetitions", "seed"} <= params.keys():
        raise _err(path, number, "approximate runs require repetitions and seed")
    if not approximate and ({"repetitions", "seed"} & params.keys()):
        raise _err(path, number, "exact runs do not accept repetitions or seed")
    if approximate:
        run.repetitions = _integer(params["repetitions"], path, number, "repetitions", False)
        run.seed = _integer(params["seed"], path, number, "seed", True)
    if run.index == "hierarchical":
        if "projection_dimension" not in params:
            raise _err(path, number, "hierarchical runs require projection_dimension")
        run.projection_dimension = _integer(params["projection_dimension"], path, number, "projection_dimension", False)
    elif "projection_dimension" in params:
        raise _err(path, number, "projection_dimension requires a hierarchical run")
    if run.backend == "cuda":
        if run.strategy not in ("direct", "gemm"):
            raise _err(path, number, "CUDA strategy must be direct or gemm")
        if s.distance == "l1" and run.index != "hierarchical" and run.strategy != "direct":
            raise _err(path, number, "CUDA L1 exact, flat, and uniform strategies must be direct")
    else:
        if run.strategy in ("direct", "gemm"):
            raise _err(path, number, "CPU runs cannot use direct or gemm strategy")
        if (s.distance == "l1" or run.index == "hierarchical") and run.strategy != "sequential":
            raise _err(path, number, "CPU L1 and hierarchical runs require sequential strategy")
    return run


# ---- dataset ---------------------------------------------------------------

def load_npy(path, descriptor: str, ndim: int) -> np.ndarray:
    """NPY arrays as the reference reads them: the exact dtype descriptor,
    C order, the expected rank, nothing after the data."""
    path = Path(path)
    try:
        with open(path, "rb") as f:
            version = np.lib.format.read_magic(f)
            if version == (1, 0):
                shape, fortran, dtype = np.lib.format.read_array_header_1_0(f)
            elif version == (2, 0):
                shape, fortran, dtype = np.lib.format.read_array_header_2_0(f)
            else:
                raise UhannInvalidArgument(f"{path}: unsupported NPY format version")
            if dtype.str != descriptor:
                raise UhannInvalidArgument(f"{path}: expected dtype {descriptor}, found {dtype.str}")
            if fortran:
                raise UhannInvalidArgument(f"{path}: Fortran-order arrays are not supported")
            if len(shape) == 0:
                raise UhannInvalidArgument(f"{path}: scalar NPY arrays are not supported")
            if len(shape) != ndim:
                raise UhannInvalidArgument(f"{path}: expected a {'two' if ndim == 2 else 'one'}-dimensional array")
            count = int(np.prod(shape, dtype=np.uint64))
            data = np.fromfile(f, dtype=dtype, count=count)
            if data.size != count:
                raise UhannInvalidArgument(f"{path}: truncated while reading NPY array data")
            if f.read(1):
                raise UhannInvalidArgument(f"{path}: unexpected trailing data")
    except OSError as error:
        raise UhannInvalidArgument(f"{path}: file does not exist or cannot be opened") from error
    return data.reshape(shape)


@dataclass
class Dataset:
    refs: np.ndarray
    queries: np.ndarray
    ref_labels: np.ndarray
    query_labels: np.ndarray
    labels_available: bool
    refs_sha256: bytes
    queries_sha256: bytes
    ref_labels_sha256: bytes | None
    query_labels_sha256: bytes | None
    hash_ms: float
    load_ms: float


def load_dataset(s: Setup) -> Dataset:
    """load_benchmark_dataset (benchmark_dataset.cpp:14-80)."""
    from .indexes import sha256_file

    has_ref_labels = s.reference_labels_path.is_file()
    if has_ref_labels != s.query_labels_path.is_file():
        raise UhannError("reference-vector and query labels must either both exist or both be absent")
    t0 = time.perf_counter()
    refs_sha, q_sha = sha256_file(s.reference_vectors_path), sha256_file(s.queries_path)
    rl_sha = sha256_file(s.reference_labels_path) if has_ref_labels else None
    ql_sha = sha256_file(s.query_labels_path) if has_ref_labels else None
    hash_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    refs = load_npy(s.reference_vectors_path, "<f4", 2)
    queries = load_npy(s.queries_path, "<f4", 2)
    for name, m in (("reference vectors", refs), ("queries", queries)):
        if not np.all(np.isfinite(m)):
            raise UhannInvalidArgument(f"{name}: matrix contains a non-finite value")
    if refs.shape[0] == 0 or refs.shape[1] == 0 or queries.shape[0] == 0 or refs.shape[1] != queries.shape[1]:
        raise UhannError("reference vectors and queries must be nonempty matrices with the same dimension")
    if has_ref_labels:
        ref_labels = load_npy(s.reference_labels_path, "<u2", 1).astype(np.int64)
        query_labels = load_npy(s.query_labels_path, "<u2", 1).astype(np.int64)
        if ref_labels.size != refs.shape[0] or query_labels.size != queries.shape[0]:
            raise UhannError("label counts do not match the matrix row counts")
    else:
        ref_labels = np.arange(refs.shape[0], dtype=np.int64)
        query_labels = np.zeros(queries.shape[0], dtype=np.int64)
    load_ms = (time.perf_counter() - t0) * 1e3
    return Dataset(refs, queries, ref_labels, query_labels, has_ref_labels, refs_sha, q_sha, rl_sha, ql_sha,
                   hash_ms, load_ms)


# ---- runs --------------------------------------------------------------------

@dataclass
class Measurement:
    batch_size: int
    trial_ms: list
    predictions: np.ndarray
    workspace_payload_bytes: int
    correct: int = 0
    agreements: int = 0
    label_agreements: int = 0
    approximation: dict | None = None


@dataclass
class Execution:
    run: Run
    build_ms: float
    space: dict
    measurements: list
    device: int | None
    device_name: str | None


def _effective_batches(run: Run, query_count: int) -> list:
    out = []
    for b in run.batch_sizes:
        b = min(b, query_count)
        if b not in out:
            out.append(b)
    return out


def _strategy(uh, run: Run, metric_l1: bool) -> int:
    if run.strategy == "direct":
        return uh.STRATEGY_DIRECT
    if run.index == "hierarchical":
        return uh.STRATEGY_DIRECT if metric_l1 else uh.STRATEGY_PROJECTION_BF16
    return uh.STRATEGY_GRAM_INT8


def _build_index(uh, s: Setup, run: Run, data: Dataset, probs, uniform):
    l1 = s.distance == "l1"
    metric = uh.METRIC_L1 if l1 else uh.METRIC_L2
    if run.index == "exact":
        index = uh.HipExactIndex(data.refs, metric, s.device)
    else:
        rng = uh.Rng(run.seed)
        if run.index == "hierarchical":
            cls = uh.HipHierarchicalL1AnnIndex if l1 else uh.HipHierarchicalL2AnnIndex
            index = cls(data.refs, probs, run.repetitions, run.projection_dimension, rng, s.device)
        else:
            p = uniform if run.index == "uniform" else probs
            sample = uh.build_coordinate_sample(p, run.repetitions, rng)
            cls = uh.HipSampledCoordinateL1AnnIndex if l1 else uh.HipSampledCoordinateL2AnnIndex
            index = cls(data.refs, sample, s.device)
    if run.strategy == "gemm" and run.index != "hierarchical":
        index.enable_mixed(uh.STRATEGY_GRAM_INT8)  # the BF16 projection needs no preparation
    return index


def execute_run(uh, s: Setup, run: Run, data: Dataset, query_count: int, probs, uniform) -> Execution:
    """execute_exact_run / execute_approximate_run for the GPU backend:
    build (timed), then per batch size warmups and timed trials of a full pass."""
    t0 = time.perf_counter()
    index = _build_index(uh, s, run, data, probs, uniform)
    build_ms = (time.perf_counter() - t0) * 1e3
    usage = index.space_usage()
    strategy = _strategy(uh, run, s.distance == "l1")
    q = np.ascontiguousarray(data.queries[:query_count])
    measurements = []
    workspace_bytes = 0
    for batch in _effective_batches(run, query_count):
        ws = index.make_query_workspace(batch)
        pred = np.zeros(query_count, np.int64)

        def one_pass():
            for b0 in range(0, query_count, batch):
                n = min(batch, query_count - b0)
                idx, _ = index.query_batch(q[b0:b0 + n], n, ws, strategy)
                pred[b0:b0 + n] = idx

        for _ in range(run.warmups):
            one_pass()
        times = []
        for _ in range(run.trials):
            t = time.perf_counter()
            one_pass()
            times.append((time.perf_counter() - t) * 1e3)
        workspace_bytes = ws.payload_bytes()
        measurements.append(Measurement(batch, times, pred.copy(), workspace_bytes))
    space = {"index_payload_bytes": usage["index_payload_bytes"], "query_workspace_payload_bytes": workspace_bytes,
             "unique_coordinates": usage["unique_query_coordinates"],
             "sampled_multiplicity": usage["sampled_multiplicity"]}
    return Execution(run, build_ms, space, measurements, s.device, uh.device_name(s.device))


def obtain_probabilities(uh, s: Setup, data: Dataset) -> dict:
    """obtain_probabilities (nearest_neighbor_benchmark.cpp:294-362)."""
    metric = uh.METRIC_L1 if s.distance == "l1" else uh.METRIC_L2
    if s.probability_policy == "load":
        t0 = time.perf_counter()
        source_sha = uh.sha256_file(s.probabilities_path)
        f = uh.load_importance_probabilities(s.probabilities_path)
        if not uh.probability_source_matches(f, metric, data.refs.shape[0], data.refs.shape[1], data.refs_sha256):
            raise UhannError("probability metadata or representative-matrix digest does not match the configured "
                             "dataset; regenerate the probability file")
        return {"probabilities": f.probabilities, "mass": uh.sampling_mass(f.probabilities), "build_ms": 0.0,
                "load_ms": (time.perf_counter() - t0) * 1e3, "source": s.probabilities_path,
                "source_sha256": source_sha, "refs_sha256": f.sha256, "gpu": None}
    t0 = time.perf_counter()
    p = uh.importance_probabilities_device(metric, data.refs, s.device)
    build_ms = (time.perf_counter() - t0) * 1e3
    gpu = {"device": s.device, "device_name": uh.device_name(s.device), "distance_backend": "direct",
           "arithmetic": "fp64", "pair_count": data.refs.shape[0] * (data.refs.shape[0] - 1) // 2, "pair_chunks": 1,
           "device_working_set_bytes": data.refs.nbytes + 8 * data.refs.shape[0] ** 2,
           "host_to_device_ms": None, "inverse_distance_ms": None, "coordinate_maximum_ms": None,
           "device_to_host_ms": None, "total_ms": build_ms}
    return {"probabilities": p, "mass": uh.sampling_mass(p), "build_ms": build_ms, "load_ms": 0.0, "source": None,
            "source_sha256": None, "refs_sha256": data.refs_sha256, "gpu": gpu}


def add_quality(data: Dataset, query_count: int, reference: np.ndarray, executions: list) -> None:
    """add_basic_quality (nearest_neighbor_benchmark.cpp:842-865)."""
    for e in executions:
        for m in e.measurements:
            p = m.predictions
            m.agreements = int((p == reference).sum())
            if data.labels_available:
                m.correct = int((data.ref_labels[p] == data.query_labels[:query_count]).sum())
                m.label_agreements = int((data.ref_labels[p] == data.ref_labels[reference]).sum())


def add_diagnostics(uh, s: Setup, data: Dataset, query_count: int, reference: np.ndarray, executions: list) -> dict:
    """add_diagnostics (nearest_neighbor_benchmark.cpp:867-945); the exact
    distances come from the GPU binary64 table."""
    out = {"build_ms": 0.0, "evaluation_ms": 0.0, "payload_bytes": 0, "geometry": None}
    if s.diagnostics == "none":
        return out
    import torch

    t0 = time.perf_counter()
    exact = (uh.HipExactL1Index if s.distance == "l1" else uh.HipExactL2Index)(data.refs, device=s.device)
    dq = torch.from_numpy(np.ascontiguousarray(data.queries[:query_count])).to(f"cuda:{s.device}")
    table = _diag.ExactDistanceTable(exact, dq, query_count)
    built = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for e in executions:
        if e.run.name == s.reference:
            continue
        first = None
        for m in e.measurements:
            if first is not None and np.array_equal(m.predictions, first.predictions):
                m.approximation = first.approximation
            elif s.diagnostics == "selected_distances":
                m.approximation = _diag.evaluate_selected_distances(table, reference, m.predictions)
            else:
                m.approximation = table.evaluate(m.predictions)
            first = first or m
    evaluated = (time.perf_counter() - t0) * 1e3
    if s.diagnostics == "selected_distances":
        out["evaluation_ms"] = built + evaluated
    else:
        out.update(build_ms=built, evaluation_ms=evaluated, payload_bytes=table.payload_bytes(),
                   geometry=table.query_geometry())
    return out


# ---- report ------------------------------------------------------------------

def _portable(p: Path) -> str:
    p = Path(p).absolute()
    try:
        rel = os.path.relpath(p)
        if not rel.startswith(".."):
            return Path(rel).as_posix()
    except ValueError:
        pass
    return p.as_posix()


def _median(values):
    v = sorted(values)
    n = len(v)
    return (v[n // 2 - 1] + v[n // 2]) / 2.0 if n % 2 == 0 else v[n // 2]


def _distribution(d):
    return {k: d[k] for k in ("count", "mean", "median", "percentile_95", "percentile_99", "maximum")}


def _approximation_json(m: dict | None):
    """write_approximation (nearest_neighbor_benchmark.cpp:1080-1139)."""
    if m is None:
        return None
    n = float(m["query_count"])
    rank = m["returned_representative_rank"]
    return {
        "distance_optimal_count": m["optimal_representative_count"],
        "distance_optimal_rate": m["optimal_representative_count"] / n,
        "non_optimal_count": m["non_optimal_count"],
        "non_optimal_rate": m["non_optimal_count"] / n,
        "reference_improvement_count": m["reference_improvement_count"],
        "distance_ratio": _distribution(m["distance_ratio"]),
        "zero_optimum_query_count": m["zero_optimum_query_count"],
        "zero_optimum_non_optimal_count": m["zero_optimum_non_optimal_count"],
        "non_optimal_ratio_count": m["conditional_non_optimal_ratio_count"],
        "non_optimal_distance_ratio_mean": m["conditional_mean_distance_ratio"],
        "non_optimal_relative_excess_mean": m["conditional_mean_relative_excess"],
        "approximation_guarantee_failures": [
            {"epsilon": f["epsilon"], "violation_count": f["violation_count"], "violation_rate": f["violation_count"] / n}
            for f in m["approximation_failures"]],
        "returned_neighbor_rank": None if rank is None or rank["count"] == 0 else _distribution(rank),
        "by_multiplicative_margin": [
            {k: b[k] for k in ("lower_inclusive", "upper_exclusive", "query_count", "non_optimal_count",
                               "ratio_eligible_count", "mean_distance_ratio")} for b in m["margin_buckets"]],
    }


def _geometry_json(g):
    if g is None:
        return None
    return {"query_count": g["query_count"], "zero_optimum_query_count": g["zero_optimum_query_count"],
            "non_unique_optimum_query_count": g["non_unique_optimum_query_count"],
            "finite_multiplicative_margin": _distribution(g["finite_multiplicative_margin"]),
            "infinite_multiplicative_margin_count": g["infinite_multiplicative_margin_count"],
            "margin_buckets": [{"lower_inclusive": b["lower_inclusive"], "upper_exclusive": b["upper_exclusive"],
                                "query_count": b["query_count"]} for b in g["margin_buckets"]]}


def _git(root: Path) -> dict:
    try:
        commit = subprocess.run(["git", "-C", str(root), "rev-parse", "HEAD"], capture_output=True, text=True,
                                check=True, timeout=10).stdout.strip()
        dirty = bool(subprocess.run(["git", "-C", str(root), "status", "--porcelain", "--untracked-files=no"],
                                    capture_output=True, text=True, check=True, timeout=10).stdout.strip())
        return {"available": True, "commit": commit, "dirty": dirty}
    except (OSError, subprocess.SubprocessError):
        return {"available": False, "commit": None, "dirty": None}


def _tool_version(cmd) -> str:
    try:
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=30).stdout
        return next((ln.strip() for ln in out.splitlines() if ln.strip()), "unknown")
    except (OSError, subprocess.SubprocessError):
        return "unknown"


def _provenance(uh, s: Setup) -> dict:
    """write_provenance (nearest_neighbor_benchmark.cpp:1196-1292). The GPU
    build and runtime fill the report's "cuda" fields (the GPU backend):
    hipcc for gfx950 and the HIP runtime."""
    from .indexes import sha256_file

    root = Path(__file__).resolve().parent.parent
    hipcc = _tool_version(["/opt/rocm/bin/hipcc", "--version"])
    try:
        import torch

        props = torch.cuda.get_device_properties(s.device)
        hip_version = str(torch.version.hip or "unknown")
        encoded = 0
        parts = [int(x) for x in hip_version.split(".")[:2] if x.isdigit()]
        if len(parts) == 2:
            encoded = parts[0] * 10_000_000 + parts[1] * 100_000
        runtime = {"device": s.device, "device_name": uh.device_name(s.device),
                   "compute_capability": getattr(props, "gcnArchName", "gfx950").split(":")[0],
                   "total_global_memory_bytes": int(props.total_memory),
                   "compiled_runtime": {"encoded": max(encoded, 1), "version": hip_version},
                   "runtime": {"encoded": max(encoded, 1), "version": hip_version},
                   "driver": {"encoded": max(encoded, 1), "version": hip_version}}
    except Exception:  # noqa: BLE001 -- no GPU: the field is null, as for a CPU-only build
        runtime = None
    mem = None
    try:
        mem = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError, AttributeError):
        pass
    model = None
    try:
        model = next((ln.split(":", 1)[1].strip() for ln in Path("/proc/cpuinfo").read_text().splitlines()
                      if ln.startswith("model name")), None)
    except OSError:
        pass
    return {
        "source": {"project_version": PROJECT_VERSION, "git": _git(root)},
        "build": {"cmake_version": "none (make + hipcc)", "build_type": "Release",
                  "cxx": {"standard": 20, "compiler_id": "Clang", "compiler_version": hipcc,
                          "flags": "-std=c++20 -O3 -ffp-contract=off"},
                  "cuda": {"enabled": True, "compiler_id": "hipcc", "compiler_version": hipcc,
                           "architectures": "gfx950",
                           "flags": "--offload-arch=gfx950 -fno-gpu-flush-denormals-to-zero -ffp-contract=off"}},
        "host": {"operating_system": {"name": platform.system(), "version": platform.release(),
                                      "architecture": platform.machine()},
                 "processor_model": model, "physical_memory_bytes": mem},
        "cuda_runtime": runtime,
        "invocation": {"arguments": ["--setup", _portable(s.path)], "setup_sha256": sha256_file(s.path).hex()},
    }


def build_report(uh, s: Setup, data: Dataset, query_count: int, probs: dict | None, diag: dict,
                 executions: list) -> dict:
    """write_json_report (nearest_neighbor_benchmark.cpp:1302-1531) as a dict."""
    def hexd(d):
        return None if d is None else d.hex()

    prob_required = any(r.index != "exact" for r in s.runs)
    p = probs or {"probabilities": np.zeros(0), "mass": 0.0, "build_ms": 0.0, "load_ms": 0.0, "source": None,
                  "source_sha256": None, "refs_sha256": None, "gpu": None}
    runs = []
    for e in executions:
        ms = []
        for m in e.measurements:
            med = _median(m.trial_ms)
            ms.append({
                "batch_size": m.batch_size, "workspace_payload_bytes": m.workspace_payload_bytes,
                "trial_ms": list(m.trial_ms), "median_ms": med,
                "median_microseconds_per_query": med * 1000.0 / query_count,
                "median_queries_per_second": query_count * 1000.0 / med,
                "exact_neighbor_agreement_count": m.agreements,
                "exact_neighbor_agreement": m.agreements / query_count,
                "correct": m.correct if data.labels_available else None,
                "accuracy": m.correct / query_count if data.labels_available else None,
                "label_agreement_with_exact": m.label_agreements / query_count if data.labels_available else None,
                "approximation": _approximation_json(m.approximation),
            })
        r = e.run
        runs.append({"name": r.name, "index": r.index, "backend": r.backend, "strategy": r.strategy,
                     "reference": r.name == s.reference, "repetitions": r.repetitions, "seed": r.seed,
                     "projection_dimension": r.projection_dimension, "warmups": r.warmups, "trials": r.trials,
                     "build_ms": e.build_ms, "space": e.space,
                     "device": None if e.device is None else {"index": e.device, "name": e.device_name},
                     "measurements": ms})
    return {
        "schema_version": SCHEMA_VERSION,
        "generated_at_utc": _dt.datetime.now(_dt.timezone.utc).strftime("%Y-%m-%dT%H:%M:%SZ"),
        "setup_file": _portable(s.path),
        "provenance": _provenance(uh, s),
        "outputs": {"json": _portable(s.json_output), "csv": _portable(s.csv_output)},
        "dataset": {
            "directory": _portable(s.dataset_directory),
            "reference_vectors_file": _portable(s.reference_vectors_path),
            "reference_labels_file": _portable(s.reference_labels_path) if data.labels_available else None,
            "queries_file": _portable(s.queries_path),
            "query_labels_file": _portable(s.query_labels_path) if data.labels_available else None,
            "reference_vectors_sha256": data.refs_sha256.hex(), "reference_labels_sha256": hexd(data.ref_labels_sha256),
            "queries_sha256": data.queries_sha256.hex(), "query_labels_sha256": hexd(data.query_labels_sha256),
            "reference_vector_count": int(data.refs.shape[0]), "query_count_available": int(data.queries.shape[0]),
            "query_count_run": query_count, "dimension": int(data.refs.shape[1]),
            "labels_available": data.labels_available, "hash_ms": data.hash_ms, "load_ms": data.load_ms},
        "settings": {"distance": s.distance, "reference_run": s.reference, "max_queries": s.maximum_queries,
                     "device": s.device, "diagnostics": s.diagnostics},
        "sampling_probabilities": {
            "required": prob_required, "policy": s.probability_policy,
            "source_file": None if p["source"] is None else _portable(p["source"]),
            "source_sha256": hexd(p["source_sha256"]), "reference_vectors_sha256": hexd(p["refs_sha256"]),
            "build_ms": p["build_ms"], "load_ms": p["load_ms"], "coordinate_count": int(np.size(p["probabilities"])),
            "sampling_mass": p["mass"], "gpu": p["gpu"]},
        "diagnostic_execution": {"build_ms": diag["build_ms"], "evaluation_ms": diag["evaluation_ms"],
                                 "payload_bytes": diag["payload_bytes"], "query_geometry": _geometry_json(diag["geometry"])},
        "runs": runs,
    }


_CSV_HEADER = ("run_name,index,backend,strategy,reference,repetitions,seed,projection_dimension,batch_size,warmups,"
               "trial,reference_vectors,dimension,queries,build_ms,elapsed_ms,microseconds_per_query,"
               "queries_per_second,correct,accuracy,exact_neighbor_agreement,label_agreement_with_exact,"
               "distance_optimal_rate,distance_ratio_mean,distance_ratio_median,distance_ratio_p95,"
               "distance_ratio_p99,distance_ratio_max,violation_rate_eps_0_001,violation_rate_eps_0_005,"
               "violation_rate_eps_0_01,violation_rate_eps_0_02,violation_rate_eps_0_05,violation_rate_eps_0_10,"
               "index_payload_bytes,workspace_payload_bytes,unique_coordinates,sampled_multiplicity,device,"
               "device_name")


def csv_text(report: dict) -> str:
    """write_csv_report (nearest_neighbor_benchmark.cpp:1546-1640): one row per trial."""
    buf = io.StringIO()
    buf.write(_CSV_HEADER + "\n")
    w = csv.writer(buf, lineterminator="\n")
    ds = report["dataset"]
    q = ds["query_count_run"]

    def opt(v):
        return "" if v is None else repr(float(v))

    for run in report["runs"]:
        for m in run["measurements"]:
            a = m["approximation"]
            for t, ms in enumerate(m["trial_ms"]):
                row = [run["name"], run["index"], run["backend"], run["strategy"], int(run["reference"]),
                       run["repetitions"], run["seed"], run["projection_dimension"], m["batch_size"], run["warmups"],
                       t, ds["reference_vector_count"], ds["dimension"], q, repr(run["build_ms"]), repr(ms),
                       repr(ms * 1000.0 / q), repr(q * 1000.0 / ms),
                       "" if m["correct"] is None else m["correct"], opt(m["accuracy"]),
                       repr(m["exact_neighbor_agreement"]), opt(m["label_agreement_with_exact"])]
                if a is None:
                    row += [""] * 12
                else:
                    d = a["distance_ratio"]
                    row += [repr(a["distance_optimal_rate"])] + [opt(d[k]) for k in ("mean", "median", "percentile_95",
                                                                                       "percentile_99", "maximum")]
                    row += [repr(f["violation_rate"]) for f in a["approximation_guarantee_failures"]]
                dev = run["device"]
                row += [run["space"]["index_payload_bytes"], m["workspace_payload_bytes"],
                        run["space"]["unique_coordinates"], run["space"]["sampled_multiplicity"],
                        "" if dev is None else dev["index"], "" if dev is None else dev["name"]]
                w.writerow(row)
    return buf.getvalue()


def run_setup(path, echo=print) -> dict:
    """run_benchmark (nearest_neighbor_benchmark.cpp:1692-1781): load, run the
    reference first, then the other runs in setup order, add quality and
    diagnostics, write the JSON report and the CSV trials."""
    import ultrahigh_ann_amd as uh

    s = parse_setup(path)
    cpu = [r.name for r in s.runs if r.backend != "cuda"]
    if cpu:
        raise UhannError("this engine runs the GPU backend only; backend=cpu runs are the reference's CPU path: "
                         + ", ".join(cpu))
    for label, p in (("reference vectors", s.reference_vectors_path), ("queries", s.queries_path)):
        if not p.is_file():
            raise UhannError(f"{label} file does not exist: {p}")
    data = load_dataset(s)
    query_count = data.queries.shape[0] if s.maximum_queries == 0 else min(s.maximum_queries, data.queries.shape[0])
    echo(f"Loaded {data.refs.shape[0]} reference vectors in {data.refs.shape[1]} dimensions and "
         f"{data.queries.shape[0]} queries; running {query_count}.")
    probs = uniform = None
    p = None
    if any(r.index != "exact" for r in s.runs):
        p = obtain_probabilities(uh, s, data)
        probs = p["probabilities"]
        uniform = uh.uniform_probabilities(data.refs.shape[1], p["mass"])
    order = [r for r in s.runs if r.name == s.reference] + [r for r in s.runs if r.name != s.reference]
    executions = []
    for i, run in enumerate(order):
        echo(f"[{i + 1}/{len(order)}] Building {run.name} ({run.backend} {run.index})...")
        executions.append(execute_run(uh, s, run, data, query_count, probs, uniform))
    reference = executions[0].measurements[0].predictions
    add_quality(data, query_count, reference, executions)
    diag = add_diagnostics(uh, s, data, query_count, reference, executions)
    report = build_report(uh, s, data, query_count, p, diag, executions)
    s.json_output.parent.mkdir(parents=True, exist_ok=True)
    s.json_output.write_text(json.dumps(report, indent=1) + "\n", encoding="utf-8")
    s.csv_output.parent.mkdir(parents=True, exist_ok=True)
    s.csv_output.write_text(csv_text(report), encoding="utf-8")
    echo(f"Wrote JSON report to {s.json_output}\nWrote CSV trials to {s.csv_output}")
    return report


def main(argv=None) -> int:
    import argparse

    ap = argparse.ArgumentParser(description="Run a benchmark setup on the MI355X engine (schema-6 report).")
    ap.add_argument("--setup", required=True)
    ap.add_argument("--validate-only", action="store_true")
    args = ap.parse_args(argv)
    if args.validate_only:
        s = parse_setup(args.setup)
        print(f"valid_setup={s.path}\nruns={len(s.runs)}\nreference={s.reference}\n"
              f"json_output={s.json_output}\ncsv_output={s.csv_output}")
        return 0
    run_setup(args.setup)
    return 0


if __name__ == "__main__":
    sys.exit(main())
