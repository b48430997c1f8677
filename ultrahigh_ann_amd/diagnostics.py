"""Exact-distance diagnostics of nearest-neighbour results.

Mirrors the reference's benchmark metrics (src/benchmarks/benchmark_metrics.cpp):
  * ExactDistanceTable (:112-192): one exact binary64 distance per
    query/reference pair -- L2 the square root of the squared distance, L1 the
    L1 distance, both in the reference CPU order -- and per query the optimum,
    the runner-up (the smallest distance at any other row than the first
    minimum), the multiplicative margin runner-up / optimum and its bucket;
  * ExactDistanceTable::evaluate (:199-300): for a prediction vector, the
    optimal count, the rank of the returned row, distance ratios, the
    conditional relative excess of the non-optimal answers, the epsilon
    violations and per-margin-bucket figures;
  * evaluate_selected_distances (:302-391): the same ratios for a candidate
    run against a reference run's picks.
The B x r table is computed on the GPU (uhann_scan_distance_table, the O(B r d)
part); the reductions over it run on the host in the reference's order
(sequential sums in query order, nearest-rank quantiles over sorted values),
so every figure equals the reference's for the same table.
"""

from __future__ import annotations

import math

import numpy as np

from ._native import UhannInvalidArgument, UhannLogicError

APPROXIMATION_EPSILONS = (0.001, 0.005, 0.01, 0.02, 0.05, 0.10)
# (lower inclusive, upper exclusive or None), benchmark_metrics.cpp:22-28
MARGIN_BUCKETS = ((1.00, 1.01), (1.01, 1.05), (1.05, 1.10), (1.10, 1.25), (1.25, None))


def _sequential_sum(values) -> float:
    """Left-to-right binary64 sum (std::accumulate)."""
    v = np.asarray(values, dtype=np.float64)
    return float(np.cumsum(v)[-1]) if v.size else 0.0


def _nearest_rank(sorted_values: np.ndarray, probability: float) -> float:
    n = sorted_values.size
    rank = int(math.ceil(probability * float(n)))
    return float(sorted_values[max(rank, 1) - 1])


def summarize(values) -> dict:
    """DistributionSummary (benchmark_metrics.cpp:56-70): count, mean of the
    sorted values, nearest-rank median / p95 / p99, maximum; None when empty."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    out = {"count": int(v.size), "mean": None, "median": None, "percentile_95": None, "percentile_99": None,
           "maximum": None}
    if v.size == 0:
        return out
    out["mean"] = _sequential_sum(v) / float(v.size)
    out["median"] = _nearest_rank(v, 0.50)
    out["percentile_95"] = _nearest_rank(v, 0.95)
    out["percentile_99"] = _nearest_rank(v, 0.99)
    out["maximum"] = float(v[-1])
    return out


def margin_bucket_index(margin: float) -> int:
    for i, (lo, hi) in enumerate(MARGIN_BUCKETS):
        if margin >= lo and (hi is None or margin < hi):
            return i
    raise UhannLogicError("multiplicative margin is smaller than one")


def _make_buckets() -> list:
    return [{"lower_inclusive": lo, "upper_exclusive": hi, "query_count": 0, "non_optimal_count": 0,
             "ratio_eligible_count": 0, "mean_distance_ratio": None} for lo, hi in MARGIN_BUCKETS]


def multiplicative_margin(optimum: float, second: float) -> float:
    if optimum > 0.0:
        return second / optimum
    if second == 0.0:
        return 1.0
    return math.inf


def _new_metrics(query_count: int) -> dict:
    return {
        "query_count": query_count,
        "optimal_representative_count": 0,
        "non_optimal_count": 0,
        "reference_improvement_count": 0,
        "zero_optimum_query_count": 0,
        "zero_optimum_non_optimal_count": 0,
        "conditional_non_optimal_ratio_count": 0,
        "conditional_mean_distance_ratio": None,
        "conditional_mean_relative_excess": None,
        "approximation_failures": [{"epsilon": e, "violation_count": 0} for e in APPROXIMATION_EPSILONS],
        "distance_ratio": None,
        "returned_representative_rank": None,
        "margin_buckets": [],
    }


class ExactDistanceTable:
    """Reference ExactDistanceTable(metric, representatives, queries,
    query_count), the table computed on the GPU of an exact index.

    ExactDistanceTable(index, device_queries, query_count) with index an
    unweighted HipExactL2Index / HipExactL1Index (or any dense scan) and
    device_queries a float32 CUDA tensor of at least query_count rows."""

    def __init__(self, index, device_queries, query_count: int, chunk_rows: int = 2048):
        import torch

        if index.rows == 0:
            raise UhannInvalidArgument("distance diagnostics expect at least one reference vector")
        if query_count <= 0 or query_count > device_queries.shape[0]:
            raise UhannInvalidArgument("distance diagnostics expect a valid positive query count")
        if device_queries.shape[1] != index.dim:
            raise UhannInvalidArgument("reference-vector and query dimensions differ")
        r = index.rows
        table = np.empty((query_count, r), np.float64)
        step = max(1, min(chunk_rows, query_count))
        buf = torch.empty((step, r), dtype=torch.float64, device=device_queries.device)
        for s in range(0, query_count, step):
            n = min(step, query_count - s)
            index.distance_table(device_queries[s:s + n], n, buf)
            table[s:s + n] = buf[:n].cpu().numpy()
        self._init_from_table(table)

    @classmethod
    def from_table(cls, table) -> "ExactDistanceTable":
        """The same diagnostics over a given binary64 table (query x reference)."""
        self = cls.__new__(cls)
        t = np.ascontiguousarray(table, dtype=np.float64)
        if t.ndim != 2 or t.shape[1] == 0:
            raise UhannInvalidArgument("distance diagnostics expect at least one reference vector")
        if t.shape[0] == 0:
            raise UhannInvalidArgument("distance diagnostics expect a valid positive query count")
        self._init_from_table(t)
        return self

    def _init_from_table(self, table: np.ndarray) -> None:
        q, r = table.shape
        self.distances = table
        first = table.argmin(axis=1)  # the first minimum, as std::ranges::min_element
        rows = np.arange(q)
        optimum = table[rows, first]
        optimum_count = (table == optimum[:, None]).sum(axis=1)
        if r > 1:
            others = table.copy()
            others[rows, first] = np.inf
            second = others.min(axis=1)
        else:
            second = np.full(q, np.inf)
        margins = np.array([multiplicative_margin(float(o), float(s)) for o, s in zip(optimum, second)])
        self.optimum = optimum
        self.margins = margins
        buckets = _make_buckets()
        finite = []
        infinite = 0
        for m in margins:
            buckets[margin_bucket_index(float(m))]["query_count"] += 1
            if math.isinf(m):
                infinite += 1
            else:
                finite.append(float(m))
        self._geometry = {
            "query_count": q,
            "zero_optimum_query_count": int((optimum == 0.0).sum()),
            "non_unique_optimum_query_count": int((optimum_count > 1).sum()),
            "infinite_multiplicative_margin_count": infinite,
            "finite_multiplicative_margin": summarize(finite),
            "margin_buckets": buckets,
        }

    def query_geometry(self) -> dict:
        return self._geometry

    def payload_bytes(self) -> int:
        return self.distances.size * 8 + self.optimum.size * 8 + self.margins.size * 8

    def evaluate(self, predictions) -> dict:
        """ExactDistanceTable::evaluate (benchmark_metrics.cpp:199-300)."""
        pred = np.asarray(predictions, dtype=np.int64)
        q, r = self.distances.shape
        if pred.shape != (q,):
            raise UhannInvalidArgument("prediction count does not match exact distance table")
        if np.any(pred < 0) or np.any(pred >= r):
            raise UhannInvalidArgument("prediction contains an out-of-range reference-vector row")
        res = _new_metrics(q)
        res["margin_buckets"] = _make_buckets()
        returned = self.distances[np.arange(q), pred]
        ranks = 1 + (self.distances < returned[:, None]).sum(axis=1)
        ratios = []
        cond_sum = 0.0
        bucket_sums = [0.0] * len(MARGIN_BUCKETS)
        for i in range(q):
            opt = float(self.optimum[i])
            ret = float(returned[i])
            optimal = ret == opt
            b = margin_bucket_index(float(self.margins[i]))
            bucket = res["margin_buckets"][b]
            bucket["query_count"] += 1
            res["optimal_representative_count"] += int(optimal)
            res["non_optimal_count"] += int(not optimal)
            bucket["non_optimal_count"] += int(not optimal)
            if opt == 0.0:
                res["zero_optimum_query_count"] += 1
                res["zero_optimum_non_optimal_count"] += int(not optimal)
            else:
                ratio = ret / opt
                ratios.append(ratio)
                bucket["ratio_eligible_count"] += 1
                bucket_sums[b] += ratio
                if not optimal:
                    cond_sum += ratio
                    res["conditional_non_optimal_ratio_count"] += 1
            for f in res["approximation_failures"]:
                violation = (not optimal) if opt == 0.0 else ret / opt > 1.0 + f["epsilon"]
                f["violation_count"] += int(violation)
        res["distance_ratio"] = summarize(ratios)
        res["returned_representative_rank"] = summarize(ranks.astype(np.float64))
        n = res["conditional_non_optimal_ratio_count"]
        if n:
            mean = cond_sum / float(n)
            res["conditional_mean_distance_ratio"] = mean
            res["conditional_mean_relative_excess"] = mean - 1.0
        for b, bucket in enumerate(res["margin_buckets"]):
            if bucket["ratio_eligible_count"]:
                bucket["mean_distance_ratio"] = bucket_sums[b] / float(bucket["ratio_eligible_count"])
        return res


def evaluate_selected_distances(table: ExactDistanceTable, reference_predictions, candidate_predictions) -> dict:
    """evaluate_selected_distances (benchmark_metrics.cpp:302-391): the
    candidate run's distances against the reference run's, per query, from the
    exact distances of the two selected rows."""
    ref = np.asarray(reference_predictions, dtype=np.int64)
    cand = np.asarray(candidate_predictions, dtype=np.int64)
    q, r = table.distances.shape
    if ref.shape != (q,) or cand.shape != (q,):
        raise UhannInvalidArgument("selected-distance diagnostics received incompatible query data")
    if np.any(ref < 0) or np.any(ref >= r) or np.any(cand < 0) or np.any(cand >= r):
        raise UhannInvalidArgument("selected-distance diagnostics received an invalid prediction")
    res = _new_metrics(q)
    rows = np.arange(q)
    optimum = table.distances[rows, ref]
    returned = table.distances[rows, cand]
    ratios = []
    cond_sum = 0.0
    for i in range(q):
        opt = float(optimum[i])
        ret = float(returned[i])
        worse = ret > opt
        res["optimal_representative_count"] += int(ret <= opt)
        res["non_optimal_count"] += int(worse)
        res["reference_improvement_count"] += int(ret < opt)
        if opt == 0.0:
            res["zero_optimum_query_count"] += 1
            res["zero_optimum_non_optimal_count"] += int(worse)
        else:
            ratio = ret / opt
            ratios.append(ratio)
            if worse:
                cond_sum += ratio
                res["conditional_non_optimal_ratio_count"] += 1
        for f in res["approximation_failures"]:
            violation = worse if opt == 0.0 else ret / opt > 1.0 + f["epsilon"]
            f["violation_count"] += int(violation)
    res["distance_ratio"] = summarize(ratios)
    n = res["conditional_non_optimal_ratio_count"]
    if n:
        mean = cond_sum / float(n)
        res["conditional_mean_distance_ratio"] = mean
        res["conditional_mean_relative_excess"] = mean - 1.0
    return res


__all__ = [
    "APPROXIMATION_EPSILONS",
    "MARGIN_BUCKETS",
    "ExactDistanceTable",
    "evaluate_selected_distances",
    "margin_bucket_index",
    "multiplicative_margin",
    "summarize",
]
