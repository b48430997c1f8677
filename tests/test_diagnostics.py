"""CPU checks of the exact-distance diagnostics (ultrahigh_ann_amd.diagnostics)
against the reference's own ExactDistanceTable / evaluate /
evaluate_selected_distances (src/benchmarks/benchmark_metrics.cpp, compiled
into oracle/_ref), on tables with zero optima, tied optima and every margin
bucket. The table itself comes from the oracle's binary64 CPU-order distances
(the GPU table is compared with the same oracle in tests/test_gpu_scan.py)."""

import math

import numpy as np
import pytest

from conftest import rng

import ultrahigh_ann_amd as uh
from ultrahigh_ann_amd import diagnostics

NAN = float("nan")


def _summary(s):
    if s is None:
        return [0.0, NAN, NAN, NAN, NAN, NAN]
    return [float(s["count"])] + [NAN if s[k] is None else s[k]
                                  for k in ("mean", "median", "percentile_95", "percentile_99", "maximum")]


def _buckets(bs):
    out = []
    for i in range(5):
        if i < len(bs):
            b = bs[i]
            out += [b["lower_inclusive"], NAN if b["upper_exclusive"] is None else b["upper_exclusive"],
                    float(b["query_count"]), float(b["non_optimal_count"]), float(b["ratio_eligible_count"]),
                    NAN if b["mean_distance_ratio"] is None else b["mean_distance_ratio"]]
        else:
            out += [NAN] * 6
    return out


def _metrics(m):
    out = [float(m[k]) for k in ("query_count", "optimal_representative_count", "non_optimal_count",
                                 "reference_improvement_count", "zero_optimum_query_count",
                                 "zero_optimum_non_optimal_count", "conditional_non_optimal_ratio_count")]
    out += [NAN if m[k] is None else m[k] for k in ("conditional_mean_distance_ratio",
                                                    "conditional_mean_relative_excess")]
    for f in m["approximation_failures"]:
        out += [f["epsilon"], float(f["violation_count"])]
    out += _summary(m["distance_ratio"]) + _summary(m["returned_representative_rank"]) + _buckets(m["margin_buckets"])
    return np.array(out)


def _geometry(g):
    out = [float(g[k]) for k in ("query_count", "zero_optimum_query_count", "non_unique_optimum_query_count",
                                 "infinite_multiplicative_margin_count")]
    return np.array(out + _summary(g["finite_multiplicative_margin"]) + _buckets(g["margin_buckets"]))


def _case(seed, r, d, b):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 2).astype(np.float32)
    refs[1] = refs[0]                      # a duplicated row: tied optima
    q = (g.random((b, d), dtype=np.float32) * 2).astype(np.float32)
    q[0] = refs[0]                         # zero optimum, tied (rows 0 and 1): margin 1
    q[1] = refs[5]                         # zero optimum, unique: infinite margin
    q[2] = refs[3] + np.float32(0.001)     # margin far above 1.25
    for k in range(3, min(b, 12)):         # near-ties between two rows: small margins
        w = np.float32(0.5 + 0.02 * k)
        q[k] = w * refs[6] + (1 - w) * refs[7]
    return refs, q


@pytest.mark.parametrize("metric_l1", [False, True])
@pytest.mark.parametrize("r,d,b", [(40, 300, 50), (9, 17, 30)])
def test_diagnostics_match_reference(oracle_mod, ref_mod, metric_l1, r, d, b):
    refs, q = _case(r + d + b, r, d, b)
    t = oracle_mod.fp64_table(metric_l1, refs, q)
    table = t if metric_l1 else np.sqrt(t)
    mine = diagnostics.ExactDistanceTable.from_table(table)
    g = rng(3)
    exact = table.argmin(1).astype(np.uint64)
    pred = np.where(g.random(b) < 0.5, exact, g.integers(0, r, b)).astype(np.uint64)
    ref_pred = np.where(g.random(b) < 0.7, exact, g.integers(0, r, b)).astype(np.uint64)
    geo, ev, sel = np.zeros(40), np.zeros(63), np.zeros(63)
    assert ref_mod.ref().ref_distance_diagnostics(int(metric_l1), ref_mod._p(refs, ref_mod._f32p), r, d,
                                                  ref_mod._p(q, ref_mod._f32p), b, ref_mod._p(pred, ref_mod._u64p),
                                                  ref_mod._p(ref_pred, ref_mod._u64p), ref_mod._p(geo, ref_mod._f64p),
                                                  ref_mod._p(ev, ref_mod._f64p), ref_mod._p(sel, ref_mod._f64p)) == 0
    np.testing.assert_array_equal(_geometry(mine.query_geometry()), geo)
    np.testing.assert_array_equal(_metrics(mine.evaluate(pred)), ev)
    np.testing.assert_array_equal(_metrics(uh.evaluate_selected_distances(mine, ref_pred, pred)), sel)
    # the case reaches zero optima, ties, an infinite margin and several buckets
    geom = mine.query_geometry()
    assert geom["zero_optimum_query_count"] >= 2 and geom["non_unique_optimum_query_count"] >= 1
    assert geom["infinite_multiplicative_margin_count"] >= 1
    assert sum(bk["query_count"] > 0 for bk in geom["margin_buckets"]) >= 3


def test_summary_and_margin_rules():
    s = diagnostics.summarize([3.0, 1.0, 2.0, 10.0])
    assert s == {"count": 4, "mean": 4.0, "median": 2.0, "percentile_95": 10.0, "percentile_99": 10.0,
                 "maximum": 10.0}
    assert diagnostics.summarize([])["mean"] is None
    assert diagnostics.multiplicative_margin(2.0, 3.0) == 1.5
    assert diagnostics.multiplicative_margin(0.0, 0.0) == 1.0
    assert math.isinf(diagnostics.multiplicative_margin(0.0, 1.0))
    assert [diagnostics.margin_bucket_index(m) for m in (1.0, 1.0099, 1.01, 1.1, 1.25, math.inf)] == [0, 0, 1, 3, 4, 4]
    with pytest.raises(uh.UhannLogicError):
        diagnostics.margin_bucket_index(0.99)
    t = diagnostics.ExactDistanceTable.from_table(np.array([[1.0, 2.0]]))
    with pytest.raises(uh.UhannInvalidArgument):
        t.evaluate([2])
    with pytest.raises(uh.UhannInvalidArgument):
        t.evaluate([0, 1])
    with pytest.raises(uh.UhannInvalidArgument):
        diagnostics.ExactDistanceTable.from_table(np.zeros((1, 0)))
