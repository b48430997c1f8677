"""The reference's own CPU-runnable case (config 1): a flat L2 importance-
sampling index over 33 class centroids (mean of 90 samples each), d = 60660,
3013 held-out queries, T = 8, sampling seeds 7, 20, 42, 1312, 2026, with exact
full-d L2 as the comparison.

For every seed the GPU flat index must return the oracle's emulation of the
reference's FP32 CUDA path bit for bit (K2 over the embedded coordinates), and
agree with the reference CPU path (binary64 over the same embedded values)
except where that path's top-two gap lies inside the FP32 error band; the
agreement-with-exact and relative-distance-excess figures computed from the
GPU answers must equal the CPU path's up to those near ties."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32
SEEDS = (7, 20, 42, 1312, 2026)


@pytest.fixture(scope="module")
def workload():
    from ultrahigh_ann_amd import synthetic

    cents, q, _ = synthetic.centroid_workload(seed=2026, device="cuda")
    return cents.cpu().numpy(), q.cpu().numpy()


def _excess(refs, q, idx, exact_idx):
    """100 (||q - returned|| / ||q - exact|| - 1) with binary64 distances."""
    r64, q64 = refs.astype(np.float64), q.astype(np.float64)
    a = np.sqrt(((r64[idx.astype(np.int64)] - q64) ** 2).sum(1))
    b = np.sqrt(((r64[exact_idx.astype(np.int64)] - q64) ** 2).sum(1))
    return 100.0 * np.where(b > 0, a / b - 1.0, 0.0)


def test_config1_flat_r33(uh, oracle_mod, workload):
    refs, q = workload
    b, d = q.shape
    probs = uh.importance_probabilities(uh.METRIC_L2, refs)
    # exact comparison: GPU FP32 exact index (the reference's direct CUDA bits) vs binary64
    ex = uh.HipExactIndex(refs)
    ex_ws = ex.make_query_workspace(b)
    exact_gpu, _ = ex.query_batch(q, b, ex_ws)
    exact_cpu, _ = oracle_mod.l2_fp64_scan(refs, q)
    # the exact GPU answer is the reference's binary32 argmin (K2 order) ...
    k2_full = oracle_mod.k2_distances(refs, q)
    np.testing.assert_array_equal(exact_gpu, oracle_mod.fp32_argmin(k2_full)[0])
    # ... and differs from the binary64 argmin only inside the binary32 tie band
    full = oracle_mod.fp64_table(False, refs, q)
    rows_all = np.arange(b)
    kf = np.ceil(d / 256) + 10
    gf = kf * 2.0**-24 / (1 - kf * 2.0**-24)
    off = exact_gpu != exact_cpu
    assert np.all(full[rows_all[off], exact_gpu[off].astype(np.int64)]
                  <= full[rows_all[off], exact_cpu[off].astype(np.int64)] * (1 + 2 * gf))
    for seed in SEEDS:
        index = uh.HipFlatL2AnnIndex(refs, probs, 8, uh.Rng(seed))
        ws = index.make_query_workspace(b)
        idx, dist = index.query_batch(q, b, ws)
        cols, mult, invp = oracle_mod.build_coordinate_sample(probs, 8, seed)
        np.testing.assert_array_equal(cols, index.sample["source_column"])
        scales, _ = oracle_mod.embedding_factors(mult, invp)
        er = (refs[:, cols.astype(np.int64)] * scales).astype(F32)
        eq = (q[:, cols.astype(np.int64)] * scales).astype(F32)
        k2 = oracle_mod.k2_distances(er, eq)
        e_idx, e_dist = oracle_mod.fp32_argmin(k2)
        np.testing.assert_array_equal(idx, e_idx)
        np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))
        # reference CPU flat path: binary64 over the same embedded values
        cpu_idx, _ = oracle_mod.l2_fp64_scan(er, eq)
        table = oracle_mod.fp64_table(False, er, eq)
        rows = np.arange(b)
        differ = idx != cpu_idx
        # a different answer only inside the FP32 band of the binary64 minimum
        k = np.ceil(len(cols) / 256) + 10
        gamma = k * 2.0**-24 / (1 - k * 2.0**-24)
        best = table[rows, cpu_idx.astype(np.int64)]
        got = table[rows, idx.astype(np.int64)]
        assert np.all(got[differ] <= best[differ] * (1 + 2 * gamma) + 1e-300)
        # the reference's quality figures from the GPU answers equal the CPU path's up to those ties
        agree_gpu = np.mean(idx == exact_cpu)
        agree_cpu = np.mean(cpu_idx == exact_cpu)
        assert abs(agree_gpu - agree_cpu) <= differ.mean()
        ex_gpu = _excess(refs, q, idx, exact_cpu)
        ex_cpu = _excess(refs, q, cpu_idx, exact_cpu)
        assert np.allclose(ex_gpu[~differ], ex_cpu[~differ])
