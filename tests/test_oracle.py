"""Pins the CPU oracle (oracle/uhann_oracle.cpp) before it is trusted as the
checker: the reference's own known-answer fixtures, closed forms that fix the
reference's FP32 summation order, error bounds against binary64, and (where
it is built) the reference's own CPU code compiled from its sources."""

import math

import numpy as np
import pytest

from conftest import rng

F32 = np.float32


# ----------------------------------------------- reference known answers


def test_exact_l2_fixture(oracle_mod):
    # tests/exact/l2/exact_l2_index_test.cpp:37-77 (3x2 references)
    refs = np.array([[0, 0], [5, 0], [2, 4]], F32)
    q = np.array([[1, 0], [3, 3], [2.5, 0], [5, 0]], F32)
    idx, _ = oracle_mod.l2_fp64_scan(refs, q)
    assert list(idx) == [0, 2, 0, 1]
    i32, _ = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    assert list(i32) == [0, 2, 0, 1]


def test_cancellation_fixture(oracle_mod):
    # tests/exact/l2/cuda_exact_l2_index_test.cpp:152-174: direct FP32 must pick 1.
    base = F32(1048576.0)
    refs = np.array([[base + F32(0.25)], [base + F32(0.125)]], F32)
    q = np.array([[base]], F32)
    d = oracle_mod.k2_distances(refs, q)
    assert d[0, 0] == F32(0.0625) and d[0, 1] == F32(0.015625)
    assert oracle_mod.fp32_argmin(d)[0][0] == 1
    # The unrefined binary32 Gram identity ||r||^2+||q||^2-2rq (K1/K4) cancels
    # to a tie -> row 0, which is what the reference test asserts for gemm.
    r2 = (refs[:, 0] * refs[:, 0]).astype(F32)
    q2 = F32(q[0, 0] * q[0, 0])
    gram = (r2 + q2 - F32(2) * (refs[:, 0] * q[0, 0]).astype(F32)).astype(F32)
    assert int(np.argmin(gram)) == 0


def test_exact_l1_and_weighted_fixture(oracle_mod):
    # tests/exact/l1/exact_l1_index_test.cpp:39-83 and
    # cuda_exact_l1_index_test.cpp:124-135 (weights {0.5, 4}, query (3,1) -> 1)
    refs = np.array([[0, 0], [5, 0], [2, 4]], F32)
    q = np.array([[1, 0], [3, 3], [2.5, 0], [5, 0]], F32)
    assert list(oracle_mod.l1_fp64_scan(refs, q)[0]) == [0, 2, 0, 1]
    assert list(oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q))[0]) == [0, 2, 0, 1]
    w = np.array([0.5, 4.0], F32)
    for mode in (0, 1):
        d = oracle_mod.k5_distances(refs, np.array([[3, 1]], F32), w, mode)
        assert oracle_mod.fp32_argmin(d)[0][0] == 1


def test_probability_closed_forms(oracle_mod):
    # tests/hnsvw25/l2/importance_sampling_test.cpp:81-86, l1 :80-84
    refs = np.array([[0, 0], [3, 4]], F32)
    p2 = oracle_mod.importance_probabilities(False, refs)
    p1 = oracle_mod.importance_probabilities(True, refs)
    assert abs(p2[0] - 9 / 25) < 1e-12 and abs(p2[1] - 16 / 25) < 1e-12
    assert abs(p1[0] - 3 / 7) < 1e-12 and abs(p1[1] - 4 / 7) < 1e-12
    # duplicates or a single row -> zeros
    assert not oracle_mod.importance_probabilities(False, np.ones((2, 3), F32)).any()
    assert not oracle_mod.importance_probabilities(True, np.ones((1, 3), F32)).any()


def test_probability_one_column(oracle_mod):
    # tests/hnsvw25/l2/importance_sampling_test.cpp:291-305 (seed 42, T=16)
    cols, mult, invp = oracle_mod.build_coordinate_sample([1.0, 0.0], 16, 42)
    assert list(cols) == [0] and list(mult) == [16] and invp[0] == 1.0
    cols, _, _ = oracle_mod.build_coordinate_sample([0.0, 0.0, 0.0], 16, 42)
    assert cols.size == 0
    with pytest.raises(ValueError):
        oracle_mod.build_coordinate_sample([1.5], 4, 1)


# ---------------------------- closed forms that fix the FP32 summation order


def _k2_single(oracle_mod, diffs):
    refs = np.array([diffs], F32)
    return oracle_mod.k2_distances(refs, np.zeros_like(refs))[0, 0]


def test_k2_chain_order(oracle_mod):
    # Columns 0, 256, 512 all land in chain 0 and are added one by one:
    # (1 + 2^-24) + 2^-24 rounds to 1 twice. Binary64 would give 1 + 2^-23.
    d = np.zeros(513, F32)
    d[0], d[256], d[512] = 1.0, 2.0**-12, 2.0**-12
    assert _k2_single(oracle_mod, d) == F32(1.0)


def test_k2_warp_tree_order(oracle_mod):
    # Chains 1 and 3 meet at shfl offset 2 before chain 0 (offset 1):
    # 1 + (2^-24 + 2^-24) = 1 + 2^-23 exactly, while a left-to-right
    # binary32 sum of the chains stays at 1.
    d = np.zeros(4, F32)
    d[0], d[1], d[3] = 1.0, 2.0**-12, 2.0**-12
    assert _k2_single(oracle_mod, d) == F32(1 + 2.0**-23)
    seq = F32(0)
    for v in (d * d):
        seq = F32(seq + v)
    assert seq == F32(1.0)


def test_k2_block_tree_order(oracle_mod):
    # Warp sums W1 and W5 are paired ((W1+W5) in the second warp_sum), then
    # joined to W0: result 1 + 2^-23. Pairing W0 with W1 first would give 1.
    d = np.zeros(256, F32)
    d[0], d[32], d[160] = 1.0, 2.0**-12, 2.0**-12
    assert _k2_single(oracle_mod, d) == F32(1 + 2.0**-23)


def test_k2_small_integer_exact(oracle_mod):
    # Every partial sum of small integer squares is exact in binary32, so any
    # order gives the integer result.
    g = rng(1)
    refs = g.integers(-20, 20, size=(7, 300)).astype(F32)
    q = g.integers(-20, 20, size=(5, 300)).astype(F32)
    exact = ((refs[None, :, :].astype(np.int64) - q[:, None, :]) ** 2).sum(-1)
    np.testing.assert_array_equal(oracle_mod.k2_distances(refs, q), exact.astype(F32))
    exact1 = np.abs(refs[None, :, :].astype(np.int64) - q[:, None, :]).sum(-1)
    np.testing.assert_array_equal(oracle_mod.k5_distances(refs, q), exact1.astype(F32))


def test_k2_error_bound_vs_fp64(oracle_mod):
    # |k2 - s| <= gamma_k * s with k = ceil(D/256) + 8 + 2 (chain, tree,
    # subtract+square), u = 2^-24.
    g = rng(2)
    D = 5000
    refs = g.random((20, D), dtype=np.float32) * 10
    q = g.random((6, D), dtype=np.float32) * 10
    k2 = oracle_mod.k2_distances(refs, q).astype(np.float64)
    exact = oracle_mod.fp64_table(False, refs, q)
    k = math.ceil(D / 256) + 10
    u = 2.0**-24
    gamma = k * u / (1 - k * u)
    assert np.all(np.abs(k2 - exact) <= gamma * exact)
    # and the emulator is not accidentally binary64
    assert np.any(k2 != exact)


def test_k7_projection_matches_order_and_bound(oracle_mod):
    g = rng(3)
    x = g.standard_normal((3, 700)).astype(F32)
    P = g.standard_normal((5, 700)).astype(F32)
    y, bad = oracle_mod.k7_project(x, P)
    assert not bad
    exact = x.astype(np.float64) @ P.astype(np.float64).T
    mag = np.abs(x).astype(np.float64) @ np.abs(P).astype(np.float64).T
    k = math.ceil(700 / 256) + 8
    assert np.all(np.abs(y - exact) <= k * 2.0**-24 * mag * 1.01)


def test_k9_median_closed_form(oracle_mod):
    rp = np.array([[0, 0, 0, 0, 0], [1, 2, 3, 4, 5]], F32)
    qp = np.array([[1, -2, 3, -4, 10]], F32)
    med, bad = oracle_mod.k9_median_l1(rp, qp)
    assert not bad
    assert med[0, 0] == 3.0  # |q| sorted 1,2,3,4,10
    assert med[0, 1] == 4.0  # 0,4,0,8,5 -> 0,0,4,5,8
    med4, _ = oracle_mod.k9_median_l1(rp[:, :4], qp[:, :4])
    assert med4[0, 0] == F32(2.5)  # 1,2,3,4 -> 2 + (3-2)*0.5
    _, bad = oracle_mod.k9_median_l1(np.array([[3e38]], F32), np.array([[-3e38]], F32))
    assert bad


def test_l1_fp64_grouping_is_transform_reduce(oracle_mod):
    # libstdc++ transform_reduce sums |d| in groups of four:
    # init + ((t0+t1)+(t2+t3)). With t = (1, 0, e, e), e = 2^-53 that is
    # 1 + 2^-52; a left-to-right sum rounds each +e back to 1.
    e = 2.0**-53
    a = np.array([[1.0, 0, 0, 0]], F32)
    b = np.array([[0.0, 0.0, -e, -e]], np.float64).astype(F32)
    _, dist = oracle_mod.l1_fp64_scan(a, b)
    assert dist[0] == 1.0 + 2.0**-52
    seq = 0.0
    for t in (1.0, 0.0, e, e):
        seq += t
    assert seq == 1.0


# ---------------------------------------- against the compiled reference


def _random_case(seed, r, d, b, scale=3.0):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * scale).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * scale).astype(F32)
    return refs, q


@pytest.mark.parametrize("seed,r,d,b", [(0, 33, 700, 9), (1, 5, 3, 4), (2, 64, 2049, 5)])
def test_fp64_scans_match_reference(ref_mod, seed, r, d, b):
    refs, q = _random_case(seed, r, d, b)
    # make a near-duplicate so ties/near-ties are exercised
    refs[3] = refs[1]
    ours, _ = ref_mod.l2_fp64_scan(refs, q)
    theirs = np.zeros(b, np.uint64)
    assert ref_mod.ref().ref_exact_l2(ref_mod._p(refs, ref_mod._f32p), r, d,
                                      ref_mod._p(q, ref_mod._f32p), b, 0,
                                      ref_mod._p(theirs, ref_mod._u64p)) == 0
    np.testing.assert_array_equal(ours, theirs)
    ours1, _ = ref_mod.l1_fp64_scan(refs, q)
    assert ref_mod.ref().ref_exact_l1(ref_mod._p(refs, ref_mod._f32p), r, d,
                                      ref_mod._p(q, ref_mod._f32p), b,
                                      ref_mod._p(theirs, ref_mod._u64p)) == 0
    np.testing.assert_array_equal(ours1, theirs)


@pytest.mark.parametrize("metric_l1", [False, True])
def test_probabilities_match_reference(ref_mod, metric_l1):
    refs, _ = _random_case(5, 12, 300, 1)
    ours = ref_mod.importance_probabilities(metric_l1, refs)
    theirs = np.zeros(300, np.float64)
    assert ref_mod.ref().ref_importance_probabilities(int(metric_l1), ref_mod._p(refs, ref_mod._f32p),
                                                      12, 300, ref_mod._p(theirs, ref_mod._f64p)) == 0
    np.testing.assert_array_equal(ours, theirs)


@pytest.mark.parametrize("reps", [1, 8, 32, 128])
@pytest.mark.parametrize("seed", [7, 20, 42, 1312, 2026])
def test_sampling_matches_reference(ref_mod, reps, seed):
    g = rng(seed)
    probs = g.random(500) ** 3
    probs[::7] = 0.0
    probs[5] = 1.0
    cols, mult, invp = ref_mod.build_coordinate_sample(probs, reps, seed)
    d = probs.size
    c2 = np.zeros(d, np.uint64)
    m2 = np.zeros(d, np.uint64)
    i2 = np.zeros(d, np.float64)
    n = ref_mod.ref().ref_build_coordinate_sample(ref_mod._p(probs, ref_mod._f64p), d, reps, seed,
                                                  ref_mod._p(c2, ref_mod._u64p),
                                                  ref_mod._p(m2, ref_mod._u64p),
                                                  ref_mod._p(i2, ref_mod._f64p))
    np.testing.assert_array_equal(cols, c2[:n])
    np.testing.assert_array_equal(mult, m2[:n])
    np.testing.assert_array_equal(invp, i2[:n])


@pytest.mark.parametrize("metric_l1", [False, True])
@pytest.mark.parametrize("p", [7, 16, 33])
def test_hierarchical_build_matches_reference(ref_mod, metric_l1, p):
    refs, _ = _random_case(9, 10, 200, 1)
    probs = ref_mod.importance_probabilities(metric_l1, refs)
    built = ref_mod.build_hierarchical(metric_l1, refs, probs, 8, p, 1729)
    d = 200
    c2 = np.zeros(d, np.uint64)
    m2 = np.zeros(d, np.uint64)
    i2 = np.zeros(d, np.float64)
    p32 = np.zeros(p * d, np.float32)
    r32 = np.zeros(10 * p, np.float32)
    k = ref_mod.ref().ref_hierarchical_fp32_build(int(metric_l1), ref_mod._p(refs, ref_mod._f32p), 10,
                                                  d, ref_mod._p(probs, ref_mod._f64p), 8, p, 1729,
                                                  ref_mod._p(c2, ref_mod._u64p),
                                                  ref_mod._p(m2, ref_mod._u64p),
                                                  ref_mod._p(i2, ref_mod._f64p),
                                                  ref_mod._p(p32, ref_mod._f32p),
                                                  ref_mod._p(r32, ref_mod._f32p))
    assert k == built["cols"].size
    np.testing.assert_array_equal(built["cols"], c2[:k])
    np.testing.assert_array_equal(built["p32"].reshape(-1), p32[: p * k])
    np.testing.assert_array_equal(built["proj_refs32"].reshape(-1), r32)


@pytest.mark.parametrize("metric_l1", [False, True])
def test_flat_and_hierarchical_queries_match_reference(ref_mod, metric_l1):
    refs, q = _random_case(11, 20, 400, 12)
    probs = ref_mod.importance_probabilities(metric_l1, refs)
    T, seed, p = 16, 42, 9
    # flat: gather (+ binary32 scale for L2) then the binary64 scan
    cols, mult, invp = ref_mod.build_coordinate_sample(probs, T, seed)
    scales, _ = ref_mod.embedding_factors(mult, invp)
    if metric_l1:
        ours, _ = ref_mod.weighted_l1_fp64_scan(refs[:, cols.astype(np.int64)], q, cols, mult, invp)
    else:
        ci = cols.astype(np.int64)
        ours, _ = ref_mod.l2_fp64_scan(refs[:, ci] * scales, q[:, ci] * scales)
    theirs = np.zeros(12, np.uint64)
    assert ref_mod.ref().ref_flat(int(metric_l1), ref_mod._p(refs, ref_mod._f32p), 20, 400,
                                  ref_mod._p(probs, ref_mod._f64p), T, seed,
                                  ref_mod._p(q, ref_mod._f32p), 12,
                                  ref_mod._p(theirs, ref_mod._u64p)) == 0
    np.testing.assert_array_equal(ours, theirs)
    built = ref_mod.build_hierarchical(metric_l1, refs, probs, T, p, seed)
    ours_h, _ = ref_mod.hierarchical_query_fp64(metric_l1, built, q)
    assert ref_mod.ref().ref_hierarchical(int(metric_l1), ref_mod._p(refs, ref_mod._f32p), 20, 400,
                                          ref_mod._p(probs, ref_mod._f64p), T, p, seed,
                                          ref_mod._p(q, ref_mod._f32p), 12,
                                          ref_mod._p(theirs, ref_mod._u64p)) == 0
    np.testing.assert_array_equal(ours_h, theirs)


def _weighted_case(seed, r=50, k=700, b=12):
    g = rng(seed)
    refs = (g.random((r, k), dtype=np.float32) * 6).astype(F32)
    q = (g.random((b, k), dtype=np.float32) * 6).astype(F32)
    mult = g.integers(1, 5, k).astype(np.uint64)
    invp = np.exp(g.uniform(np.log(1.0), np.log(100.0), k))  # weights over more than two orders of magnitude
    w64 = mult.astype(np.float64) * invp                     # sampled_coordinate_l1_ann_index.cpp:63-83
    w32 = w64.astype(F32)                                    # the CUDA embedding's float weights
    return refs, q, mult, invp, w64, w32


def test_weighted_k5_within_gamma_of_binary64(oracle_mod):
    """The weighted K5 emulation (both contraction modes) stays within the
    binary32 error bound gamma_k (k = ceil(dim / 256) + 11: the chain, the
    8-level tree, the subtract, the product and the weight's rounding to
    binary32) of the reference's binary64 weighted L1; dropping the weights
    or applying them to the wrong columns leaves that bound."""
    refs, q, mult, invp, w64, w32 = _weighted_case(3)
    k = refs.shape[1]
    exact = (w64[None, None, :] * np.abs(q.astype(np.float64)[:, None, :] - refs.astype(np.float64)[None, :, :])).sum(2)
    # numpy's binary64 sum agrees with the oracle's (pinned to the reference's
    # SampledCoordinateL1AnnIndex) at the argmin
    cols = np.arange(k, dtype=np.uint64)
    e_idx, e_dist = oracle_mod.weighted_l1_fp64_scan(refs, q, cols, mult, invp)
    np.testing.assert_array_equal(e_idx, exact.argmin(1))
    np.testing.assert_allclose(e_dist, exact.min(1), rtol=1e-12)
    kk = -(-k // 256) + 11
    u = 2.0**-24
    gamma = kk * u / (1 - kk * u)
    for mode in (0, 1):
        got = oracle_mod.k5_distances(refs, q, w32, mode).astype(np.float64)
        assert np.all(np.abs(got - exact) <= gamma * exact)
    bad_drop = oracle_mod.k5_distances(refs, q).astype(np.float64)
    assert np.mean(np.abs(bad_drop - exact) > gamma * exact) > 0.99
    bad_cols = oracle_mod.k5_distances(refs, q, np.roll(w32, 1), 0).astype(np.float64)
    assert np.mean(np.abs(bad_cols - exact) > gamma * exact) > 0.99


def test_weighted_k5_argmin_outside_tie_band(oracle_mod):
    """Both contraction modes return the binary64 weighted argmin wherever
    the binary64 gap between the two nearest rows exceeds the tie band."""
    for seed in range(4):
        refs, q, mult, invp, w64, w32 = _weighted_case(40 + seed, r=300, k=900, b=40)
        exact = (w64[None, None, :] * np.abs(q.astype(np.float64)[:, None, :]
                                             - refs.astype(np.float64)[None, :, :])).sum(2)
        srt = np.sort(exact, axis=1)
        kk = -(-refs.shape[1] // 256) + 11
        gamma = kk * 2.0**-24 / (1 - kk * 2.0**-24)
        clear = (srt[:, 1] - srt[:, 0]) > 2 * gamma * srt[:, 1]
        for mode in (0, 1):
            idx, _ = oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q, w32, mode))
            np.testing.assert_array_equal(idx[clear], exact.argmin(1)[clear])


def test_weighted_l1_closed_form(oracle_mod):
    """Weights (10, 1, 0.01) on rows e0, 3 e1, 50 e2 and the zero query give
    10, 3, fl(0.01f * 50): the weighted argmin is row 2, while unweighted
    (1, 3, 50) or reversed weights (0.01, 3, 500) pick row 0."""
    refs = np.array([[1, 0, 0], [0, 3, 0], [0, 0, 50]], F32)
    q = np.zeros((1, 3), F32)
    w = np.array([10, 1, 0.01], F32)
    for mode in (0, 1):
        d = oracle_mod.k5_distances(refs, q, w, mode)[0]
        assert d[0] == F32(10) and d[1] == F32(3) and d[2] == F32(w[2] * F32(50))
        assert oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q, w, mode))[0][0] == 2
    assert oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q))[0][0] == 0
    assert oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q, w[::-1].copy(), 0))[0][0] == 0
    cols = np.arange(3, dtype=np.uint64)
    mult = np.array([10, 1, 1], np.uint64)
    invp = np.array([1.0, 1.0, 0.01])
    idx, dist = oracle_mod.weighted_l1_fp64_scan(refs, q, cols, mult, invp)
    assert idx[0] == 2 and abs(dist[0] - 0.5) < 1e-15
