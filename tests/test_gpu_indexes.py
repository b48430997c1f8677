"""GPU parity of the flat and hierarchical indexes through the C ABI.

Fixtures follow the reference's CUDA tests
(tests/coordinate_sampling/cuda_sampled_coordinate_l2_ann_index_test.cpp,
tests/hnsvw25/l{1,2}/hierarchical/cuda_hierarchical_l*_ann_index_test.cpp).
Random cases are checked bit for bit against the oracle's emulation of the
reference's CUDA arithmetic: flat = embedded references/queries then K2 (L2)
or weighted K5 (L1); hierarchical = K7 projection of the gathered coordinates
then K2 over p (L2) or the K9 median (L1)."""

import numpy as np
import pytest

from conftest import bf16_round, rng

pytestmark = pytest.mark.gpu
F32 = np.float32
NAN, INF = np.float32(np.nan), np.float32(np.inf)

REPS = np.array([[0, 99, 0, -5], [4, -99, 1, -5], [2, 0, 5, -5]], F32)
Q4 = np.array([[0.5, 10, 0.25, 20], [3.8, -10, 1.1, 20], [2, 10, 0.5, 20], [2, -10, 5, 20]], F32)


def _emb(x, cols, scales=None):
    g = np.ascontiguousarray(x[:, cols.astype(np.int64)], F32)
    return g if scales is None else (g * scales).astype(F32)


def test_flat_l2_reference_fixture(uh, oracle_mod):
    sample = uh.make_sample([0, 2], [2, 1], [0.5, 4.0])
    index = uh.HipSampledCoordinateL2AnnIndex(REPS, sample)
    ws = index.make_query_workspace(4)
    use = index.space_usage()
    assert use["index_payload_bytes"] == 3 * 2 * 4 + 3 * 4 + 2 * 8 + 2 * 4
    assert use["unique_query_coordinates"] == 2 and use["sampled_multiplicity"] == 3
    idx, dist = index.query_batch(Q4, 4, ws)
    scales, _ = oracle_mod.embedding_factors(sample["multiplicity"], sample["inverse_probability"])
    cols = sample["source_column"]
    exp64, _ = oracle_mod.l2_fp64_scan(_emb(REPS, cols, scales), _emb(Q4, cols, scales))
    np.testing.assert_array_equal(idx, exp64)
    assert idx[2] == 0  # tie resolves to the lowest row
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(_emb(REPS, cols, scales), _emb(Q4, cols, scales)))
    np.testing.assert_array_equal(dist.view(np.uint32), exp_dist.view(np.uint32))
    # unsampled non-finite coordinates are never read
    assert index.query(np.array([0.5, NAN, 0.25, INF], F32), ws) == index.query(np.array([0.5, 0, 0.25, 0], F32), ws)
    with pytest.raises(uh.UhannInvalidArgument):
        index.query(np.array([NAN, 0, 0, 0], F32), ws)
    # empty sample -> row 0
    empty = uh.HipSampledCoordinateL2AnnIndex(REPS, uh.make_sample([], [], []))
    ews = empty.make_query_workspace(2)
    assert list(empty.query_batch(Q4[:2], 2, ews)[0]) == [0, 0]
    # invalid sample
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipSampledCoordinateL2AnnIndex(REPS, uh.make_sample([0], [0], [1.0]))
    with pytest.raises(uh.UhannOutOfRange):
        uh.HipSampledCoordinateL2AnnIndex(REPS, uh.make_sample([9], [1], [1.0]))
    # foreign workspace and capacity
    other = uh.HipSampledCoordinateL2AnnIndex(REPS, sample)
    with pytest.raises(uh.UhannInvalidArgument):
        other.query(Q4[0], ws)
    with pytest.raises(uh.UhannInvalidArgument):
        index.query_batch(np.zeros((5, 4), F32), 5, ws)


def test_flat_seed_42_matches_cpu_flat(uh, oracle_mod, ref_mod):
    # cuda_sampled_coordinate_l2_ann_index_test.cpp:169-186
    probs = np.array([1.0, 0.0, 1.0, 0.0])
    index = uh.HipFlatL2AnnIndex(REPS, probs, 2, uh.Rng(42))
    ws = index.make_query_workspace(4)
    idx, _ = index.query_batch(Q4, 4, ws)
    theirs = np.zeros(4, np.uint64)
    assert ref_mod.ref().ref_flat(0, ref_mod._p(REPS, ref_mod._f32p), 3, 4, ref_mod._p(probs, ref_mod._f64p), 2, 42,
                                  ref_mod._p(Q4, ref_mod._f32p), 4, ref_mod._p(theirs, ref_mod._u64p)) == 0
    np.testing.assert_array_equal(idx, theirs)


def test_flat_l1_reference_fixture(uh, oracle_mod):
    sample = uh.make_sample([0, 2], [2, 1], [0.5, 4.0])
    index = uh.HipSampledCoordinateL1AnnIndex(REPS, sample)
    ws = index.make_query_workspace(4)
    idx, dist = index.query_batch(Q4, 4, ws)
    cols, mult, invp = sample["source_column"], sample["multiplicity"], sample["inverse_probability"]
    exp64, _ = oracle_mod.weighted_l1_fp64_scan(_emb(REPS, cols), Q4, cols, mult, invp)
    np.testing.assert_array_equal(idx, exp64)
    _, weights = oracle_mod.embedding_factors(mult, invp)
    e_idx, e_dist = oracle_mod.fp32_argmin(oracle_mod.k5_distances(_emb(REPS, cols), _emb(Q4, cols), weights, 0))
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))


def _random_index_case(seed, r, d, b):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 3).astype(F32)
    return refs, q


@pytest.mark.parametrize("metric", ["l2", "l1"])
@pytest.mark.parametrize("r,d,b,T", [(60, 900, 17, 16), (300, 4000, 40, 64), (7, 33, 5, 3)])
def test_flat_parity_bit_exact(uh, oracle_mod, metric, r, d, b, T):
    import torch

    refs, q = _random_index_case(r + d, r, d, b)
    l1 = metric == "l1"
    probs = oracle_mod.importance_probabilities(l1, refs)
    cls = uh.HipFlatL1AnnIndex if l1 else uh.HipFlatL2AnnIndex
    index = cls(refs, probs, T, uh.Rng(2026))
    ws = index.make_query_workspace(b)
    cols, mult, invp = oracle_mod.build_coordinate_sample(probs, T, 2026)
    np.testing.assert_array_equal(index.sample["source_column"], cols)
    scales, weights = oracle_mod.embedding_factors(mult, invp)
    if l1:
        emul = oracle_mod.k5_distances(_emb(refs, cols), _emb(q, cols), weights, 0)
    else:
        emul = oracle_mod.k2_distances(_emb(refs, cols, scales), _emb(q, cols, scales))
    e_idx, e_dist = oracle_mod.fp32_argmin(emul)
    idx, dist = index.query_batch(q, b, ws)
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))
    didx, ddist = index.query_device_batch(torch.from_numpy(q).cuda(), b, ws)
    np.testing.assert_array_equal(didx, e_idx)
    np.testing.assert_array_equal(ddist.view(np.uint32), e_dist.view(np.uint32))


@pytest.mark.parametrize("metric", ["l2", "l1"])
def test_uniform_reference_fixture(uh, oracle_mod, metric):
    # tests/coordinate_sampling/uniform_ann_index_test.cpp:26-77
    reps = np.array([[0, 0], [5, 5]], F32)
    cls = uh.HipUniformL1AnnIndex if metric == "l1" else uh.HipUniformL2AnnIndex
    queries = np.array([[1, 1], [4, 4]], F32)
    index = cls(reps, 2.0, 4, uh.Rng(42))
    ws = index.make_query_workspace(2)
    assert index.query(queries[0], ws) == 0
    assert index.query(queries[1], ws) == 1
    idx, _ = index.query_batch(queries, 2, ws)
    assert idx.tolist() == [0, 1]
    usage = index.space_usage()
    assert usage["unique_query_coordinates"] == 2 and usage["sampled_multiplicity"] == 2 * 4
    empty = cls(reps, 0.0, 4, uh.Rng(42))
    ews = empty.make_query_workspace(1)
    assert empty.query(queries[1], ews) == 0  # all-zero distances tie to row 0
    eu = empty.space_usage()
    assert eu["unique_query_coordinates"] == 0 and eu["sampled_multiplicity"] == 0
    if oracle_mod.ref_available():
        out = np.zeros(2, np.uint64)
        usage_ref = np.zeros(2, np.uint64)
        assert oracle_mod.ref().ref_uniform(int(metric == "l1"), oracle_mod._p(reps, oracle_mod._f32p), 2, 2, 2.0, 4, 42,
                                            oracle_mod._p(queries, oracle_mod._f32p), 2,
                                            oracle_mod._p(out, oracle_mod._u64p),
                                            oracle_mod._p(usage_ref, oracle_mod._u64p)) == 0
        assert out.tolist() == idx.tolist() and usage_ref.tolist() == [2, 8]


@pytest.mark.parametrize("metric", ["l2", "l1"])
@pytest.mark.parametrize("r,d,b,T", [(60, 900, 17, 16), (300, 4000, 40, 64)])
def test_uniform_parity_bit_exact(uh, oracle_mod, metric, r, d, b, T):
    """The uniform kind is the flat scan over the mass-matched uniform
    sample: same sample as the oracle's draw from mass / d, bit-identical
    distances to the oracle's K2 / weighted K5 emulation."""
    refs, q = _random_index_case(r * 3 + d, r, d, b)
    l1 = metric == "l1"
    mass = uh.sampling_mass(oracle_mod.importance_probabilities(l1, refs))
    cls = uh.HipUniformL1AnnIndex if l1 else uh.HipUniformL2AnnIndex
    index = cls(refs, mass, T, uh.Rng(1312))
    ws = index.make_query_workspace(b)
    cols, mult, invp = oracle_mod.build_coordinate_sample(np.full(d, mass / d), T, 1312)
    np.testing.assert_array_equal(index.sample["source_column"], cols)
    np.testing.assert_array_equal(index.sample["multiplicity"], mult)
    scales, weights = oracle_mod.embedding_factors(mult, invp)
    if l1:
        emul = oracle_mod.k5_distances(_emb(refs, cols), _emb(q, cols), weights, 0)
    else:
        emul = oracle_mod.k2_distances(_emb(refs, cols, scales), _emb(q, cols, scales))
    e_idx, e_dist = oracle_mod.fp32_argmin(emul)
    idx, dist = index.query_batch(q, b, ws)
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))


@pytest.mark.parametrize("strategy", ["int8", "bf16"])
def test_flat_l2_mixed_matches_direct(uh, oracle_mod, strategy):
    """The Gram strategies of the flat L2 index meet the exact index's bar on
    the embedded (gathered, scaled) coordinates: the direct path's row and
    binary32 bits, or another row only through the binary64 near-tie re-rank
    with a binary64 distance no larger; every returned distance is the K2
    distance of its row."""
    refs, q = _random_index_case(5, 400, 6000, 30)
    probs = oracle_mod.importance_probabilities(False, refs)
    index = uh.HipFlatL2AnnIndex(refs, probs, 128, uh.Rng(7))
    strat = uh.STRATEGY_GRAM_INT8 if strategy == "int8" else uh.STRATEGY_GRAM_BF16
    index.enable_mixed(strat)
    ws = index.make_query_workspace(30)
    d_idx, d_dist = index.query_batch(q, 30, ws)
    m_idx, m_dist = index.query_batch(q, 30, ws, strat)
    cols, mult, invp = (index.sample[f] for f in ("source_column", "multiplicity", "inverse_probability"))
    scales, _ = oracle_mod.embedding_factors(mult, invp)
    er, eq = _emb(refs, cols, scales), _emb(q, cols, scales)
    k2 = oracle_mod.k2_distances(er, eq)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(k2)
    np.testing.assert_array_equal(d_idx, exp_idx)
    rows = np.arange(30)
    np.testing.assert_array_equal(m_dist.view(np.uint32), k2[rows, m_idx.astype(np.int64)].view(np.uint32))
    diff = m_idx != d_idx
    table = oracle_mod.fp64_table(False, er, eq)
    assert np.all(table[rows[diff], m_idx[diff].astype(np.int64)] <= table[rows[diff], d_idx[diff].astype(np.int64)])
    if strategy == "int8":
        assert ws.shortlist_stats()["certified"] + ws.mixed_stats()["escalated"] == 30


@pytest.mark.parametrize("seed", [7, 42])
def test_flat_l1_matches_reference_cpu_outside_tie_band(uh, ref_mod, seed):
    """Flat L1 on the GPU (binary32 weighted K5 order) against the
    reference's own CPU FlatL1AnnIndex (binary64, oracle/_ref) built from the
    same probabilities and seed: same rows wherever the binary64 gap between
    the two nearest rows exceeds the binary32 tie band; a different row
    otherwise is within that band of the reference's."""
    r, d, b, T = 500, 3000, 64, 32
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 5).astype(F32)
    q = (refs[g.integers(0, r, b)] + g.random((b, d), dtype=np.float32) * 2).astype(F32)
    probs = ref_mod.importance_probabilities(True, refs)
    ref_idx = np.zeros(b, np.uint64)
    assert ref_mod.ref().ref_flat(1, ref_mod._p(refs, ref_mod._f32p), r, d, ref_mod._p(probs, ref_mod._f64p), T,
                                  seed, ref_mod._p(q, ref_mod._f32p), b, ref_mod._p(ref_idx, ref_mod._u64p)) == 0
    index = uh.HipFlatL1AnnIndex(refs, probs, T, uh.Rng(seed))
    ws = index.make_query_workspace(b)
    idx, _ = index.query_batch(q, b, ws)
    cols, mult, invp = (index.sample[f] for f in ("source_column", "multiplicity", "inverse_probability"))
    w64 = mult.astype(np.float64) * invp
    sub_r = refs[:, cols.astype(np.int64)].astype(np.float64)
    sub_q = q[:, cols.astype(np.int64)].astype(np.float64)
    exact = (w64[None, None, :] * np.abs(sub_q[:, None, :] - sub_r[None, :, :])).sum(2)
    np.testing.assert_array_equal(ref_idx, exact.argmin(1))
    kk = -(-cols.size // 256) + 11
    gamma = kk * 2.0**-24 / (1 - kk * 2.0**-24)
    rows = np.arange(b)
    diff = idx != ref_idx
    band = (2 * gamma) * exact[rows, ref_idx.astype(np.int64)]
    assert np.all(exact[rows[diff], idx[diff].astype(np.int64)] - exact[rows[diff], ref_idx[diff].astype(np.int64)]
                  <= band[diff])


def _hier_expect(oracle_mod, metric_l1, refs, q, probs, T, p, seed):
    built = oracle_mod.build_hierarchical(metric_l1, refs, probs, T, p, seed)
    x = _emb(q, built["cols"])
    y, bad = oracle_mod.k7_project(x, built["p32"])
    assert not bad
    if metric_l1:
        dist, _ = oracle_mod.k9_median_l1(built["proj_refs32"], y)
    else:
        dist = oracle_mod.k2_distances(built["proj_refs32"], y)
    return built, oracle_mod.fp32_argmin(dist)


def test_hierarchical_l2_reference_fixture(uh, oracle_mod):
    # cuda_hierarchical_l2_ann_index_test.cpp:50-187
    probs = np.array([1.0, 0.0, 1.0, 0.0])
    q = np.array([[0.2, 10, 0.1, 20], [3.9, -10, 1.1, 20], [2.1, 10, 4.8, 20], [1.0, -10, 2.0, 20]], F32)
    r = uh.Rng(1729)
    index = uh.HipHierarchicalL2AnnIndex(REPS, probs, 1, 7, r)
    r2 = uh.Rng(1729)
    s2 = uh.build_coordinate_sample(probs, 1, r2)
    uh.build_projection(s2, 7, uh.METRIC_L2, r2)
    assert r == r2  # construction consumed exactly the sample and projection draws
    use = index.space_usage()
    assert use["index_payload_bytes"] == 3 * 7 * 4 + 3 * 4 + 2 * 24 + 7 * 2 * 4
    assert use["unique_query_coordinates"] == 2 and use["sampled_multiplicity"] == 2
    ws = index.make_query_workspace(4)
    idx, dist = index.query_batch(q, 4, ws)
    built = oracle_mod.build_hierarchical(False, REPS, probs, 1, 7, 1729)
    cpu_idx, _ = oracle_mod.hierarchical_query_fp64(False, built, q)
    np.testing.assert_array_equal(idx, cpu_idx)
    _, (e_idx, e_dist) = _hier_expect(oracle_mod, False, REPS, q, probs, 1, 7, 1729)
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))
    assert index.query(np.array([0.2, NAN, 0.1, INF], F32), ws) == cpu_idx[0]
    with pytest.raises(uh.UhannInvalidArgument):
        index.query(np.array([NAN, 0, 0, 0], F32), ws)


@pytest.mark.parametrize("metric_l1,p", [(False, 31), (True, 17)])
def test_hierarchical_wide_identity(uh, metric_l1, p):
    # d = 513 (wider than one block), identity sample: each query is its own row
    rows, d = 4, 513
    col = np.arange(d)
    refs = np.array([[2.0 * r + 0.01 * ((c % 7) - 3) for c in col] for r in range(rows)], F32)
    sample = uh.make_sample(col, np.ones(d), np.ones(d))
    cls = uh.HipHierarchicalL1AnnIndex if metric_l1 else uh.HipHierarchicalL2AnnIndex
    index = cls(refs, None, 0, p, uh.Rng(1729), sample=sample)
    ws = index.make_query_workspace(rows)
    idx, _ = index.query_batch(refs, rows, ws)
    assert list(idx) == list(range(rows))


def test_hierarchical_l1_reference_fixture(uh, oracle_mod):
    # cuda_hierarchical_l1_ann_index_test.cpp:50-178
    refs = np.array([[0, 99], [5, -99]], F32)
    probs = np.array([1.0, 0.0])
    index = uh.HipHierarchicalL1AnnIndex(refs, probs, 16, 6, uh.Rng(1729))
    use = index.space_usage()
    assert use["index_payload_bytes"] == 24 + 6 * 4 + 2 * 6 * 4
    assert use["unique_query_coordinates"] == 1 and use["sampled_multiplicity"] == 16
    ws = index.make_query_workspace(4)
    q = np.array([[1, 100], [4, -100], [2.5, 5], [5, 0]], F32)
    idx, _ = index.query_batch(q, 4, ws)
    built = oracle_mod.build_hierarchical(True, refs, probs, 16, 6, 1729)
    cpu_idx, _ = oracle_mod.hierarchical_query_fp64(True, built, q)
    np.testing.assert_array_equal(idx, cpu_idx)
    assert idx[2] == 0
    assert index.query(np.array([1, NAN], F32), ws) == cpu_idx[0]
    with pytest.raises(uh.UhannInvalidArgument):
        index.query(np.array([INF, 0], F32), ws)


@pytest.mark.parametrize("metric_l1", [False, True])
@pytest.mark.parametrize("r,d,b,T,p", [(80, 1200, 21, 32, 16), (250, 3000, 33, 32, 64), (40, 700, 9, 8, 6),
                                       (60, 900, 11, 16, 32), (50, 800, 7, 16, 128), (30, 500, 5, 8, 8)])
def test_hierarchical_parity_bit_exact(uh, oracle_mod, metric_l1, r, d, b, T, p):
    refs, q = _random_index_case(r * 7 + p, r, d, b)
    probs = oracle_mod.importance_probabilities(metric_l1, refs)
    cls = uh.HipHierarchicalL1AnnIndex if metric_l1 else uh.HipHierarchicalL2AnnIndex
    index = cls(refs, probs, T, p, uh.Rng(11))
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws)
    _, (e_idx, e_dist) = _hier_expect(oracle_mod, metric_l1, refs, q, probs, T, p, 11)
    np.testing.assert_array_equal(idx, e_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), e_dist.view(np.uint32))


_bf16 = bf16_round


def test_bf16_rounding_helper():
    # the helper itself: ties go to even, 1 + 2^-8 is a tie between 1 and 1 + 2^-7
    x = np.array([1 + 2.0**-8, 1 + 3 * 2.0**-8, 1 + 2.0**-9, -(1 + 2.0**-8), 3.4e38], F32)
    np.testing.assert_array_equal(_bf16(x), [1.0, 1 + 2.0**-6, 1.0, -1.0, np.inf])


@pytest.mark.parametrize("r,d,b,T,p", [(300, 4000, 70, 32, 64), (90, 1500, 33, 32, 16), (40, 700, 9, 8, 33)])
def test_hierarchical_l2_bf16_projection(uh, oracle_mod, r, d, b, T, p):
    """PROJECTION_BF16 = exact projection of the bf16-rounded operands up to
    binary32 accumulation: distances within the derived bound, argmin equal
    outside that band."""
    refs, q = _random_index_case(r * 5 + p, r, d, b)
    probs = oracle_mod.importance_probabilities(False, refs)
    index = uh.HipHierarchicalL2AnnIndex(refs, probs, T, p, uh.Rng(23))
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws, uh.STRATEGY_PROJECTION_BF16)
    built = oracle_mod.build_hierarchical(False, refs, probs, T, p, 23)
    k = built["cols"].shape[0]
    xb = _bf16(_emb(q, built["cols"]))
    pb = _bf16(built["p32"])
    y = xb @ pb.T
    # |y_gpu - y| <= (ceil(k/16) + 8) u32 sum_j |x_j P_j| (products are exact in binary32)
    eps = ((k + 15) // 16 + 8) * 2.0**-24 * (np.abs(xb) @ np.abs(pb).T)
    pr = built["proj_refs32"].astype(np.float64)
    diff = y[:, None, :] - pr[None, :, :]
    d64 = (diff**2).sum(axis=2)
    tol = 2 * (np.abs(diff) * eps[:, None, :]).sum(axis=2) + (eps**2).sum(axis=1)[:, None] + (p + 10) * 2.0**-24 * d64
    rows = np.arange(b)
    chosen = idx.astype(np.int64)
    assert np.all(np.abs(dist.astype(np.float64) - d64[rows, chosen]) <= tol[rows, chosen])
    best = d64.argmin(axis=1)
    near = d64[rows, chosen] - d64[rows, best] <= tol[rows, chosen] + tol[rows, best]
    assert np.all((chosen == best) | near)
    assert np.mean(chosen == best) > 0.9
    # the binary32 path is unchanged by the bf16 copy
    idx32, _ = index.query_batch(q, b, ws)
    _, (e_idx, _) = _hier_expect(oracle_mod, False, refs, q, probs, T, p, 23)
    np.testing.assert_array_equal(idx32, e_idx)


def test_bf16_projection_wide_identity_and_l1_rejects(uh):
    rows, d = 4, 513
    col = np.arange(d)
    refs = np.array([[2.0 * r + 0.01 * ((c % 7) - 3) for c in col] for r in range(rows)], F32)
    sample = uh.make_sample(col, np.ones(d), np.ones(d))
    index = uh.HipHierarchicalL2AnnIndex(refs, None, 0, 31, uh.Rng(1729), sample=sample)
    ws = index.make_query_workspace(rows)
    idx, _ = index.query_batch(refs, rows, ws, uh.STRATEGY_PROJECTION_BF16)
    assert list(idx) == list(range(rows))
    l1 = uh.HipHierarchicalL1AnnIndex(refs, None, 0, 17, uh.Rng(1729), sample=sample)
    with pytest.raises(uh.UhannInvalidArgument):
        l1.query_batch(refs, rows, l1.make_query_workspace(rows), uh.STRATEGY_PROJECTION_BF16)


def test_hierarchical_single_row_and_errors(uh):
    refs = np.array([[1.0, 2.0, 3.0]], F32)
    index = uh.HipHierarchicalL2AnnIndex(refs, np.array([1.0, 1.0, 1.0]), 2, 3, uh.Rng(3))
    ws = index.make_query_workspace(2)
    assert list(index.query_batch(np.ones((2, 3), F32), 2, ws)[0]) == [0, 0]
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipHierarchicalL2AnnIndex(refs, np.array([1.0, 1.0, 1.0]), 2, 0, uh.Rng(3))
    with pytest.raises(uh.UhannError):
        uh.HipHierarchicalL2AnnIndex(np.ones((2, 3), F32), np.zeros(3), 2, 3, uh.Rng(3))


def test_async_status_reports_sampled_nan(uh):
    import torch

    sample = uh.make_sample([0, 2], [2, 1], [0.5, 4.0])
    index = uh.HipSampledCoordinateL2AnnIndex(REPS, sample)
    ws = index.make_query_workspace(2)
    q = torch.tensor([[0.5, 0, 0.25, 0], [NAN, 0, 0, 0]], dtype=torch.float32, device="cuda")
    oi = torch.zeros(2, dtype=torch.int64, device="cuda")
    od = torch.zeros(2, dtype=torch.float32, device="cuda")
    index.query_device_async(q, 2, ws, oi, od)
    with pytest.raises(uh.UhannInvalidArgument):
        ws.status()
    ws.status()  # cleared


@pytest.mark.parametrize("metric", [0, 1])
@pytest.mark.parametrize("r,d", [(70, 1031), (130, 257), (2, 5), (9, 64)])
def test_device_probabilities_bit_exact(uh, oracle_mod, metric, r, d):
    g = rng(r + d + metric)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    refs[r // 2] = refs[0]  # a zero-distance pair
    gpu = uh.importance_probabilities_device(metric, refs)
    cpu = oracle_mod.importance_probabilities(metric == 1, refs)
    np.testing.assert_array_equal(gpu.view(np.uint64), cpu.view(np.uint64))


def test_flat_autotune_keeps_bits(uh, oracle_mod):
    import torch

    refs, q = _random_index_case(21, 120, 2000, 33)
    probs = oracle_mod.importance_probabilities(False, refs)
    index = uh.HipFlatL2AnnIndex(refs, probs, 32, uh.Rng(3))
    ws = index.make_query_workspace(33)
    before, bd = index.query_batch(q, 33, ws)
    table = index.autotune(torch.from_numpy(q).cuda(), 33, ws)
    assert table
    after, ad = index.query_batch(q, 33, ws)
    np.testing.assert_array_equal(before, after)
    np.testing.assert_array_equal(bd.view(np.uint32), ad.view(np.uint32))
