"""The benchmark configurations at their full sizes, on the GPU, against the
oracle (synthetic log-expression data from ultrahigh_ann_amd.synthetic,
generated on the device and copied to the host, where the oracle replays the
reference's arithmetic on the same values).

Config 2 -- r = 7028, d = 60660, batch 4096:
  * exact L2 / L1 direct: bit-identical to the K2 / K5 replay on 32 queries
    spread over the batch; every L2 answer is the binary64 argmin or inside
    the binary32 tie band of it (binary64 table on the GPU, pinned to the CPU
    order by test_gpu_scan.py);
  * the certified INT8 and the statistical BF16 Gram strategies: the direct
    path's row and bits on all 4096 queries, another row only through the
    binary64 near-tie re-rank (never worse in binary64);
  * flat L2 / L1 at T = 8, 32, 128: bit-identical to the K2 / weighted K5
    replay over the embedded coordinates (all queries at T = 8 and, for L1,
    T = 32; 1024, 512 or 256 queries spread over the batch otherwise); flat
    L2 INT8 equals flat L2 direct on all queries.
Config 3 -- hierarchical, T = 32, p = 16, 32, 64: L2 bit-identical to the K7
  projection + K2 replay on all 4096 queries; L1 to K7 + the K9 median on
  1024 (p = 16, 32) or 512 (p = 64) queries spread over the batch.
Config 4 -- r = 65536, batch 8192: FP32 direct, INT8 and BF16 give the same
  rows and bits on all 8192 queries (or a binary64 near tie); the direct
  distances equal the K2 replay on 4 queries; every strategy is scored on
  all 8192 queries against the GPU binary64 table; the uncertified FP8
  shortlist must reach agreement >= 0.999 with at most 0.1 % relative
  distance excess (its measured quality; it makes no guarantee).
Config 5 -- 64 indexes drawn as the benchmark draws them (r log-uniform in
  [33, 7028], T log-uniform in [4, 256], flat or hierarchical, L1 or L2,
  p in {16, 32, 64}) at d = 60660, served together: every request (1-256
  queries for a random index) equals its facade alone, bit for bit; one
  request per kind is replayed by the oracle."""

import numpy as np
import pytest

from conftest import rng
from test_gpu_indexes import _emb, _hier_expect

pytestmark = pytest.mark.gpu
F32 = np.float32
R, D, B = 7028, 60660, 4096


def _spread(n, count):
    return np.unique(np.linspace(0, n - 1, count).astype(np.int64))


def _gamma(k):
    return k * 2.0**-24 / (1 - k * 2.0**-24)


@pytest.fixture(scope="module")
def cfg2(uh):
    import torch

    from ultrahigh_ann_amd import synthetic

    refs, q = synthetic.log_expression(R, B, D, seed=2026, device="cuda")
    torch.cuda.synchronize()
    data = {"refs": refs, "q": q, "refs_np": refs.cpu().numpy(), "q_np": q.cpu().numpy(), "probs": {}}
    yield data
    del data
    torch.cuda.empty_cache()


def _probs(uh, cfg, metric):
    if metric not in cfg["probs"]:
        cfg["probs"][metric] = uh.importance_probabilities_device(metric, cfg["refs_np"])
    return cfg["probs"][metric]


def _device_query(uh, index, q, b, ws, strategy):
    import torch

    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_async(q, b, ws, oi, od, strategy=strategy, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    return oi.cpu().numpy().astype(np.uint64), od.cpu().numpy()


def _fp64_table(index, q, b):
    import torch

    table = torch.empty((b, index.rows), dtype=torch.float64, device="cuda")
    for s in range(0, b, 1024):
        e = min(b, s + 1024)
        index.distance_table(q[s:e], e - s, table[s:e])
    torch.cuda.synchronize()
    return table


def _gather_rows(table, idx):
    import torch

    gi = torch.as_tensor(idx.astype(np.int64), device="cuda")
    return table[torch.arange(gi.numel(), device="cuda"), gi].cpu().numpy()


def _same_or_fp64_tie(table, d_idx, d_dist, m_idx, m_dist):
    diff = m_idx != d_idx
    np.testing.assert_array_equal(m_dist[~diff].view(np.uint32), d_dist[~diff].view(np.uint32))
    if diff.any():
        got, want = _gather_rows(table, m_idx), _gather_rows(table, d_idx)
        assert np.all(got[diff] <= want[diff])
    return diff


@pytest.mark.parametrize("metric", ["l2", "l1"])
def test_config2_exact_direct(uh, oracle_mod, cfg2, metric):
    refs_np, q_np = cfg2["refs_np"], cfg2["q_np"]
    cls = uh.HipExactL2Index if metric == "l2" else uh.HipExactL1Index
    index = cls(refs_np)
    ws = index.make_query_workspace(B)
    idx, dist = _device_query(uh, index, cfg2["q"], B, ws, uh.STRATEGY_DIRECT)
    pick = _spread(B, 32)
    emul = (oracle_mod.k2_distances(refs_np, q_np[pick]) if metric == "l2"
            else oracle_mod.k5_distances(refs_np, q_np[pick]))
    e_idx, e_dist = oracle_mod.fp32_argmin(emul)
    np.testing.assert_array_equal(idx[pick], e_idx)
    np.testing.assert_array_equal(dist[pick].view(np.uint32), e_dist.view(np.uint32))
    table = _fp64_table(index, cfg2["q"], B)
    best = table.argmin(1).cpu().numpy().astype(np.uint64)
    off = idx != best
    if off.any():
        got, opt = _gather_rows(table, idx), _gather_rows(table, best)
        band = 2 * _gamma(np.ceil(D / 256) + 10)
        # L2 tables hold square roots: the squared-distance band halves
        assert np.all(got[off] <= opt[off] * (1 + (band / 2 if metric == "l2" else band)))


def test_config2_exact_mixed_full_batch(uh, oracle_mod, cfg2):
    refs_np = cfg2["refs_np"]
    index = uh.HipExactL2Index(refs_np)
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    index.enable_mixed(uh.STRATEGY_GRAM_BF16)
    ws = index.make_query_workspace(B)
    d_idx, d_dist = _device_query(uh, index, cfg2["q"], B, ws, uh.STRATEGY_DIRECT)
    table = _fp64_table(index, cfg2["q"], B)
    for strat in (uh.STRATEGY_GRAM_INT8, uh.STRATEGY_GRAM_BF16):
        m_idx, m_dist = _device_query(uh, index, cfg2["q"], B, ws, strat)
        _same_or_fp64_tie(table, d_idx, d_dist, m_idx, m_dist)
        if strat == uh.STRATEGY_GRAM_INT8:
            assert ws.shortlist_stats()["certified"] + ws.mixed_stats()["escalated"] == B


FLAT_SAMPLE = {("l2", 8): B, ("l2", 32): 1024, ("l2", 128): 256, ("l1", 8): B, ("l1", 32): B, ("l1", 128): 512}


@pytest.mark.parametrize("metric", ["l2", "l1"])
@pytest.mark.parametrize("T", [8, 32, 128])
def test_config2_flat(uh, oracle_mod, cfg2, metric, T):
    refs_np, q_np = cfg2["refs_np"], cfg2["q_np"]
    m = uh.METRIC_L2 if metric == "l2" else uh.METRIC_L1
    probs = _probs(uh, cfg2, m)
    cls = uh.HipFlatL2AnnIndex if metric == "l2" else uh.HipFlatL1AnnIndex
    index = cls(refs_np, probs, T, uh.Rng(2026))
    ws = index.make_query_workspace(B)
    idx, dist = _device_query(uh, index, cfg2["q"], B, ws, uh.STRATEGY_DIRECT)
    cols, mult, invp = (index.sample[f] for f in ("source_column", "multiplicity", "inverse_probability"))
    ref_cols, ref_mult, _ = oracle_mod.build_coordinate_sample(probs, T, 2026)
    np.testing.assert_array_equal(cols, ref_cols)
    np.testing.assert_array_equal(mult, ref_mult)
    scales, weights = oracle_mod.embedding_factors(mult, invp)
    pick = _spread(B, FLAT_SAMPLE[(metric, T)])
    if metric == "l2":
        emul = oracle_mod.k2_distances(_emb(refs_np, cols, scales), _emb(q_np[pick], cols, scales))
    else:
        emul = oracle_mod.k5_distances(_emb(refs_np, cols), _emb(q_np[pick], cols), weights, 0)
    e_idx, e_dist = oracle_mod.fp32_argmin(emul)
    np.testing.assert_array_equal(idx[pick], e_idx)
    np.testing.assert_array_equal(dist[pick].view(np.uint32), e_dist.view(np.uint32))
    if metric == "l2":
        index.enable_mixed(uh.STRATEGY_GRAM_INT8)
        m_idx, m_dist = _device_query(uh, index, cfg2["q"], B, ws, uh.STRATEGY_GRAM_INT8)
        diff = m_idx != idx
        np.testing.assert_array_equal(m_dist[~diff].view(np.uint32), dist[~diff].view(np.uint32))
        # a different row only from the binary64 near-tie re-rank over the embedded values
        if diff.any():
            rows = np.flatnonzero(diff)
            er = _emb(refs_np, cols, scales)
            eq = _emb(q_np[rows], cols, scales)
            t64 = oracle_mod.fp64_table(False, er, eq)
            r_ = np.arange(rows.size)
            assert np.all(t64[r_, m_idx[rows].astype(np.int64)] <= t64[r_, idx[rows].astype(np.int64)])


@pytest.mark.parametrize("metric", ["l2", "l1"])
@pytest.mark.parametrize("p", [16, 32, 64])
def test_config3_hierarchical(uh, oracle_mod, cfg2, metric, p):
    refs_np, q_np = cfg2["refs_np"], cfg2["q_np"]
    m = uh.METRIC_L2 if metric == "l2" else uh.METRIC_L1
    probs = _probs(uh, cfg2, m)
    cls = uh.HipHierarchicalL2AnnIndex if metric == "l2" else uh.HipHierarchicalL1AnnIndex
    index = cls(refs_np, probs, 32, p, uh.Rng(2026))
    ws = index.make_query_workspace(B)
    idx, dist = _device_query(uh, index, cfg2["q"], B, ws, uh.STRATEGY_DIRECT)
    pick = np.arange(B) if metric == "l2" else _spread(B, 1024 if p < 64 else 512)
    built, (e_idx, e_dist) = _hier_expect(oracle_mod, metric == "l1", refs_np, q_np[pick], probs, 32, p, 2026)
    np.testing.assert_array_equal(index.sample["source_column"], built["cols"])
    np.testing.assert_array_equal(idx[pick], e_idx)
    np.testing.assert_array_equal(dist[pick].view(np.uint32), e_dist.view(np.uint32))


def test_config4_mixed_r65536(uh, oracle_mod):
    import torch

    from ultrahigh_ann_amd import synthetic

    r, b = 65536, 8192
    refs, q = synthetic.log_expression(r, b, D, seed=2027, device="cuda")
    refs_np = refs.cpu().numpy()
    del refs
    torch.cuda.empty_cache()
    index = uh.HipExactL2Index(refs_np)
    for s in (uh.STRATEGY_GRAM_INT8, uh.STRATEGY_GRAM_BF16, uh.STRATEGY_GRAM_FP8):
        index.enable_mixed(s)
    ws = index.make_query_workspace(b)
    d_idx, d_dist = _device_query(uh, index, q, b, ws, uh.STRATEGY_DIRECT)
    pick = _spread(b, 4)
    e_idx, e_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs_np, q[pick].cpu().numpy()))
    np.testing.assert_array_equal(d_idx[pick], e_idx)
    np.testing.assert_array_equal(d_dist[pick].view(np.uint32), e_dist.view(np.uint32))
    table = _fp64_table(index, q, b)
    best = table.argmin(1).cpu().numpy().astype(np.uint64)
    for s in (uh.STRATEGY_GRAM_INT8, uh.STRATEGY_GRAM_BF16):
        m_idx, m_dist = _device_query(uh, index, q, b, ws, s)
        _same_or_fp64_tie(table, d_idx, d_dist, m_idx, m_dist)
    assert ws.mixed_stats()["escalated"] == 0
    f_idx, f_dist = _device_query(uh, index, q, b, ws, uh.STRATEGY_GRAM_FP8)
    got, opt = _gather_rows(table, f_idx), _gather_rows(table, best)
    assert np.mean(f_idx == best) >= 0.999
    assert np.max(got / opt - 1.0) <= 1e-3
    # every FP8 answer carries its row's binary32 reference-order distance
    same = f_idx == d_idx
    np.testing.assert_array_equal(f_dist[same].view(np.uint32), d_dist[same].view(np.uint32))
    del index, ws, table, q
    torch.cuda.empty_cache()


def test_config5_serving_64_indexes(uh, oracle_mod, cfg2):
    refs_np, q_np = cfg2["refs_np"], cfg2["q_np"]
    g = rng(2031)
    n = 64
    rows = np.exp(g.uniform(np.log(33), np.log(R), n)).astype(np.int64)
    reps = np.exp(g.uniform(np.log(4), np.log(256), n)).astype(np.int64)
    kinds = g.integers(0, 2, n)
    metrics = g.integers(0, 2, n)
    pdims = g.choice([16, 32, 64], n)
    starts = [int(g.integers(0, R - r + 1)) for r in rows]
    srv = uh.Server(hbm_budget_bytes=64 << 30, max_tick_queries=8192, max_tick_requests=512,
                    scratch_floats=64 << 20, staging_floats=1 << 20, use_graph=True)
    try:
        specs = []
        for i in range(n):
            sub = refs_np[starts[i]:starts[i] + rows[i]]
            m = int(metrics[i])
            probs = uh.importance_probabilities_device(m, sub)
            seed = 500 + i
            if kinds[i] == 0:
                sid = srv.register_flat(sub, probs, int(reps[i]), uh.Rng(seed), m)
                cls = uh.HipFlatL1AnnIndex if m else uh.HipFlatL2AnnIndex
                facade = cls(sub, probs, int(reps[i]), uh.Rng(seed))
            else:
                sid = srv.register_hierarchical(sub, probs, int(reps[i]), int(pdims[i]), uh.Rng(seed), m)
                cls = uh.HipHierarchicalL1AnnIndex if m else uh.HipHierarchicalL2AnnIndex
                facade = cls(sub, probs, int(reps[i]), int(pdims[i]), uh.Rng(seed))
            specs.append((sid, facade, sub, probs, seed))
        reqs = []
        for _ in range(160):
            i = int(g.integers(0, n))
            c = int(g.integers(1, 257))
            o = int(g.integers(0, B - c + 1))
            reqs.append((i, o, c))
        # device-resident queries, as the benchmark serves them
        handles = srv.run_trace([0.0] * len(reqs), [specs[i][0] for i, _, _ in reqs],
                                [cfg2["q"][o:o + c] for _, o, c in reqs], None)
        checked_kind = set()
        for (i, o, c), h in zip(reqs, handles):
            got_idx, got_dist = h.wait(timeout=120)
            sid, facade, sub, probs, seed = specs[i]
            ws = facade.make_query_workspace(c)
            f_idx, f_dist = facade.query_batch(q_np[o:o + c], c, ws)
            np.testing.assert_array_equal(got_idx, f_idx)
            np.testing.assert_array_equal(got_dist.view(np.uint32), f_dist.view(np.uint32))
            kind = (int(kinds[i]), int(metrics[i]))
            if kind in checked_kind:
                continue
            checked_kind.add(kind)
            qs = q_np[o:o + c]
            if kinds[i] == 0:
                cols, mult, invp = (facade.sample[f] for f in ("source_column", "multiplicity",
                                                               "inverse_probability"))
                scales, weights = oracle_mod.embedding_factors(mult, invp)
                emul = (oracle_mod.k5_distances(_emb(sub, cols), _emb(qs, cols), weights, 0) if metrics[i]
                        else oracle_mod.k2_distances(_emb(sub, cols, scales), _emb(qs, cols, scales)))
                e_idx, e_dist = oracle_mod.fp32_argmin(emul)
            else:
                _, (e_idx, e_dist) = _hier_expect(oracle_mod, bool(metrics[i]), sub, qs, probs, int(reps[i]),
                                                  int(pdims[i]), seed)
            np.testing.assert_array_equal(got_idx, e_idx)
            np.testing.assert_array_equal(got_dist.view(np.uint32), e_dist.view(np.uint32))
        assert len(checked_kind) == 4
    finally:
        srv.close()
