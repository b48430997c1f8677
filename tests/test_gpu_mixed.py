"""GPU checks of the mixed-precision exact L2 strategy (BF16 Gram on MFMA +
bounded candidate set + exact re-rank + binary64 near ties).

Bar: every result is the binary32 direct path's result (bit-identical
distance), except where the binary64 re-rank chose another row, in which case
that row's binary64 distance is not larger; and the result agrees with the
binary64 oracle at least as often as the binary32 path does. The reference's
cancellation fixture (tests/exact/l2/cuda_exact_l2_index_test.cpp:152-174),
where its unrefined GEMM strategy returns the wrong row, must return the right
one."""

import numpy as np
import pytest

from conftest import bf16_round, check_against_direct, rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def _check_against_direct(uh, oracle_mod, refs, q):
    return check_against_direct(uh, oracle_mod, refs, q, uh.STRATEGY_GRAM_BF16)


@pytest.mark.parametrize("r,d,b", [(300, 3000, 40), (129, 257, 130), (1000, 8192, 64), (7, 65, 3)])
def test_mixed_random(uh, oracle_mod, r, d, b):
    g = rng(r + d)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    refs[-1] = refs[0]  # duplicate rows: the lower index must win
    q[0] = refs[0]
    stats, _, _ = _check_against_direct(uh, oracle_mod, refs, q)
    assert stats["escalated"] <= b


def test_mixed_log_expression(uh, oracle_mod):
    import torch

    from ultrahigh_ann_amd import synthetic

    refs, q = synthetic.log_expression(700, 96, 20000, seed=11, device="cuda")
    refs, q = refs.cpu().numpy(), q.cpu().numpy()
    stats, agree_m, agree_f = _check_against_direct(uh, oracle_mod, refs, q)
    # the candidate sets stay small on this data
    assert stats["escalated"] == 0
    assert stats["candidates"] <= 16 * 96
    assert agree_m == 1.0
    del torch


def test_mixed_cancellation_fixture(uh):
    base = F32(1048576.0)
    index = uh.HipDenseL2ScanIndex(np.array([[base + F32(0.25)], [base + F32(0.125)]], F32))
    index.enable_mixed()
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(np.array([base], F32), 1, ws, uh.STRATEGY_GRAM_BF16)
    assert idx[0] == 1 and dist[0] == F32(0.015625)


def test_mixed_escalation_to_direct_scan(uh, oracle_mod):
    # 100 identical rows: every one is a candidate (> 64) -> the query is
    # answered by the exact direct scan, which returns the lowest row.
    g = rng(5)
    base = (g.random(512, dtype=np.float32) * 2).astype(F32)
    refs = np.tile(base, (100, 1))
    refs[37] += 0.5
    q = np.stack([base + F32(0.01), base]).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed()
    ws = index.make_query_workspace(2)
    idx, dist = index.query_batch(q, 2, ws, uh.STRATEGY_GRAM_BF16)
    assert ws.mixed_stats()["escalated"] == 2
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    np.testing.assert_array_equal(idx, exp_idx)
    np.testing.assert_array_equal(dist, exp_dist)


def test_mixed_escalation_mixed_batch(uh, oracle_mod):
    """Escalated and ordinary queries in one batch: 70 identical rows put more
    than 64 candidates in the first 20 queries (several 8-query tiles of the
    device-side direct scan), the other queries sit near distinct rows. Every
    answer equals the oracle's binary32 argmin; the device path (strided
    queries) agrees; a second batch without escalations is not affected by
    the first one's list."""
    import torch

    g = rng(9)
    d = 700
    base = (g.random(d, dtype=np.float32) * 2).astype(F32)
    refs = np.concatenate([np.tile(base, (70, 1)), (g.random((40, d), dtype=np.float32) * 2 + 3).astype(F32)])
    near_base = base + (g.random((20, d), dtype=np.float32) * 0.01).astype(F32)
    near_rows = refs[70 + g.integers(0, 40, 17)] + (g.random((17, d), dtype=np.float32) * 0.01).astype(F32)
    q = np.concatenate([near_base, near_rows]).astype(F32)
    b = q.shape[0]
    index = uh.HipExactL2Index(refs)
    index.enable_mixed()
    ws = index.make_query_workspace(b)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    idx, dist = index.query_batch(q, b, ws, uh.STRATEGY_GRAM_BF16)
    assert ws.mixed_stats()["escalated"] == 20
    np.testing.assert_array_equal(idx, exp_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), exp_dist.view(np.uint32))
    padded = torch.zeros((b, d + 3), dtype=torch.float32, device="cuda")
    padded[:, :d] = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_async(padded, b, ws, oi, od, strategy=uh.STRATEGY_GRAM_BF16,
                             stream=torch.cuda.current_stream(), ld_queries=d + 3)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))
    tail = q[20:]
    t_idx, t_dist = index.query_batch(tail, tail.shape[0], ws, uh.STRATEGY_GRAM_BF16)
    assert ws.mixed_stats()["escalated"] == 0
    np.testing.assert_array_equal(t_idx, exp_idx[20:])
    np.testing.assert_array_equal(t_dist.view(np.uint32), exp_dist[20:].view(np.uint32))


def test_mixed_requires_enable_and_l2(uh):
    refs = np.ones((3, 4), F32)
    index = uh.HipExactL2Index(refs)
    ws = index.make_query_workspace(1)
    with pytest.raises(uh.UhannInvalidArgument):
        index.query_batch(np.ones(4, F32), 1, ws, uh.STRATEGY_GRAM_BF16)
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipExactL1Index(refs).enable_mixed()


def test_mixed_stage_timing(uh):
    g = rng(9)
    refs = (g.random((256, 1024), dtype=np.float32)).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed()
    ws = index.make_query_workspace(16)
    ws.set_timing(True)
    index.query_batch(refs[:16].copy(), 16, ws, uh.STRATEGY_GRAM_BF16)
    t = ws.stage_times()
    assert len(t) == 4 and all(x >= 0 for x in t) and t[1] > 0


@pytest.mark.parametrize("r,d,b", [(300, 1000, 70), (517, 3000, 33), (40, 64, 300)])
def test_gram_estimates_every_tile(uh, r, d, b):
    """Every Gram block shape gives the same bits (each estimate accumulates
    the same K steps in the same order), and the estimates are the exact
    ||q||^2 + ||v||^2 - 2 <bf16(q), bf16(v)> up to binary32 accumulation."""
    import torch

    g = rng(r * 3 + d)
    refs = (g.random((r, d), dtype=np.float32) * 4 - 1).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4 - 1).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed()
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    outs = []
    for tile in range(8):
        out = torch.full((b, r + 3), float("nan"), dtype=torch.float32, device="cuda")
        index.gram_estimates(dq, b, ws, out, tile)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    for o in outs:
        assert np.all(np.isnan(o[:, r:]))  # row stride respected, nothing written past the references
        np.testing.assert_array_equal(o[:, :r].view(np.uint32), outs[0][:, :r].view(np.uint32))
    est = outs[0][:, :r].astype(np.float64)
    qb, vb = bf16_round(q), bf16_round(refs)
    nq = (q.astype(np.float64) ** 2).sum(1)
    nv = (refs.astype(np.float64) ** 2).sum(1)
    want = nq[:, None] + nv[None, :] - 2.0 * (qb @ vb.T)
    steps = (d + 63) // 64 * 4
    bound = 2.0 * (steps + 8) * 2.0**-24 * (np.abs(qb) @ np.abs(vb).T) + 4 * 2.0**-24 * (nq[:, None] + nv[None, :])
    assert np.all(np.abs(est - want) <= bound)
    # a plausible mistake -- truncating instead of rounding to bf16 -- is outside that bound
    trunc = lambda a: (np.ascontiguousarray(a, F32).view(np.uint32) & 0xFFFF0000).view(F32).astype(np.float64)
    wrong = nq[:, None] + nv[None, :] - 2.0 * (trunc(q) @ trunc(refs).T)
    assert np.any(np.abs(est - wrong) > bound)


@pytest.mark.parametrize("r,d,b", [(300, 4096, 70), (4200, 4096, 4100), (700, 2048, 600)])
def test_gram_split_tail(uh, r, d, b):
    """The queries' Gram configuration splits the tiles of a partial last
    round along K (here: every tile when there are fewer tiles than compute
    units, and the 33-tile tail of a 289-tile grid) and sums the parts in
    order. Tiles that stay whole keep the whole-tile bits; split tiles stay
    within the binary32 accumulation bound of the exact
    ||q||^2 + ||v||^2 - 2 <bf16(q), bf16(v)>, and differ from the whole-tile
    bits somewhere (the split ran)."""
    import torch

    g = rng(r + d + b)
    refs = (g.random((r, d), dtype=np.float32) * 4 - 1).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4 - 1).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed()
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    outs = {}
    for tile in (3, -1, -1):
        out = torch.full((b, r + 1), float("nan"), dtype=torch.float32, device="cuda")
        index.gram_estimates(dq, b, ws, out, tile)
        torch.cuda.synchronize()
        outs.setdefault(tile, []).append(out.cpu().numpy())
    whole = outs[3][0]
    split, split2 = outs[-1]
    assert np.all(np.isnan(split[:, r:]))
    np.testing.assert_array_equal(split.view(np.uint32), split2.view(np.uint32))  # deterministic
    split = split[:, :r].astype(np.float64)
    assert not np.array_equal(split, whole[:, :r].astype(np.float64))
    qb, vb = bf16_round(q), bf16_round(refs)
    nq = (q.astype(np.float64) ** 2).sum(1)
    nv = (refs.astype(np.float64) ** 2).sum(1)
    want = nq[:, None] + nv[None, :] - 2.0 * (qb @ vb.T)
    steps = (d + 63) // 64 * 4
    bound = 2.0 * (steps + 8) * 2.0**-24 * (np.abs(qb) @ np.abs(vb).T) + 4 * 2.0**-24 * (nq[:, None] + nv[None, :])
    assert np.all(np.abs(split - want) <= bound)
    if b * r > 256 * 256 * 256:
        # 289 tiles: the first 256 in the grouped order stay whole, so most
        # estimates keep the whole-tile bits
        same = split == whole[:, :r].astype(np.float64)
        assert same.mean() > 0.8


ODD_SHAPES = [(1, 70, 5), (2, 1, 3), (255, 63, 257), (256, 64, 256), (257, 65, 255), (511, 129, 2),
              (33, 1000, 1), (600, 127, 513)]


@pytest.mark.parametrize("r,d,b", ODD_SHAPES)
def test_mixed_and_fp8_odd_shapes(uh, oracle_mod, r, d, b):
    """Shapes around the 256-row Gram tiles, the 64-wide (BF16) and 128-wide
    (FP8) K steps, one reference, one coordinate and one query: the BF16
    strategy meets the same bar as the larger cases, and the FP8 strategy
    returns a row whose distance is its binary32 reference distance."""
    g = rng(r * 7 + d * 3 + b)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    if r > 2:
        refs[-1] = refs[0]
        q[0] = refs[0]
    _check_against_direct(uh, oracle_mod, refs, q)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws, uh.STRATEGY_GRAM_FP8)
    k2 = oracle_mod.k2_distances(refs, q)
    rows = np.arange(b)
    np.testing.assert_array_equal(dist.view(np.uint32), k2[rows, idx.astype(np.int64)].view(np.uint32))
    if r <= 16:  # the shortlist holds every row: the binary32 argmin (or a binary64 near tie)
        exp_idx, _ = oracle_mod.fp32_argmin(k2)
        table = oracle_mod.fp64_table(False, refs, q)
        same = idx == exp_idx
        assert np.all(same | (table[rows, idx.astype(np.int64)] <= table[rows, exp_idx.astype(np.int64)]))
