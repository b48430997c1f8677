"""GPU checks that one batch of the index query path (direct binary32 scan,
BF16 Gram + certified re-rank with device-side escalation, FP8 shortlist) can
be captured into a HIP graph through the asynchronous entry point and
replayed: each replay, with eager work and synchronisations between the
replays, returns the eager call's results bit for bit (indices and binary32
distances), including a batch that escalates queries to the direct scan."""

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def _batches(seed, r, d, b, escalate):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    qa = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    qb = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    if escalate:
        refs[:80] = refs[0]          # 80 identical rows: more than 64 candidates
        qb[: b // 2] = refs[0] + F32(0.01)
    return refs, qa, qb


# Escalation to the direct scan exists in the certified strategies (BF16, INT8).
@pytest.mark.parametrize("strategy,escalate", [("direct", False), ("gram_bf16", False), ("gram_bf16", True),
                                               ("gram_fp8", False), ("gram_int8", False), ("gram_int8", True)])
def test_query_graph_replay(uh, strategy, escalate):
    import torch

    r, d, b = 300, 1000, 40
    refs, qa, qb = _batches(31 + int(escalate), r, d, b, escalate)
    strat = {"direct": uh.STRATEGY_DIRECT, "gram_bf16": uh.STRATEGY_GRAM_BF16, "gram_fp8": uh.STRATEGY_GRAM_FP8,
             "gram_int8": uh.STRATEGY_GRAM_INT8}[strategy]
    index = uh.HipExactL2Index(refs)
    if strat != uh.STRATEGY_DIRECT:
        index.enable_mixed(strat)
    ws = index.make_query_workspace(b)
    eager = {0: index.query_batch(qa, b, ws, strat), 1: index.query_batch(qb, b, ws, strat)}
    if escalate:
        assert ws.mixed_stats()["escalated"] >= b // 2
    dq = torch.from_numpy(qa).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up: workspace buffers exist before the capture
        index.query_device_async(dq, b, ws, oi, od, strategy=strat, stream=side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        index.query_device_async(dq, b, ws, oi, od, strategy=strat, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    hq = {0: torch.from_numpy(qa).cuda(), 1: torch.from_numpy(qb).cuda()}
    for rep, which in enumerate((0, 1, 0, 1, 1)):
        dq.copy_(hq[which])
        oi.fill_(-1)
        od.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        e_idx, e_dist = eager[which]
        np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), e_idx, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), e_dist.view(np.uint32), err_msg=f"replay {rep}")
        # eager work on the same workspace between two replays
        o_idx, _ = index.query_batch(qb if which == 0 else qa, b, ws, strat)
        np.testing.assert_array_equal(o_idx, eager[1 - which][0])


def test_shared_query_tile_graph_replay(uh, oracle_mod):
    """The shared-query scan tile (16 queries, four groups of four rows per
    block) captured into a HIP graph: every replay, with eager calls on the
    same workspace between them, returns the oracle's rows and distances."""
    import torch

    r, d, b = 300, 1000, 16
    tile = (2, 16, 4, 4)
    refs, qa, qb = _batches(57, r, d, b, False)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    expect = {}
    for which, q in ((0, qa), (1, qb)):
        idx, dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
        expect[which] = (idx, dist)
    dq = torch.from_numpy(qa).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        index.query_device_tile(dq, b, ws, oi, od, tile, stream=side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        index.query_device_tile(dq, b, ws, oi, od, tile, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    hq = {0: torch.from_numpy(qa).cuda(), 1: torch.from_numpy(qb).cuda()}
    for rep, which in enumerate((0, 1, 1, 0, 1)):
        dq.copy_(hq[which])
        oi.fill_(-1)
        od.fill_(-1.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        e_idx, e_dist = expect[which]
        np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), e_idx, err_msg=f"replay {rep}")
        np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), e_dist.view(np.uint32), err_msg=f"replay {rep}")
        o_idx, _ = index.query_batch(qb if which == 0 else qa, b, ws)
        np.testing.assert_array_equal(o_idx, expect[1 - which][0])
