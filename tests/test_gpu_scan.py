"""GPU parity of the dense scan engine (exact L2 / L1, plain binary32 path)
through the C ABI, against the oracle's emulation of the reference's CUDA
kernels (bit-identical distances, hence bit-identical argmin) and the
reference's own CUDA test fixtures (tests/exact/l{1,2}/cuda_exact_l*_test.cpp)."""

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
F32 = np.float32

TILES = [(1, 8, 8, 8), (1, 12, 12, 8), (1, 16, 8, 4), (1, 8, 16, 2), (1, 4, 16, 1), (1, 2, 16, 1),
         (1, 1, 16, 1), (4, 1, 8, 1), (4, 2, 8, 1), (4, 4, 8, 1), (4, 8, 4, 2),
         (1, 1, 4, 1), (1, 2, 8, 1), (1, 4, 8, 1), (1, 16, 4, 4), (4, 1, 4, 1), (4, 2, 4, 1), (4, 4, 4, 1),
         (4, 1, 2, 1), (4, 1, 1, 1), (4, 2, 2, 1), (4, 4, 2, 1),
         (2, 16, 4, 4), (2, 8, 4, 4), (2, 8, 4, 2),
         (0, 1, 0, 0), (0, 2, 0, 0), (0, 4, 0, 0), (0, 4, 8, 0), (0, 8, 8, 0), (0, 0, 0, 0)]


def test_exact_l2_reference_fixture(uh):
    # tests/exact/l2/cuda_exact_l2_index_test.cpp:83-150
    refs = np.array([[0, 0], [5, 0], [2, 4]], F32)
    index = uh.HipExactL2Index(refs)
    ws = index.make_query_workspace(4)
    assert index.device() == 0 and uh.device_name(0)
    assert ws.maximum_batch_size() == 4
    assert index.payload_bytes() == 6 * 4 + 3 * 4
    q = np.array([1, 0, 3, 3, 2.5, 0, 5, 0], F32)
    idx, dist = index.query_batch(q, 4, ws)
    assert list(idx) == [0, 2, 0, 1]  # the tie (2.5, 0) resolves to row 0
    assert list(dist) == [1.0, 2.0, 6.25, 0.0]
    assert index.query(q[:2], ws) == 0


def test_cancellation_fixture(uh):
    # cuda_exact_l2_index_test.cpp:152-174: direct binary32 must pick row 1
    base = F32(1048576.0)
    index = uh.HipDenseL2ScanIndex(np.array([[base + F32(0.25)], [base + F32(0.125)]], F32))
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(np.array([base], F32), 1, ws)
    assert idx[0] == 1 and dist[0] == F32(0.015625)


def test_error_paths(uh):
    # cuda_exact_l2_index_test.cpp:176-246
    refs = np.array([[0, 0], [5, 0], [2, 4]], F32)
    index = uh.HipExactL2Index(refs)
    ws = index.make_query_workspace(4)
    with pytest.raises(uh.UhannInvalidArgument):
        index.query_batch(np.zeros(10, F32), 5, ws)  # capacity
    with pytest.raises(uh.UhannInvalidArgument):
        index.query_batch(np.zeros(3, F32), 2, ws)  # dimensions
    with pytest.raises(uh.UhannInvalidArgument):
        index.query(np.array([np.nan, 0], F32), ws)  # non-finite query
    other = uh.HipExactL2Index(refs)
    with pytest.raises(uh.UhannInvalidArgument):
        other.query(np.array([1, 0], F32), ws)  # foreign workspace
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipExactL2Index(np.array([[np.inf, 0]], F32))
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipExactL2Index(np.zeros((0, 3), F32))
    with pytest.raises(uh.UhannInvalidArgument):
        index.make_query_workspace(0)
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipDenseL1ScanIndex(refs, np.array([-1.0, 1.0], F32))
    # count == 0 is a no-op
    idx, _ = index.query_batch(np.zeros(0, F32), 0, ws)
    assert idx.size == 0


def test_exact_l1_reference_fixture(uh):
    # tests/exact/l1/cuda_exact_l1_index_test.cpp:98-135
    refs = np.array([[0, 0], [5, 0], [2, 4]], F32)
    index = uh.HipExactL1Index(refs)
    ws = index.make_query_workspace(4)
    assert index.payload_bytes() == 6 * 4
    idx, _ = index.query_batch(np.array([1, 0, 3, 3, 2.5, 0, 5, 0], F32), 4, ws)
    assert list(idx) == [0, 2, 0, 1]
    weighted = uh.HipDenseL1ScanIndex(refs, np.array([0.5, 4.0], F32))
    assert weighted.payload_bytes() == 6 * 4 + 2 * 4
    wws = weighted.make_query_workspace(1)
    assert weighted.query(np.array([3, 1], F32), wws) == 1


def _data(seed, r, d, b, dup=True):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    if dup and r > 3:
        refs[r - 1] = refs[1]  # exact duplicate: the lower index must win
        q[0] = refs[1]
    return refs, q


@pytest.mark.parametrize(
    "r,d,b",
    [(37, 1000, 19), (5, 3, 7), (64, 256, 33), (130, 4097, 5), (1, 70, 3), (300, 60660, 2)],
)
@pytest.mark.parametrize("metric", ["l2", "l1", "l1w"])
@pytest.mark.parametrize("pairs", ["auto", "off"])
def test_parity_bit_exact(uh, oracle_mod, metric, r, d, b, pairs, monkeypatch):
    if pairs == "off":
        monkeypatch.setenv("UHANN_PAIRS_MAX_DIM", "0")
    refs, q = _data(r * 31 + d, r, d, b)
    if metric == "l2":
        index = uh.HipDenseL2ScanIndex(refs)
        expect = oracle_mod.k2_distances(refs, q)
    elif metric == "l1":
        index = uh.HipDenseL1ScanIndex(refs)
        expect = oracle_mod.k5_distances(refs, q)
    else:
        w = (rng(d).random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
        index = uh.HipDenseL1ScanIndex(refs, w)
        expect = oracle_mod.k5_distances(refs, q, w, 0)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws)
    np.testing.assert_array_equal(idx, exp_idx)
    np.testing.assert_array_equal(dist.view(np.uint32), exp_dist.view(np.uint32))


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("d", [1000, 515, 4096, 33, 257])
def test_every_tile_bit_exact(uh, oracle_mod, tile, d):
    import torch

    r, b = 45, 13
    refs, q = _data(7 + d, r, d, b)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    expect = oracle_mod.k2_distances(refs, q)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    if (tile[0] == 4 and d % 4 != 0) or (tile[:2] == (0, 0) and d > 128):
        with pytest.raises(uh.UhannInvalidArgument):
            index.query_device_tile(dq, b, ws, oi, od, tile)
        return
    index.query_device_tile(dq, b, ws, oi, od, tile, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))


@pytest.mark.parametrize("tile", [(0, 1, 0, 0), (0, 2, 0, 0), (0, 4, 0, 0), (0, 4, 8, 0), (0, 8, 8, 0)])
@pytest.mark.parametrize("metric", ["l2", "l1", "l1w"])
@pytest.mark.parametrize("d", [222, 300, 32, 1000, 4096, 1, 6, 35, 256, 257, 1536, 1537])
def test_pair_tiles_many_reference_blocks(uh, oracle_mod, tile, metric, d):
    # Pair-per-thread scan over several 256-row reference blocks (the last one
    # partial), whole 32-column chunks and tails, empty groups (d = 32, 222),
    # several terms per chain (d = 1000, 4096: the widest rows), chain
    # classes without columns (d = 1, 6, 35), and the L1 weights, for every
    # queries-per-thread width of both bodies (whole 32-chain groups; chain
    # classes, tr = 8). d = 256 / 257: the last row with one unit per chain
    # class and the first with two; d = 1536 / 1537: the last row whose
    # query columns the chain-class body stages whole and the first it stages
    # a pass's two groups at a time.
    import torch

    r, b = 600, 9
    refs, q = _data(11 * d + tile[1], r, d, b)
    if metric == "l2":
        index = uh.HipDenseL2ScanIndex(refs)
        expect = oracle_mod.k2_distances(refs, q)
    elif metric == "l1":
        index = uh.HipDenseL1ScanIndex(refs)
        expect = oracle_mod.k5_distances(refs, q)
    else:
        w = (rng(d).random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
        index = uh.HipDenseL1ScanIndex(refs, w)
        expect = oracle_mod.k5_distances(refs, q, w, 0)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_tile(dq, b, ws, oi, od, tile, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))


@pytest.mark.parametrize("d", [1, 5, 16, 17, 31, 32, 33, 63, 64, 65, 100, 127, 128])
@pytest.mark.parametrize("r,b", [(1, 3), (130, 300), (300, 257)])
def test_short_rows_bit_exact(uh, oracle_mod, d, r, b):
    """The thread-per-query scan for rows of at most 128 coordinates (one term
    per chain, empty chains and groups left out): L2, L1 and weighted L1
    distances bit-identical to the oracle's replay of the reference's
    kernels, through the forced tile and through the automatic choice
    (batches of 256 or more queries)."""
    import torch

    refs, q = _data(d * 13 + r + b, r, d, b)
    w = (rng(d + 7).random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
    cases = ((uh.HipDenseL2ScanIndex(refs), oracle_mod.k2_distances(refs, q)),
             (uh.HipDenseL1ScanIndex(refs), oracle_mod.k5_distances(refs, q)),
             (uh.HipDenseL1ScanIndex(refs, w), oracle_mod.k5_distances(refs, q, w, 0)))
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    for index, expect in cases:
        exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
        ws = index.make_query_workspace(b)
        index.query_device_tile(dq, b, ws, oi, od, (0, 0, 0, 0), stream=torch.cuda.current_stream())
        torch.cuda.synchronize()
        np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
        np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))
        idx, dist = index.query_batch(q, b, ws)
        np.testing.assert_array_equal(idx, exp_idx)
        np.testing.assert_array_equal(dist.view(np.uint32), exp_dist.view(np.uint32))


STREAM_TILES = [t for t in TILES if t[0] in (4, 2)]


@pytest.mark.parametrize("tile", STREAM_TILES)
@pytest.mark.parametrize("d", [1000, 516, 4096, 3000])
def test_stream_tiles_bit_exact(uh, oracle_mod, tile, d):
    """A batch that fits one query tile streams its references (rows read one
    after the other per wave): same bits as the oracle, also for a partial
    query tile and with an L1 weighted index."""
    import torch

    r = 45
    for b in sorted({tile[1], max(1, tile[1] - 1)}):
        refs, q = _data(11 + d + b, r, d, b)
        dq = torch.from_numpy(q).cuda()
        oi = torch.zeros(b, dtype=torch.int64, device="cuda")
        od = torch.zeros(b, dtype=torch.float32, device="cuda")
        w = (rng(d + 1).random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
        for index, expect in ((uh.HipDenseL2ScanIndex(refs), oracle_mod.k2_distances(refs, q)),
                              (uh.HipDenseL1ScanIndex(refs, w), oracle_mod.k5_distances(refs, q, w, 0))):
            ws = index.make_query_workspace(b)
            exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
            index.query_device_tile(dq, b, ws, oi, od, tile, stream=torch.cuda.current_stream())
            torch.cuda.synchronize()
            np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
            np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))


@pytest.mark.parametrize("tile", [t for t in TILES if t[0] == 2])
@pytest.mark.parametrize("metric", ["l1", "l1w"])
@pytest.mark.parametrize("d", [1000, 257, 33, 1])
def test_shared_query_tiles_l1(uh, oracle_mod, tile, metric, d):
    """Shared-query tiles (one chain per thread, the block's queries staged in
    LDS) on L1 and weighted L1 indexes with several query tiles, the last one
    partial, and several reference blocks, the last one partial: the
    weighted form stages one step per LDS buffer, the others two."""
    import torch

    r, b = 45, 2 * tile[1] + 3
    refs, q = _data(17 + d + tile[1], r, d, b)
    if metric == "l1":
        index = uh.HipDenseL1ScanIndex(refs)
        expect = oracle_mod.k5_distances(refs, q)
    else:
        w = (rng(d + 5).random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
        index = uh.HipDenseL1ScanIndex(refs, w)
        expect = oracle_mod.k5_distances(refs, q, w, 0)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_tile(dq, b, ws, oi, od, tile, stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))


@pytest.mark.parametrize("tile", [t for t in TILES if t[0] == 2])
def test_shared_query_tiles_strided_queries(uh, oracle_mod, tile):
    """Shared-query tiles on device query rows with a stride that is not a
    multiple of four floats (ld = d + 13): same bits as the oracle."""
    import torch

    r, d = 45, 1000
    b = 2 * tile[1] + 3
    refs, q = _data(23 + tile[1] + tile[3], r, d, b)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    padded = torch.full((b, d + 13), float("nan"), dtype=torch.float32, device="cuda")
    padded[:, :d] = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_tile(padded, b, ws, oi, od, tile, stream=torch.cuda.current_stream(), ld_queries=d + 13)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy().view(np.uint32), exp_dist.view(np.uint32))


def test_device_batch_and_async(uh, oracle_mod):
    import torch

    refs, q = _data(3, 50, 777, 20)
    index = uh.HipDenseL1ScanIndex(refs)
    ws = index.make_query_workspace(20)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k5_distances(refs, q))
    dq = torch.from_numpy(q).cuda()
    idx, dist = index.query_device_batch(dq, 20, ws)
    np.testing.assert_array_equal(idx, exp_idx)
    np.testing.assert_array_equal(dist, exp_dist)
    # strided device queries (row stride > dim)
    padded = torch.zeros((20, 800), dtype=torch.float32, device="cuda")
    padded[:, :777] = dq
    oi = torch.zeros(20, dtype=torch.int64, device="cuda")
    od = torch.zeros(20, dtype=torch.float32, device="cuda")
    index.query_device_async(padded, 20, ws, oi, od, stream=torch.cuda.current_stream(), ld_queries=800)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(oi.cpu().numpy().astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy(), exp_dist)


def test_zero_dimension_and_overflow(uh, oracle_mod):
    # dim == 0: every distance is 0 and row 0 wins (cuda_dense_l1_scan_index.cu:487-490)
    index = uh.HipDenseL2ScanIndex(np.zeros((4, 0), F32))
    ws = index.make_query_workspace(2)
    idx, dist = index.query_batch(np.zeros(0, F32), 2, ws)
    assert list(idx) == [0, 0] and list(dist) == [0.0, 0.0]
    # huge finite values overflow to +inf; the lowest index among +inf wins
    refs = np.array([[3e38, 3e38], [-3e38, -3e38], [3e38, -3e38]], F32)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(np.array([-3e38, 3e38], F32), 1, ws)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, np.array([[-3e38, 3e38]], F32)))
    assert idx[0] == exp_idx[0] and np.isinf(dist[0]) == np.isinf(exp_dist[0])


@pytest.mark.parametrize("d", [300, 5000])
def test_autotune_keeps_bits(uh, oracle_mod, d):
    import torch

    r, b = 200, 40
    refs, q = _data(11 + d, r, d, b)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    table = index.autotune(dq, b, ws)
    assert len(table) >= 6 and all(row["us"] > 0 for row in table)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    for n in (1, 3, 17, 40):
        idx, dist = index.query_device_batch(dq[:n], n, ws)
        np.testing.assert_array_equal(idx, exp_idx[:n])
        np.testing.assert_array_equal(dist.view(np.uint32), exp_dist[:n].view(np.uint32))


def test_direct_kernel_timing_keeps_results(uh, oracle_mod):
    """Stage timing around the scan kernel (events on the workspace stream)
    reports a positive kernel time and leaves the results unchanged."""
    import torch

    refs, q = _data(21, 90, 5000, 3)
    index = uh.HipExactIndex(refs)
    ws = index.make_query_workspace(3)
    dq = torch.from_numpy(q).cuda()
    plain = index.query_device_batch(dq, 3, ws)
    ws.set_timing(True)
    timed_run = index.query_device_batch(dq, 3, ws)
    times = ws.stage_times()
    ws.set_timing(False)
    np.testing.assert_array_equal(plain[0], timed_run[0])
    np.testing.assert_array_equal(plain[1].view(np.uint32), timed_run[1].view(np.uint32))
    assert times[0] > 0.0 and times[1:] == [0.0, 0.0, 0.0]
    exp_idx, _ = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    np.testing.assert_array_equal(plain[0], exp_idx)


@pytest.mark.parametrize("metric_l1", [False, True])
@pytest.mark.parametrize("r,d,b", [(7, 33, 5), (130, 1000, 70), (65, 4097, 129)])
def test_distance_table_bit_exact(uh, oracle_mod, metric_l1, r, d, b):
    """The GPU diagnostic table is the reference CPU order's binary64
    distance, bit for bit (L2 with a correctly rounded square root), for
    ragged tiles; the diagnostics over it equal the same computed from the
    oracle's table."""
    import torch

    g = np.random.default_rng(r * d + b)
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(np.float32)
    q = (g.random((b, d), dtype=np.float32) * 3).astype(np.float32)
    q[0] = refs[r // 2]  # a zero distance
    index = uh.HipExactL1Index(refs) if metric_l1 else uh.HipExactL2Index(refs)
    dq = torch.from_numpy(q).cuda()
    out = torch.full((b, r + 2), float("nan"), dtype=torch.float64, device="cuda")
    index.distance_table(dq, b, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    assert np.all(np.isnan(got[:, r:]))
    t = oracle_mod.fp64_table(metric_l1, refs, q)
    want = t if metric_l1 else np.sqrt(t)
    np.testing.assert_array_equal(got[:, :r].view(np.uint64), want.view(np.uint64))
    table = uh.ExactDistanceTable(index, dq, b, chunk_rows=32)
    np.testing.assert_array_equal(table.distances.view(np.uint64), want.view(np.uint64))
    pred = np.asarray(index.query_batch(q, b, index.make_query_workspace(b))[0])
    assert table.evaluate(pred) == uh.ExactDistanceTable.from_table(want).evaluate(pred)


def test_distance_table_rejects_weighted(uh):
    import torch

    index = uh.HipDenseL1ScanIndex(np.ones((3, 4), np.float32), weights=np.ones(4, np.float32))
    out = torch.zeros((1, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(uh.UhannInvalidArgument):
        index.distance_table(torch.ones((1, 4), device="cuda"), 1, out)


@pytest.mark.parametrize("tq", [1, 2])
def test_pair_scan_batches_beyond_grid_limit(uh, oracle_mod, tq):
    """More query tiles than one grid dimension holds (65535): the pair scan
    runs the batch in slices. Queries around the slice boundaries and a
    spread sample agree with the oracle bit for bit."""
    import torch

    r, d, b = 300, 222, 140000
    g = rng(77 + tq)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_tile(dq, b, ws, oi, od, (0, tq, 0, 0), stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    edge = 65535 * tq
    pick = np.unique(np.concatenate([np.arange(edge - 3, edge + 3), np.arange(2 * edge - 3, min(b, 2 * edge + 3)),
                                     np.arange(0, b, 97), [b - 1]]))
    pick = pick[pick < b]
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q[pick]))
    np.testing.assert_array_equal(oi.cpu().numpy()[pick].astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy()[pick].view(np.uint32), exp_dist.view(np.uint32))


def test_distance_table_batches_beyond_grid_limit(uh, oracle_mod):
    """More query tiles than one grid dimension holds (65535 x 64 queries):
    the binary64 table runs the batch in slices; rows around the slice
    boundary and a spread sample equal the oracle's table bit for bit."""
    import torch

    r, d = 3, 2
    b = 65535 * 64 + 70
    g = np.random.default_rng(5)
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(np.float32)
    q = (g.random((b, d), dtype=np.float32) * 3).astype(np.float32)
    index = uh.HipExactL2Index(refs)
    dq = torch.from_numpy(q).cuda()
    out = torch.full((b, r), float("nan"), dtype=torch.float64, device="cuda")
    index.distance_table(dq, b, out)
    torch.cuda.synchronize()
    edge = 65535 * 64
    pick = np.unique(np.concatenate([np.arange(edge - 3, edge + 3), np.arange(0, b, 99991), np.arange(b - 3, b)]))
    got = out.cpu().numpy()[pick]
    want = np.sqrt(oracle_mod.fp64_table(False, refs, q[pick]))
    np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    del dq, out


@pytest.mark.parametrize("metric", ["l2", "l1w"])
def test_short_row_scan_batches_beyond_grid_limit(uh, oracle_mod, metric):
    """More query blocks than one grid dimension holds (65535 x 256 queries):
    the thread-per-query scan runs the batch in slices. Queries around the
    slice boundary and a spread sample agree with the oracle bit for bit."""
    import torch

    r, d = 9, 3
    b = 65535 * 256 + 300
    g = rng(91)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    if metric == "l2":
        index = uh.HipDenseL2ScanIndex(refs)
    else:
        w = (g.random(d, dtype=np.float32) * 3 + 0.1).astype(F32)
        index = uh.HipDenseL1ScanIndex(refs, w)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    oi = torch.zeros(b, dtype=torch.int64, device="cuda")
    od = torch.zeros(b, dtype=torch.float32, device="cuda")
    index.query_device_tile(dq, b, ws, oi, od, (0, 0, 0, 0), stream=torch.cuda.current_stream())
    torch.cuda.synchronize()
    edge = 65535 * 256
    pick = np.unique(np.concatenate([np.arange(edge - 3, edge + 3), np.arange(0, b, 104729), np.arange(b - 3, b)]))
    sub = q[pick]
    expect = oracle_mod.k2_distances(refs, sub) if metric == "l2" else oracle_mod.k5_distances(refs, sub, w, 0)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(expect)
    np.testing.assert_array_equal(oi.cpu().numpy()[pick].astype(np.uint64), exp_idx)
    np.testing.assert_array_equal(od.cpu().numpy()[pick].view(np.uint32), exp_dist.view(np.uint32))
    del dq, oi, od, ws


def test_autotune_large_batch_short_rows(uh, oracle_mod):
    """Autotuning up to 65536 queries on short rows tries the pair tiles at
    every bucket (the tq = 1 tile needs 65536 query tiles there) and keeps
    the bits."""
    import torch

    r, d, b = 200, 222, 65536
    refs, q = _data(5, r, d, b)
    index = uh.HipDenseL2ScanIndex(refs)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    table = index.autotune(dq, b, ws)
    assert len(table) == 17
    idx, dist = index.query_device_batch(dq, b, ws)
    pick = np.arange(0, b, 331)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q[pick]))
    np.testing.assert_array_equal(idx[pick], exp_idx)
    np.testing.assert_array_equal(dist[pick].view(np.uint32), exp_dist.view(np.uint32))
