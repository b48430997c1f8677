"""GPU tests of the serving loop (uhann_server_*, ultrahigh_ann_amd.Server).

Every served request must return the bits the one-index facade returns for
the same queries (the facades are pinned to the oracle by the other GPU
tests), whether it runs alone or shares its tick with requests for other
indexes, with graph replay or direct launches, across tick boundaries and
after autotuning. Admission by HBM budget and per-request errors are checked
from the host."""

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def _data(seed, r, d, b):
    g = rng(seed)
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 3).astype(F32)
    return refs, q


# name, kind, metric, rows, dim, T, p, bf16 projection, batch
CASES = [
    ("exact_l2", "exact", 0, 40, 1500, 0, 0, False, 9),
    ("exact_l1", "exact", 1, 33, 1500, 0, 0, False, 4),
    ("exact_l2_wide", "exact", 0, 20, 5000, 0, 0, False, 3),
    ("exact_l1_wide", "exact", 1, 21, 5000, 0, 0, False, 5),
    ("flat_l2", "flat", 0, 120, 1500, 16, 0, False, 23),
    ("flat_l1", "flat", 1, 90, 1500, 16, 0, False, 17),
    ("flat_l2_wide", "flat", 0, 50, 5000, 3000, 0, False, 11),
    ("flat_l1_wide", "flat", 1, 45, 5000, 2600, 0, False, 6),
    ("hier_l2", "hier", 0, 150, 1500, 16, 16, False, 40),
    ("hier_l2_bf16", "hier", 0, 150, 1500, 16, 33, True, 35),
    ("hier_l1", "hier", 1, 80, 1500, 16, 9, False, 21),
    ("hier_l1_p64", "hier", 1, 300, 1500, 24, 64, False, 7),
    ("hier_one_row", "hier", 0, 1, 1500, 4, 5, False, 3),
    # about 1750 sampled coordinates: the weighted-L1 pair stage for rows
    # longer than the short pair stages take
    ("flat_l1_mid", "flat", 1, 40, 4000, 800, 0, False, 13),
    # unweighted L1 rows of at most 1024 coordinates: the short pair stage
    ("exact_l1_short", "exact", 1, 300, 700, 0, 0, False, 6),
]


def _build(uh, srv, case, seed):
    name, kind, metric, r, d, T, p, bf16, b = case
    refs, q = _data(seed, r, d, b)
    strategy = uh.STRATEGY_PROJECTION_BF16 if bf16 else uh.STRATEGY_DIRECT
    if kind == "exact":
        facade = uh.HipExactIndex(refs, metric)
        sid = srv.register_exact(refs, metric)
    else:
        probs = uh.importance_probabilities(metric, refs) if r > 1 else np.full(d, 0.5)
        if kind == "flat":
            cls = uh.HipFlatL1AnnIndex if metric else uh.HipFlatL2AnnIndex
            facade = cls(refs, probs, T, uh.Rng(seed))
            sid = srv.register_flat(refs, probs, T, uh.Rng(seed), metric)
        else:
            cls = uh.HipHierarchicalL1AnnIndex if metric else uh.HipHierarchicalL2AnnIndex
            facade = cls(refs, probs, T, p, uh.Rng(seed))
            sid = srv.register_hierarchical(refs, probs, T, p, uh.Rng(seed), metric, strategy)
    ws = facade.make_query_workspace(b)
    idx, dist = facade.query_batch(q, b, ws, strategy)
    return sid, q, idx, dist


def _same(got, want):
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1].view(np.uint32), want[1].view(np.uint32))


@pytest.fixture(scope="module")
def served(uh):
    srv = uh.Server(hbm_budget_bytes=1 << 30)
    built = [(c[0],) + _build(uh, srv, c, 100 + i) for i, c in enumerate(CASES)]
    yield srv, built
    srv.close()


def test_mix_matches_facades(uh, served):
    import torch

    srv, built = served
    # half the requests from device memory, half from host memory; device
    # queries must be complete before they are submitted
    srcs = [torch.from_numpy(q).cuda() if i % 2 else q for i, (_, _, q, _, _) in enumerate(built)]
    torch.cuda.synchronize()
    reqs = [(name, srv.submit(sid, src), idx, dist) for (name, sid, _, idx, dist), src in zip(built, srcs)]
    for name, req, idx, dist in reqs:
        got = req.wait(timeout=60)
        _same(got, (idx, dist))
        assert got[0].dtype == np.uint64, name


def test_alone_equals_mix_and_trace(uh, served):
    srv, built = served
    alone = [srv.submit(sid, q).wait(timeout=60) for _, sid, q, _, _ in built]
    # the whole set again, arriving together through the native trace replay
    reqs = srv.run_trace([0.0] * len(built), [b[1] for b in built], [b[2] for b in built], None)
    for (name, *_), a, r in zip(built, alone, reqs):
        _same(r.wait(timeout=60), a)
    st = srv.stats()
    assert st["requests"] >= 2 * len(built) and st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["resident_indexes"] == len(built) and st["pending_indexes"] == 0


@pytest.mark.parametrize("use_graph,tq,max_q", [(False, 1, 8192), (True, 4, 40), (True, 2, 64)])
def test_graph_direct_and_tick_boundaries(uh, use_graph, tq, max_q):
    # max_q = 40 forces the requests over many ticks (each tick holds <= 40 rows)
    cases = [c for c in CASES if c[8] <= max_q]
    with uh.Server(hbm_budget_bytes=1 << 30, use_graph=use_graph, pairs_tq=tq, max_tick_queries=max_q) as srv:
        built = [_build(uh, srv, c, 300 + i) for i, c in enumerate(cases)]
        reqs = [(srv.submit(sid, q), idx, dist) for sid, q, idx, dist in built for _ in range(3)]
        for req, idx, dist in reqs:
            _same(req.wait(timeout=60), (idx, dist))
        assert srv.stats()["ticks"] >= 2


def test_admission_by_hbm_budget(uh):
    a_refs, q = _data(1, 200, 1500, 8)
    with uh.Server(hbm_budget_bytes=1 << 30) as probe:
        sid = probe.register_exact(a_refs)
        _, bytes_one = probe.state(sid)
    # room for two exact indexes of this shape, not three
    with uh.Server(hbm_budget_bytes=int(bytes_one * 2.5)) as srv:
        ids = [srv.register_exact(_data(s, 200, 1500, 1)[0]) for s in (1, 2, 3)]
        assert [srv.state(i)[0] for i in ids] == [uh.INDEX_RESIDENT, uh.INDEX_RESIDENT, uh.INDEX_PENDING]
        c_refs = _data(3, 200, 1500, 1)[0]
        fac = uh.HipExactIndex(c_refs)
        want = fac.query_batch(q, 8, fac.make_query_workspace(8))
        waiting = srv.submit(ids[2], q)
        with pytest.raises(TimeoutError):
            waiting.wait(timeout=0.3)
        assert srv.stats()["pending_indexes"] == 1
        srv.remove(ids[0])
        assert srv.state(ids[2])[0] == uh.INDEX_RESIDENT
        _same(waiting.wait(timeout=60), want)
        with pytest.raises(uh.UhannInvalidArgument):
            srv.submit(ids[0], q)  # removed
        with pytest.raises(uh.UhannInvalidArgument):
            srv.register_exact(np.ones((2000, 1500), F32))  # larger than the whole budget


def test_request_errors_stay_per_request(uh):
    refs, q = _data(5, 60, 1200, 4)
    probs = uh.importance_probabilities(0, refs)
    with uh.Server(hbm_budget_bytes=1 << 30, max_tick_queries=64) as srv:
        flat = srv.register_flat(refs, probs, 8, uh.Rng(9))
        exact = srv.register_exact(refs)
        sample = uh.HipFlatL2AnnIndex(refs, probs, 8, uh.Rng(9)).sample
        sampled = int(sample["source_column"][0])
        unsampled = int(np.setdiff1d(np.arange(1200), sample["source_column"].astype(np.int64))[0])
        bad = q.copy()
        bad[1, sampled] = np.nan
        ok = q.copy()
        ok[1, unsampled] = np.inf  # never read by the flat index
        r_bad = srv.submit(flat, bad)
        r_ok = srv.submit(flat, ok)
        r_exact_bad = srv.submit(exact, ok)  # the exact index reads every value
        r_exact = srv.submit(exact, q)
        with pytest.raises(uh.UhannInvalidArgument):
            r_bad.wait(timeout=60)
        with pytest.raises(uh.UhannInvalidArgument):
            r_exact_bad.wait(timeout=60)
        r_ok.wait(timeout=60)
        fac = uh.HipExactIndex(refs)
        _same(r_exact.wait(timeout=60), fac.query_batch(q, 4, fac.make_query_workspace(4)))
        with pytest.raises(uh.UhannInvalidArgument):
            srv.submit(12345, q)
        with pytest.raises(uh.UhannInvalidArgument):
            srv.submit(exact, np.zeros((65, 1200), F32))  # more rows than a tick holds
        assert len(srv.submit(exact, np.zeros((0, 1200), F32)).wait(timeout=60)[0]) == 0


def test_autotune_keeps_results(uh, served):
    import torch

    srv, built = served
    width = max(b[2].shape[1] for b in built)
    dq = torch.zeros((32, width), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    tq = srv.autotune(dq, 16)
    assert tq in (1, 2, 4)
    for name, sid, q, idx, dist in built:
        _same(srv.submit(sid, q).wait(timeout=60), (idx, dist))
