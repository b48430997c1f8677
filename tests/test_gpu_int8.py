"""GPU checks of the certified GRAM_INT8 strategy (16-bit fixed-point rows in
two int8 planes, exact int32 MFMA sums, deterministic candidate bound, exact
re-rank, binary64 near ties, direct scan for escalated queries).

* The row conversion equals the numpy model (conftest.int8_model) byte for
  byte; the stored norms are never below the binary64 values and at most
  2^-19 above them.
* The estimates equal a host replay of H + X / 256 in exact integers and the
  kernel's binary64 combination, bit for bit, for every block shape and for
  the K-split tail; they lie within the deterministic bound of the binary64
  squared distance.
* Results meet the bar of conftest.check_against_direct (the binary32 direct
  path's answer; another row only through the binary64 near-tie re-rank,
  never worse in binary64) on random, log-expression-like, log2(count + 1)
  (counts 1-3 and Poisson at d = 60660, values that round the same way every
  time) and signed data, odd shapes, the reference's cancellation fixture
  (tests/exact/l2/cuda_exact_l2_index_test.cpp:152-174) and a pair of rows
  whose coherent bf16 rounding makes the BF16 strategy's statistical bound
  exclude the nearest row (the BF16 strategy returns the wrong row there; the
  INT8 strategy returns the direct path's row)."""

import numpy as np
import pytest

from conftest import check_against_direct, int8_model, rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def coherent_pair(d=60660):
    """One query of ones and two rows: row 0 all log2(3) (bf16 rounds every
    coordinate up by the same 9.7e-4), row 1 a mix of the two bf16 values
    around it (no rounding). Row 1 is nearer by 57.6; the BF16 estimate puts
    row 0 ahead by 60."""
    q = np.ones((1, d), F32)
    r0 = np.full(d, np.log2(F32(3)), F32)
    r1 = np.full(d, F32(1.578125), F32)
    r1[: int(0.77 * d)] = F32(1.5859375)
    return np.stack([r0, r1]), q


def _stats_view(stats):
    s = stats.cpu().numpy()
    rec = s.view(np.uint8).reshape(-1, 32)
    norm2 = rec[:, :8].copy().view(np.float64)[:, 0]
    f = rec[:, 8:32].copy().view(np.float32)
    return {"norm2": norm2, "scale": f[:, 0], "err": f[:, 1], "hat": f[:, 2], "lo": f[:, 3], "norm": f[:, 4]}


def _host_bytes(a, b, dim):
    width = max(128, (dim + 63) // 64 * 128)
    out = np.zeros((a.shape[0], width), np.uint8)
    for c in range((dim + 63) // 64):
        lo, hi = 64 * c, min(dim, 64 * c + 64)
        out[:, 128 * c: 128 * c + hi - lo] = (a[:, lo:hi] & 255).astype(np.uint8)
        out[:, 128 * c + 64: 128 * c + 64 + hi - lo] = (b[:, lo:hi] & 255).astype(np.uint8)
    return out


@pytest.mark.parametrize("n,dim", [(5, 1), (3, 63), (4, 64), (7, 65), (9, 1000), (3, 60660)])
def test_int8_rows_match_model(uh, n, dim):
    import torch

    g = rng(n * 31 + dim)
    x = (g.standard_normal((n, dim)) * 10 ** g.uniform(-5, 5, (n, 1))).astype(F32)
    x[0] = 0.0
    if n > 3:
        x[1] = np.log2(g.integers(1, 4, dim) + 1.0).astype(F32)
        x[2, : dim // 2 + 1] = -3.25
    dev = torch.from_numpy(x).cuda()
    got_bytes, got_stats = uh.int8_rows(dev, n, dim)
    torch.cuda.synchronize()
    a, b, s, codes = int8_model(x)
    np.testing.assert_array_equal(got_bytes.cpu().numpy(), _host_bytes(a, b, dim))
    st = _stats_view(got_stats)
    np.testing.assert_array_equal(st["scale"], s)
    xd = x.astype(np.float64)
    hat = s.astype(np.float64)[:, None] * codes / 256.0
    want = {
        "norm": np.sqrt((xd * xd).sum(1)),
        "err": np.sqrt(((xd - hat) ** 2).sum(1)),
        "hat": np.sqrt((hat * hat).sum(1)),
        "lo": s.astype(np.float64) * np.sqrt((b.astype(np.float64) ** 2).sum(1)) / 256.0,
    }
    np.testing.assert_allclose(st["norm2"], (xd * xd).sum(1), rtol=1e-12)
    for k, w in want.items():
        got = st[k].astype(np.float64)
        assert np.all(got >= w * (1 - 1e-13)), k  # rounded up (binary64 sums of both sides agree to 1e-13)
        assert np.all(got <= w * (1 + 2.0**-19) + 1e-300), k


def _replay_estimates(q, refs, qst, vst):
    aq, bq, sq, _ = int8_model(q)
    av, bv, sv, _ = int8_model(refs)
    h = (aq @ av.T).astype(np.float64)
    x = (aq @ bv.T + bq @ av.T).astype(np.float64)
    t = h + x * 0.00390625
    dot = (qst["scale"].astype(np.float64)[:, None] * vst["scale"].astype(np.float64)[None, :]) * t
    return ((qst["norm2"][:, None] + vst["norm2"][None, :]) - 2.0 * dot).astype(F32)


@pytest.mark.parametrize("r,d,b", [(300, 1000, 70), (517, 3000, 33), (40, 64, 300), (4200, 2048, 1100)])
def test_int8_estimates_bit_exact_and_bounded(uh, r, d, b):
    import torch

    g = rng(r * 3 + d)
    refs = (g.random((r, d), dtype=np.float32) * 4 - 1).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4 - 1).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    ws = index.make_query_workspace(b)
    dq = torch.from_numpy(q).cuda()
    outs = []
    for tile in (-1, 0, 1, 2, 3, 4, 5, 6, 7):
        out = torch.full((b, r + 3), float("nan"), dtype=torch.float32, device="cuda")
        index.int8_estimates(dq, b, ws, out, tile)
        torch.cuda.synchronize()
        outs.append(out.cpu().numpy())
    for o in outs:
        assert np.all(np.isnan(o[:, r:]))
        np.testing.assert_array_equal(o[:, :r].view(np.uint32), outs[0][:, :r].view(np.uint32))
    _, qstats = uh.int8_rows(dq, b, d)
    _, vstats = uh.int8_rows(torch.from_numpy(refs).cuda(), r, d)
    torch.cuda.synchronize()
    qst, vst = _stats_view(qstats), _stats_view(vstats)
    want = _replay_estimates(q, refs, qst, vst)
    np.testing.assert_array_equal(outs[0][:, :r].view(np.uint32), want.view(np.uint32))
    est = outs[0][:, :r].astype(np.float64)
    diff = q.astype(np.float64)[:, None, :] - refs.astype(np.float64)[None, :, :] if b * r * d < 4e8 else None
    if diff is not None:
        exact = (diff * diff).sum(2)
        beta = (qst["norm"].astype(np.float64)[:, None] * vst["err"][None, :]
                + qst["err"].astype(np.float64)[:, None] * vst["hat"][None, :]
                + qst["lo"].astype(np.float64)[:, None] * vst["lo"][None, :])
        tau = 2 * beta + 2.0**-23 * np.abs(est) + 1e-10 * (qst["norm2"][:, None] + vst["norm2"][None, :])
        assert np.all(np.abs(est - exact) <= tau)


@pytest.mark.parametrize("r,d,b", [(300, 3000, 40), (129, 257, 130), (1000, 8192, 64), (7, 65, 3)])
def test_int8_random(uh, oracle_mod, r, d, b):
    g = rng(r + d)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    refs[-1] = refs[0]
    q[0] = refs[0]
    stats, agree, _ = check_against_direct(uh, oracle_mod, refs, q, uh.STRATEGY_GRAM_INT8)
    assert stats["escalated"] + stats["certified"] == b


def test_int8_log_expression(uh, oracle_mod):
    from ultrahigh_ann_amd import synthetic

    refs, q = synthetic.log_expression(700, 96, 20000, seed=11, device="cuda")
    refs, q = refs.cpu().numpy(), q.cpu().numpy()
    stats, agree_m, _ = check_against_direct(uh, oracle_mod, refs, q, uh.STRATEGY_GRAM_INT8)
    assert stats["escalated"] == 0 and stats["certified"] == 96
    assert stats["candidates"] <= 8 * 96
    assert agree_m == 1.0


@pytest.mark.parametrize("kind", ["counts13", "poisson", "signed"])
def test_int8_repeated_values_d60660(uh, oracle_mod, kind):
    """log2(count + 1) rows at the benchmark's d: few distinct values, each
    rounded the same way in bf16, so the BF16 error grows like d."""
    g = rng(17 + len(kind))
    d, r, b = 60660, 600, 48
    if kind == "counts13":
        refs = np.log2(g.integers(1, 4, (r, d)) + 1.0).astype(F32)
        q = np.log2(g.integers(1, 4, (b, d)) + 1.0).astype(F32)
    elif kind == "poisson":
        refs = np.log2(g.poisson(3.0, (r, d)) + 1.0).astype(F32)
        q = np.log2(g.poisson(3.0, (b, d)) + 1.0).astype(F32)
    else:
        refs = (g.standard_normal((r, d)) * 2).astype(F32)
        q = (refs[g.integers(0, r, b)] + g.standard_normal((b, d)) * 0.5).astype(F32)
    stats, _, _ = check_against_direct(uh, oracle_mod, refs, q, uh.STRATEGY_GRAM_INT8)
    assert stats["escalated"] + stats["certified"] == b


def test_int8_coherent_pair(uh, oracle_mod):
    refs, q = coherent_pair()
    k2 = oracle_mod.k2_distances(refs, q)
    exp_idx, exp_dist = oracle_mod.fp32_argmin(k2)
    assert exp_idx[0] == 1
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    index.enable_mixed(uh.STRATEGY_GRAM_BF16)
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(q, 1, ws, uh.STRATEGY_GRAM_INT8)
    assert idx[0] == 1 and dist.view(np.uint32)[0] == exp_dist.view(np.uint32)[0]
    assert ws.mixed_stats()["candidates"] == 1
    # The statistical BF16 bound is not a certificate: here it keeps only row 0.
    b_idx, _ = index.query_batch(q, 1, ws, uh.STRATEGY_GRAM_BF16)
    assert b_idx[0] == 0


def test_int8_cancellation_fixture(uh):
    base = F32(1048576.0)
    index = uh.HipDenseL2ScanIndex(np.array([[base + F32(0.25)], [base + F32(0.125)]], F32))
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(np.array([base], F32), 1, ws, uh.STRATEGY_GRAM_INT8)
    assert idx[0] == 1 and dist[0] == F32(0.015625)


def test_int8_escalation_to_direct_scan(uh, oracle_mod):
    g = rng(5)
    base = (g.random(512, dtype=np.float32) * 2).astype(F32)
    refs = np.tile(base, (100, 1))
    refs[37] += 0.5
    q = np.stack([base + F32(0.01), base]).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    ws = index.make_query_workspace(2)
    idx, dist = index.query_batch(q, 2, ws, uh.STRATEGY_GRAM_INT8)
    assert ws.mixed_stats()["escalated"] == 2
    exp_idx, exp_dist = oracle_mod.fp32_argmin(oracle_mod.k2_distances(refs, q))
    np.testing.assert_array_equal(idx, exp_idx)
    np.testing.assert_array_equal(dist, exp_dist)


ODD_SHAPES = [(1, 70, 5), (2, 1, 3), (255, 63, 257), (256, 64, 256), (257, 65, 255), (511, 129, 2),
              (33, 1000, 1), (600, 127, 513)]


@pytest.mark.parametrize("r,d,b", ODD_SHAPES)
def test_int8_odd_shapes(uh, oracle_mod, r, d, b):
    g = rng(r * 7 + d * 3 + b)
    refs = (g.random((r, d), dtype=np.float32) * 4).astype(F32)
    q = (g.random((b, d), dtype=np.float32) * 4).astype(F32)
    if r > 2:
        refs[-1] = refs[0]
        q[0] = refs[0]
    check_against_direct(uh, oracle_mod, refs, q, uh.STRATEGY_GRAM_INT8)


def test_int8_limits_and_errors(uh):
    index = uh.HipExactL2Index(np.ones((3, 4), F32))
    ws = index.make_query_workspace(1)
    with pytest.raises(uh.UhannInvalidArgument):
        index.query_batch(np.ones(4, F32), 1, ws, uh.STRATEGY_GRAM_INT8)  # not enabled
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipExactL1Index(np.ones((3, 4), F32)).enable_mixed(uh.STRATEGY_GRAM_INT8)
    with pytest.raises(uh.UhannLengthError):
        uh.HipExactL2Index(np.ones((2, 65537), F32)).enable_mixed(uh.STRATEGY_GRAM_INT8)


def test_int8_flat_index(uh, oracle_mod):
    """The flat L2 index runs the strategy on its gathered, scaled queries:
    same rows and bits as its direct path (no near ties on this data)."""
    g = rng(21)
    d, r, b = 3000, 400, 64
    refs = (g.random((r, d), dtype=np.float32) * 3).astype(F32)
    q = (refs[g.integers(0, r, b)] + g.random((b, d), dtype=np.float32) * 0.3).astype(F32)
    cols = np.sort(g.choice(d, 500, replace=False))
    sample = uh.make_sample(cols, g.integers(1, 4, cols.size), 1.0 / g.uniform(0.05, 1.0, cols.size))
    index = uh.HipSampledCoordinateL2AnnIndex(refs, sample)
    index.enable_mixed(uh.STRATEGY_GRAM_INT8)
    ws = index.make_query_workspace(b)
    d_idx, d_dist = index.query_batch(q, b, ws, uh.STRATEGY_DIRECT)
    m_idx, m_dist = index.query_batch(q, b, ws, uh.STRATEGY_GRAM_INT8)
    np.testing.assert_array_equal(m_idx, d_idx)
    np.testing.assert_array_equal(m_dist.view(np.uint32), d_dist.view(np.uint32))
