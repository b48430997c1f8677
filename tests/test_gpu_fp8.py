"""GPU checks of the FP8 shortlist strategy (UHANN_STRATEGY_GRAM_FP8).

The e4m3 estimates are pinned against an independent restatement of the
quantisation (power-of-two row scale 2^ceil(log2(max|x|/448)), round to
nearest even, subnormals as multiples of 2^-9) and exact products of the
quantised values, within the binary32 accumulation bound; a plausible
mistake (a non-power-of-two scale max|x|/448) falls outside that bound.
Query results must be a shortlisted row's binary32 reference-order distance,
bit for bit, and the binary32 direct answer whenever that row is in the
shortlist."""

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu
F32 = np.float32


def row_scales(x, power_of_two=True):
    mx = np.abs(x.astype(F32)).max(axis=1).astype(np.float64)
    if not power_of_two:
        return np.where(mx > 0, mx / 448.0, 1.0)
    f, e = np.frexp(mx)
    c = np.where(f <= 0.875, e - 9, e - 8)
    c = np.where(mx > 0, np.maximum(c, -126), 0)
    return np.ldexp(1.0, c)


def e4m3_round(v):
    """Round |v| <= 448 (binary64) to e4m3, nearest even; returns binary64."""
    a = np.abs(v)
    _, e = np.frexp(np.where(a > 0, a, 1.0))
    quantum = np.ldexp(1.0, e - 1 - 3)  # 3 mantissa bits below the leading one
    normal = np.round(a / quantum) * quantum
    sub = np.round(a * 512.0) / 512.0
    out = np.where(a >= 2.0**-6, normal, sub)
    return np.sign(v) * out


def fp8_rows(x, power_of_two=True):
    s = row_scales(x, power_of_two)
    return e4m3_round(x.astype(np.float64) / s[:, None]), s


def test_e4m3_helper_known_values():
    v = np.array([1.0, 1.0625, 1.1875, 448.0, 447.0, 2.0**-6, 2.0**-7, 3 * 2.0**-10, 2.0**-10, -17.0, 0.0])
    want = np.array([1.0, 1.0, 1.25, 448.0, 448.0, 2.0**-6, 2.0**-7, 2.0**-8, 0.0, -16.0, 0.0])
    np.testing.assert_array_equal(e4m3_round(v), want)


def _case(seed, r, d, b):
    g = rng(seed)
    refs = (g.standard_normal((r, d)) * np.exp(g.standard_normal((r, 1)))).astype(F32)
    q = (g.standard_normal((b, d)) * 2).astype(F32)
    refs[:, : d // 8] *= 1e-3  # small coordinates: e4m3 subnormals after scaling
    refs[1] = 0.0              # zero row: scale 1
    return refs, q


# (300, 4096, 40) and (600, 8192, 300) split their tiles along K (fewer tiles than compute units)
@pytest.mark.parametrize("r,d,b", [(300, 1000, 40), (77, 3000, 9), (40, 130, 70), (300, 4096, 40), (600, 8192, 300)])
def test_fp8_estimates_match_emulation(uh, r, d, b):
    import torch

    refs, q = _case(r + d, r, d, b)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(b)
    out = torch.full((b, r), float("nan"), dtype=torch.float32, device="cuda")
    index.fp8_estimates(torch.from_numpy(q).cuda(), b, ws, out)
    torch.cuda.synchronize()
    est = out.cpu().numpy().astype(np.float64)
    q8, sq = fp8_rows(q)
    v8, sv = fp8_rows(refs)
    nq = (q.astype(np.float64) ** 2).sum(1)
    nv = (refs.astype(np.float64) ** 2).sum(1)
    dot = q8 @ v8.T
    want = nq[:, None] + nv[None, :] - 2.0 * sq[:, None] * sv[None, :] * dot
    # The FP8 MFMA sums its 64 products per instruction with less than
    # binary32 precision: errors up to 2.4 u per (padded) product, relative
    # to sum |q8 v8|, were measured (integer data whose partial sums are
    # representable come out exact). The bound allows 8 u per product.
    products = (d + 127) // 128 * 128
    bound = (8.0 * products * 2.0**-24 * sq[:, None] * sv[None, :] * (np.abs(q8) @ np.abs(v8).T)
             + 4 * 2.0**-24 * (nq[:, None] + nv[None, :]))
    assert np.all(np.abs(est - want) <= bound)
    # a plausible mistake -- the unrounded scale max|x| / 448 -- is outside the bound
    q8w, sqw = fp8_rows(q, False)
    v8w, svw = fp8_rows(refs, False)
    wrong = nq[:, None] + nv[None, :] - 2.0 * sqw[:, None] * svw[None, :] * (q8w @ v8w.T)
    assert np.any(np.abs(est - wrong) > bound)


@pytest.mark.parametrize("d", [384, 4096])
def test_fp8_estimates_exact_on_small_integers(uh, d):
    """Integers |x| <= 8 are exact in e4m3 under the power-of-two scale and
    every partial sum is representable: the estimates are exact (d = 4096
    also sums K parts)."""
    import torch

    g = rng(1)
    refs = g.integers(-8, 9, (40, d)).astype(F32)
    q = g.integers(-8, 9, (33, d)).astype(F32)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(33)
    out = torch.zeros((33, 40), dtype=torch.float32, device="cuda")
    index.fp8_estimates(torch.from_numpy(q).cuda(), 33, ws, out)
    torch.cuda.synchronize()
    qd, rd = q.astype(np.float64), refs.astype(np.float64)
    want = (qd**2).sum(1)[:, None] + (rd**2).sum(1)[None, :] - 2.0 * qd @ rd.T
    np.testing.assert_array_equal(out.cpu().numpy().astype(np.float64), want)


def _check_results(uh, oracle_mod, refs, q, index, ws, idx, dist, k=16):
    import torch

    b, r = q.shape[0], refs.shape[0]
    k2 = oracle_mod.k2_distances(refs, q)
    rows = np.arange(b)
    chosen = idx.astype(np.int64)
    # the answer's distance is its binary32 reference-order distance, bit for bit
    np.testing.assert_array_equal(dist.view(np.uint32), k2[rows, chosen].view(np.uint32))
    out = torch.zeros((b, r), dtype=torch.float32, device="cuda")
    index.fp8_estimates(torch.from_numpy(q).cuda(), b, ws, out)
    torch.cuda.synchronize()
    est = out.cpu().numpy()
    # shortlist = k smallest estimates (ties to the lower row)
    order = np.lexsort((np.tile(np.arange(r), (b, 1)), est), axis=1)[:, : min(k, r)]
    in_short = np.array([chosen[i] in order[i] for i in range(b)])
    assert in_short.all()
    exp_idx, _ = oracle_mod.fp32_argmin(k2)
    best_in = np.array([exp_idx[i] in order[i] for i in range(b)])
    table = oracle_mod.fp64_table(False, refs, q)
    # when the binary32 answer is shortlisted the result is it, or a binary64 near tie that is not worse
    same = chosen == exp_idx.astype(np.int64)
    ok = same | (table[rows, chosen] <= table[rows, exp_idx.astype(np.int64)])
    assert np.all(ok[best_in])
    return np.mean(same)


@pytest.mark.parametrize("r,d,b", [(300, 1000, 40), (1000, 4096, 33)])
def test_fp8_strategy_random(uh, oracle_mod, r, d, b):
    refs, q = _case(7 * r + d, r, d, b)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws, uh.STRATEGY_GRAM_FP8)
    _check_results(uh, oracle_mod, refs, q, index, ws, idx, dist)
    st = ws.mixed_stats()
    assert st["candidates"] == b * min(16, r) and st["escalated"] == 0
    assert 0 <= ws.shortlist_stats()["certified"] <= b


@pytest.mark.parametrize("k", [1, 4, 8])
def test_fp8_shortlist_length(uh, oracle_mod, k):
    """A shorter shortlist re-ranks exactly the k smallest estimates: the
    answer is a shortlisted row's binary32 distance, and the direct answer
    when that row is shortlisted; lengths outside 1..16 are rejected."""
    r, d, b = 300, 1000, 40
    refs, q = _case(7 * r + d, r, d, b)
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    index.set_fp8_shortlist(k)
    ws = index.make_query_workspace(b)
    idx, dist = index.query_batch(q, b, ws, uh.STRATEGY_GRAM_FP8)
    _check_results(uh, oracle_mod, refs, q, index, ws, idx, dist, k)
    assert ws.mixed_stats()["candidates"] == b * k
    for bad in (0, 17):
        with pytest.raises(uh.UhannInvalidArgument):
            index.set_fp8_shortlist(bad)


def test_fp8_strategy_log_expression(uh, oracle_mod):
    from ultrahigh_ann_amd import synthetic

    refs, q = synthetic.log_expression(700, 64, 20000, seed=12, device="cuda")
    refs, q = refs.cpu().numpy(), q.cpu().numpy()
    index = uh.HipExactL2Index(refs)
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(64)
    idx, dist = index.query_batch(q, 64, ws, uh.STRATEGY_GRAM_FP8)
    agree = _check_results(uh, oracle_mod, refs, q, index, ws, idx, dist)
    assert agree == 1.0  # on this data the exact neighbour is always shortlisted


def test_fp8_cancellation_fixture_and_flat(uh, oracle_mod):
    base = F32(1048576.0)
    index = uh.HipDenseL2ScanIndex(np.array([[base + F32(0.25)], [base + F32(0.125)]], F32))
    index.enable_mixed(uh.STRATEGY_GRAM_FP8)
    ws = index.make_query_workspace(1)
    idx, dist = index.query_batch(np.array([base], F32), 1, ws, uh.STRATEGY_GRAM_FP8)
    assert idx[0] == 1 and dist[0] == F32(0.015625)
    # flat L2 index: FP8 over the gathered, scaled coordinates
    g = rng(3)
    refs = (g.random((200, 900), dtype=np.float32) * 3).astype(F32)
    q = (g.random((30, 900), dtype=np.float32) * 3).astype(F32)
    probs = oracle_mod.importance_probabilities(False, refs)
    flat = uh.HipFlatL2AnnIndex(refs, probs, 32, uh.Rng(5))
    fws = flat.make_query_workspace(30)
    d_idx, _ = flat.query_batch(q, 30, fws)
    flat.enable_mixed(uh.STRATEGY_GRAM_FP8)
    f_idx, _ = flat.query_batch(q, 30, fws, uh.STRATEGY_GRAM_FP8)
    assert np.mean(f_idx == d_idx) >= 0.9
    with pytest.raises(uh.UhannInvalidArgument):
        uh.HipExactL1Index(refs).enable_mixed(uh.STRATEGY_GRAM_FP8)
