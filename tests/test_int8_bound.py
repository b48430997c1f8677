"""CPU checks of the GRAM_INT8 arithmetic and its deterministic bound, on a
numpy model of the row conversion (int8_model in conftest.py, the kernel's
binary32 steps).

* The codes: a in [-127, 127], b in [-128, 127], n = 256 a + b, and the
  fixed-point value s n / 256 is within s / 512 of x (rint) up to the
  binary32 rounding of x (32512 / max|x|).
* The estimate ||q||^2 + ||v||^2 - 2 s_q s_v (H + X / 256) with exact
  integer sums H, X stays within the bound
  2 (||q|| ||e_v|| + ||e_q|| ||x^_v|| + lo_q lo_v) of the binary64 squared
  distance on log-expression-like data, on log2(count + 1) data whose values
  repeat (where a model of independent rounding errors fails) and on signed
  data. A plausible mistake (leaving out the low plane, X = 0) breaks it.
* Coherent rounding: on the fixed pair of rows of test_gpu_int8.py the
  statistical BF16 bound excludes the true nearest row, the INT8 bound keeps
  it and separates the two rows.
"""

import numpy as np
import pytest

from conftest import bf16_round, int8_model, rng

F32 = np.float32


def _stats(x, a, b, s, n):
    xd = x.astype(np.float64)
    hat = s.astype(np.float64)[:, None] * n.astype(np.float64) / 256.0
    return {
        "norm2": (xd * xd).sum(1),
        "norm": np.sqrt((xd * xd).sum(1)),
        "err": np.sqrt(((xd - hat) ** 2).sum(1)),
        "hat": np.sqrt((hat * hat).sum(1)),
        "lo": s.astype(np.float64) * np.sqrt((b.astype(np.float64) ** 2).sum(1)) / 256.0,
    }


def _estimate(q, refs, drop_lo=False):
    aq, bq, sq, nq_ = int8_model(q)
    av, bv, sv, nv_ = int8_model(refs)
    h = aq @ av.T
    x = np.zeros_like(h) if drop_lo else aq @ bv.T + bq @ av.T
    stq, stv = _stats(q, aq, bq, sq, nq_), _stats(refs, av, bv, sv, nv_)
    dot = (sq.astype(np.float64)[:, None] * sv.astype(np.float64)[None, :]) * (h + x / 256.0)
    est = stq["norm2"][:, None] + stv["norm2"][None, :] - 2.0 * dot
    beta = (stq["norm"][:, None] * stv["err"][None, :] + stq["err"][:, None] * stv["hat"][None, :]
            + stq["lo"][:, None] * stv["lo"][None, :])
    return est, 2.0 * beta


def _exact(q, refs):
    diff = q.astype(np.float64)[:, None, :] - refs.astype(np.float64)[None, :, :]
    return (diff * diff).sum(2)


def _data(kind, n, d, g):
    if kind == "logexpr":
        base = np.exp(1.5 * g.standard_normal(d))
        x = base * np.exp(0.7 * g.standard_normal((n, d))) * (g.random((n, d)) >= 0.3)
        return np.log2(x + 1).astype(F32)
    if kind == "counts13":
        return np.log2(g.integers(1, 4, (n, d)) + 1.0).astype(F32)
    if kind == "poisson":
        return np.log2(g.poisson(3.0, (n, d)) + 1.0).astype(F32)
    return (g.standard_normal((n, d)) * 3).astype(F32)


def test_codes_ranges_and_rounding():
    g = rng(3)
    x = np.concatenate([(g.standard_normal((40, 333)) * 10 ** g.uniform(-6, 6, (40, 1))).astype(F32),
                        np.zeros((1, 333), F32), np.full((1, 333), -7.5, F32)])
    a, b, s, n = int8_model(x)
    assert a.min() >= -127 and a.max() <= 127
    assert b.min() >= -128 and b.max() <= 127
    np.testing.assert_array_equal(n, 256 * a + b)
    hat = s.astype(np.float64)[:, None] * n / 256.0
    # rint of a binary32 product: within half a step plus the product's rounding
    step = s.astype(np.float64)[:, None] / 256.0
    assert np.all(np.abs(x - hat) <= step * (0.5 + 1e-3) + np.abs(x) * 2.0**-22)
    assert np.all(n[-2] == 0) and s[-2] == 1.0
    assert np.all(n[-1] == -32512)


@pytest.mark.parametrize("kind", ["logexpr", "counts13", "poisson", "signed"])
def test_bound_holds(kind):
    g = rng(len(kind))
    d = 3000
    refs = _data(kind, 60, d, g)
    q = _data(kind, 12, d, g)
    est, tau = _estimate(q, refs)
    ex = _exact(q, refs)
    assert np.all(np.abs(est - ex) <= tau)
    # the bound is far from vacuous: below the gap to the second-nearest row
    # for most queries on these data
    srt = np.sort(ex, axis=1)
    assert np.median(tau.min(1)) < np.median(srt[:, 1] - srt[:, 0]) or kind == "signed"


def test_dropping_the_low_plane_breaks_the_bound():
    g = rng(8)
    refs = _data("logexpr", 40, 2000, g)
    q = _data("logexpr", 8, 2000, g)
    est, tau = _estimate(q, refs, drop_lo=True)
    assert np.any(np.abs(est - _exact(q, refs)) > tau)


def test_coherent_pair_statistical_bound_fails_int8_holds():
    from test_gpu_int8 import coherent_pair

    refs, q = coherent_pair()
    ex = _exact(q, refs)[0]
    assert ex[1] < ex[0] - 50  # row 1 is the nearest
    # BF16 Gram estimate and the statistical (Hoeffding, K = 7) bound
    qb, rb = bf16_round(q), bf16_round(refs)
    nq = (q.astype(np.float64) ** 2).sum(1)
    nv = (refs.astype(np.float64) ** 2).sum(1)
    approx = (nq[:, None] + nv[None, :] - 2 * qb @ rb.T)[0]
    l4 = lambda v: (v ** 4).sum(-1) ** 0.25
    d = q.shape[1]
    tau_stat = (2 * (7 * (l4(q)[0] * l4(refs - rb) + l4(rb) * l4(q - qb)[0])
                     + 7 * np.sqrt((d + 15) // 16) * 2.0**-24 * np.sqrt((qb ** 2).sum()) * np.sqrt((rb ** 2).sum(1)))
                + 4 * 2.0**-24 * (nq[0] + nv))
    upper = (approx + tau_stat).min()
    assert not (approx[1] - tau_stat[1] <= upper)  # the true nearest row is excluded
    est, tau = _estimate(q, refs)
    assert np.all(np.abs(est - ex[None, :]) <= tau)
    assert est[0, 1] + tau[0, 1] < est[0, 0] - tau[0, 0]  # INT8 separates the rows
