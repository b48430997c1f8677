"""CPU checks of the engine's host-side index construction (C ABI:
uhann_rng_*, uhann_build_coordinate_sample, uhann_build_projection,
uhann_importance_probabilities). The index must be identical to the
reference's for the same seed, so every draw is compared with the oracle's
restatement and, where it is built, with the reference's own code."""

import numpy as np
import pytest

from conftest import rng as np_rng

import ultrahigh_ann_amd as uh


def _probs(seed, d=400):
    g = np_rng(seed)
    p = g.random(d) ** 3
    p[::9] = 0.0
    p[3] = 1.0
    return p


@pytest.mark.parametrize("reps", [1, 8, 32, 128])
@pytest.mark.parametrize("seed", [7, 20, 42, 1312, 2026])
def test_sample_matches_oracle(oracle_mod, reps, seed):
    p = _probs(seed)
    s = uh.build_coordinate_sample(p, reps, uh.Rng(seed))
    cols, mult, invp = oracle_mod.build_coordinate_sample(p, reps, seed)
    np.testing.assert_array_equal(s["source_column"], cols)
    np.testing.assert_array_equal(s["multiplicity"], mult)
    np.testing.assert_array_equal(s["inverse_probability"], invp)


def test_sample_known_answers():
    # tests/hnsvw25/l2/importance_sampling_test.cpp:291-305
    s = uh.build_coordinate_sample([1.0, 0.0], 16, uh.Rng(42))
    assert s.size == 1 and s[0]["source_column"] == 0 and s[0]["multiplicity"] == 16
    assert s[0]["inverse_probability"] == 1.0
    with pytest.raises(uh.UhannDomainError):
        uh.build_coordinate_sample([0.5, 1.5], 4, uh.Rng(1))
    with pytest.raises(uh.UhannDomainError):
        uh.build_coordinate_sample([np.nan], 4, uh.Rng(1))
    # zero probabilities draw nothing: the engine is untouched
    a, b = uh.Rng(5), uh.Rng(5)
    uh.build_coordinate_sample([0.0, 0.0, 0.0], 9, a)
    assert a == b


@pytest.mark.parametrize("metric", [uh.METRIC_L2, uh.METRIC_L1])
@pytest.mark.parametrize("p", [7, 16, 64])
def test_projection_matches_oracle(oracle_mod, metric, p):
    g = np_rng(3)
    refs = (g.random((12, 300), dtype=np.float32) * 3).astype(np.float32)
    probs = oracle_mod.importance_probabilities(metric == uh.METRIC_L1, refs)
    built = oracle_mod.build_hierarchical(metric == uh.METRIC_L1, refs, probs, 16, p, 1729)
    r = uh.Rng(1729)
    s = uh.build_coordinate_sample(probs, 16, r)
    proj = uh.build_projection(s, p, metric, r)
    np.testing.assert_array_equal(s["source_column"], built["cols"])
    np.testing.assert_array_equal(proj, built["p64"])


@pytest.mark.parametrize("metric", [uh.METRIC_L2, uh.METRIC_L1])
def test_probabilities_match_oracle(oracle_mod, metric):
    g = np_rng(4)
    refs = (g.random((30, 257), dtype=np.float32) * 5).astype(np.float32)
    refs[7] = refs[2]  # a zero-distance pair is skipped
    ours = uh.importance_probabilities(metric, refs)
    np.testing.assert_array_equal(ours, oracle_mod.importance_probabilities(metric == uh.METRIC_L1, refs))


def test_probabilities_closed_form():
    refs = np.array([[0, 0], [3, 4]], np.float32)
    np.testing.assert_allclose(uh.importance_probabilities(uh.METRIC_L2, refs), [9 / 25, 16 / 25], rtol=0, atol=1e-15)
    np.testing.assert_allclose(uh.importance_probabilities(uh.METRIC_L1, refs), [3 / 7, 4 / 7], rtol=0, atol=1e-15)
    with pytest.raises(uh.UhannInvalidArgument):
        uh.importance_probabilities(uh.METRIC_L2, np.array([[np.inf, 0]], np.float32))


@pytest.mark.parametrize("metric_l1", [False, True])
def test_hierarchical_draws_match_reference(ref_mod, metric_l1):
    g = np_rng(8)
    refs = (g.random((9, 150), dtype=np.float32) * 2).astype(np.float32)
    probs = ref_mod.importance_probabilities(metric_l1, refs)
    p = 17
    r = uh.Rng(99)
    s = uh.build_coordinate_sample(probs, 8, r)
    proj = uh.build_projection(s, p, uh.METRIC_L1 if metric_l1 else uh.METRIC_L2, r)
    d = 150
    c2 = np.zeros(d, np.uint64)
    m2 = np.zeros(d, np.uint64)
    i2 = np.zeros(d, np.float64)
    p32 = np.zeros(p * d, np.float32)
    r32 = np.zeros(9 * p, np.float32)
    k = ref_mod.ref().ref_hierarchical_fp32_build(int(metric_l1), ref_mod._p(refs, ref_mod._f32p), 9, d,
                                                  ref_mod._p(probs, ref_mod._f64p), 8, p, 99,
                                                  ref_mod._p(c2, ref_mod._u64p), ref_mod._p(m2, ref_mod._u64p),
                                                  ref_mod._p(i2, ref_mod._f64p), ref_mod._p(p32, ref_mod._f32p),
                                                  ref_mod._p(r32, ref_mod._f32p))
    assert k == s.size
    np.testing.assert_array_equal(s["source_column"], c2[:k])
    np.testing.assert_array_equal(proj.astype(np.float32).reshape(-1), p32[: p * k])


def test_rng_equality_and_draws():
    a, b = uh.Rng(1), uh.Rng(1)
    assert a == b
    x = a.next()
    assert a != b
    assert b.next() == x
    assert a == b
    # std::mt19937_64 with the default seed 5489 yields 14514284786278117030 first
    assert uh.Rng(5489).next() == 14514284786278117030


def test_uniform_probabilities_reference_fixture():
    # tests/coordinate_sampling/uniform_probabilities_test.cpp:26-86
    importance = np.array([0.5, 0.25, 0.0, 0.25])
    mass = uh.sampling_mass(importance)
    assert abs(mass - 1.0) < 1e-12
    uniform = uh.uniform_probabilities(importance.size, mass)
    assert uniform.shape == (4,)
    assert np.all(np.abs(uniform - 0.25) < 1e-12)
    assert abs(uh.sampling_mass(uniform) - mass) < 1e-12
    with pytest.raises(uh.UhannDomainError):
        uh.sampling_mass(np.array([np.nan]))
    with pytest.raises(uh.UhannDomainError):
        uh.uniform_probabilities(2, 3.0)
    # the remaining domain checks of the same functions
    with pytest.raises(uh.UhannDomainError):
        uh.sampling_mass(np.array([0.5, 1.5]))
    with pytest.raises(uh.UhannDomainError):
        uh.sampling_mass(np.array([-0.25]))
    with pytest.raises(uh.UhannDomainError):
        uh.uniform_probabilities(3, -1.0)
    with pytest.raises(uh.UhannDomainError):
        uh.uniform_probabilities(3, np.inf)
    assert uh.uniform_probabilities(0, 0.0).size == 0
    assert uh.uniform_probabilities(5, 5.0).tolist() == [1.0] * 5


def test_uniform_probabilities_match_reference(ref_mod):
    """Sequential-sum mass and mass / d, bit for bit with the reference's own
    functions on importance probabilities of random data."""
    g = np_rng(77)
    refs = (g.random((40, 3000), dtype=np.float32) * 3).astype(np.float32)
    for metric_l1 in (False, True):
        probs = ref_mod.importance_probabilities(metric_l1, refs)
        mine = uh.sampling_mass(probs)
        theirs = np.zeros(1, np.float64)
        assert ref_mod.ref().ref_sampling_mass(ref_mod._p(probs, ref_mod._f64p), probs.size,
                                               ref_mod._p(theirs, ref_mod._f64p)) == 0
        assert mine == theirs[0]
        # a different summation order would change the last bits; the sum in order is
        # also what numpy's sequential cumsum gives
        assert mine == np.cumsum(probs)[-1]
        u_mine = uh.uniform_probabilities(probs.size, mine)
        u_theirs = np.zeros(probs.size, np.float64)
        assert ref_mod.ref().ref_uniform_probabilities(probs.size, mine, ref_mod._p(u_theirs, ref_mod._f64p)) == 0
        np.testing.assert_array_equal(u_mine, u_theirs)
    bad = np.zeros(1, np.float64)
    assert ref_mod.ref().ref_uniform_probabilities(2, 3.0, ref_mod._p(bad, ref_mod._f64p)) == 1
