"""Pins of the oracle's relaxation restatement (orc_relax_point,
scale/letkf/letkf_tools.f90:215-268): properties the formulas fix, and an
independent numpy restatement of the weight-space form."""
import numpy as np
import pytest

import oracle
from letkf_amd import synth


def _case(p, k, seed, nrow=3):
    h, rd, rl, dp = synth.local_problem(p, k, seed=seed)
    core = oracle.letkf_core(h, rd, rl, dp, 1.03, transm=True, pao=True)
    rng = np.random.default_rng(seed + 1)
    xb = rng.standard_normal((nrow, k)) * np.array([[1.0], [3.0], [0.2]])[:nrow] + \
        np.array([[5.0], [-2.0], [1.0]])[:nrow]
    return core, xb


def _plain(core, xb):
    """Unrelaxed analysis x-mean + X'(Wa + w-bar 1^T)."""
    g = xb - xb.mean(axis=1, keepdims=True)
    return xb.mean(axis=1, keepdims=True) + g @ (core["trans"] + core["transm"][:, None])


@pytest.mark.parametrize("p,k", [(40, 20), (0, 12), (150, 50)])
def test_no_relaxation_is_letkf(p, k):
    core, xb = _case(p, k, 3)
    xa, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0)
    np.testing.assert_allclose(xa, _plain(core, xb), rtol=0, atol=1e-12)


def test_rtpp_limits():
    core, xb = _case(60, 20, 5)
    g = xb - xb.mean(axis=1, keepdims=True)
    a0, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 1, 0.0)
    np.testing.assert_allclose(a0, _plain(core, xb), atol=1e-12)
    a1, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 1, 1.0)
    # alpha = 1: analysis perturbations are the background's, the mean still moves
    np.testing.assert_allclose(a1 - a1.mean(axis=1, keepdims=True), g, atol=1e-12)
    np.testing.assert_allclose(a1.mean(axis=1), _plain(core, xb).mean(axis=1), atol=1e-12)


def test_rtps_restores_background_spread():
    core, xb = _case(80, 24, 7)
    a, infl = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 2, 1.0)
    sb = xb.std(axis=1, ddof=1)
    sa = a.std(axis=1, ddof=1)
    np.testing.assert_allclose(sa, sb, rtol=1e-12)
    assert np.all(infl > 1.0)
    # alpha = 0.5: spread halfway (sigma_a_new = 0.5 sigma_b + 0.5 sigma_a)
    a5, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 2, 0.5)
    s0 = _plain(core, xb).std(axis=1, ddof=1)
    np.testing.assert_allclose(a5.std(axis=1, ddof=1), 0.5 * sb + 0.5 * s0, rtol=1e-12)


def test_beta_zero_gives_background_and_beta_interpolates():
    core, xb = _case(30, 16, 9)
    a, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 1, 0.3, beta=0.0)
    np.testing.assert_allclose(a, xb, atol=1e-12)
    a1, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0, beta=1.0)
    ah, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0, beta=0.25)
    np.testing.assert_allclose(ah, 0.25 * a1 + 0.75 * xb, atol=1e-12)


def test_q_spread_limit():
    core, xb = _case(50, 20, 11)
    xb[2] = np.abs(xb[2]) * 1e-3 + 2e-3      # a humidity-like row with large relative spread
    plain, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0)
    qs = plain[2].std(ddof=1) / plain[2].mean()
    lim = 0.5 * qs
    a, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0,
                              q_sprd_max=lim, q_row=2)
    np.testing.assert_allclose(a[2].std(ddof=1) / a[2].mean(), lim, rtol=1e-12)
    np.testing.assert_allclose(a[2].mean(), plain[2].mean(), rtol=1e-14)
    np.testing.assert_allclose(a[:2], plain[:2], atol=0)
    # a limit above the spread changes nothing
    b, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 0, 0.0,
                              q_sprd_max=2 * qs, q_row=2)
    np.testing.assert_array_equal(b, plain)


def test_rtps_numpy_restatement():
    core, xb = _case(120, 40, 13)
    g = xb - xb.mean(axis=1, keepdims=True)
    k = xb.shape[1]
    a, infl = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], 2, 0.7, beta=0.6)
    for r in range(xb.shape[0]):
        var_g = g[r] @ g[r]
        var_a = g[r] @ core["pao"] @ g[r]
        f = 0.7 * np.sqrt(var_g / (var_a * (k - 1))) - 0.7 + 1.0
        T = (core["trans"] * f + core["transm"][:, None]) * 0.6 + 0.4 * np.eye(k)
        np.testing.assert_allclose(a[r], xb[r].mean() + g[r] @ T, rtol=0, atol=1e-12 * np.abs(xb[r]).max())
        assert abs(infl[r] - f) < 1e-13
