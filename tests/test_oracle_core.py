"""Pins of the oracle's letkf_core / mtx_eigen restatement (CPU only).

* bit-identical to the reference Fortran (oracle/_ref built from
  common/common_letkf.f90, common_mtx.f90, netlib.f, netlibblas.f);
* closed forms: p = 0 (common_letkf.f90:82-96) and p = 1 (Sherman-Morrison);
* the dense Kalman-filter identities the LETKF transform satisfies
  (Hunt et al. 2007): mean update X w = K d with K = aPb H^T (H aPb H^T + R)^-1
  and analysis covariance Xa' Xa'^T / (k-1) = (I - K H) aPb, rho == 1;
* mtx_eigen against numpy.linalg.eigh (A V = V L, V^T V = I, descending).
A mutation check shows the Kalman identity fails when the mean term w-bar is
dropped from T, so the identity discriminates.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import synth

CASES = [(0, 20), (1, 20), (21, 20), (200, 40), (1000, 50), (50, 64), (5, 3), (300, 96)]
OPTS = [dict(), dict(transm=True), dict(pao=True), dict(transm=True, pao=True, rdiag_wloc=True),
        dict(minfl=1.1)]


@pytest.mark.parametrize("p,k", CASES)
@pytest.mark.parametrize("opts", OPTS)
def test_c_restatement_bit_identical_to_reference(ref_lib, p, k, opts):
    h, rd, rl, dp = synth.local_problem(p, k, seed=p * 7 + k)
    a = oracle.letkf_core(h, rd, rl, dp, 1.04, impl="c", **opts)
    b = oracle.letkf_core(h, rd, rl, dp, 1.04, impl="ref", **opts)
    assert np.array_equal(a["trans"], b["trans"])
    assert a["parm_infl"] == b["parm_infl"]
    if opts.get("transm"):
        assert np.array_equal(a["transm"], b["transm"])
    if opts.get("pao"):
        assert np.array_equal(a["pao"], b["pao"])


@pytest.mark.parametrize("n", [2, 3, 7, 20, 50])
def test_mtx_eigen_bit_identical_to_reference(ref_lib, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n))
    a = a @ a.T + n * np.eye(n)
    e1, v1, r1, _ = oracle.mtx_eigen(a, impl="c")
    e2, v2, r2, _ = oracle.mtx_eigen(a, impl="ref")
    assert np.array_equal(e1, e2) and np.array_equal(v1, v2) and r1 == r2


@pytest.mark.parametrize("n", [1, 2, 5, 20, 64])
def test_mtx_eigen_against_numpy(n):
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal((n, n))
    a = a @ a.T + 0.5 * np.eye(n)
    ev, vec, nrank, st = oracle.mtx_eigen(a)
    assert st == 0 and nrank == n
    assert np.all(np.diff(ev) <= 0)  # descending
    np.testing.assert_allclose(ev, np.linalg.eigvalsh(a)[::-1], rtol=1e-12, atol=1e-12 * ev[0])
    np.testing.assert_allclose(a @ vec, vec * ev, atol=1e-11 * ev[0])
    np.testing.assert_allclose(vec.T @ vec, np.eye(n), atol=1e-12)


def test_mtx_eigen_truncation():
    """lambda < lambda_max sqrt(eps) are zeroed with their vectors (common_mtx.f90:66-74)."""
    a = np.diag([1.0, 1e-9, 3.0])
    ev, vec, nrank, st = oracle.mtx_eigen(a)
    assert st == 0 and nrank == 2
    assert ev[-1] == 0.0 and np.all(vec[:, -1] == 0.0)


def test_p0_closed_form():
    k, a = 20, 1.07
    r = oracle.letkf_core(np.zeros((0, k)), [], [], [], a, transm=True, pao=True)
    np.testing.assert_array_equal(r["trans"], np.sqrt(a) * np.eye(k))
    np.testing.assert_array_equal(r["transm"], np.zeros(k))
    np.testing.assert_array_equal(r["pao"], a / (k - 1) * np.eye(k))
    assert r["parm_infl"] == a


@pytest.mark.parametrize("k", [3, 20, 50])
def test_p1_sherman_morrison(k):
    rng = np.random.default_rng(k)
    y = rng.standard_normal(k)
    rd, rl, d, a = 0.7, 0.4, 0.9, 1.05
    r = oracle.letkf_core(y[None, :], [rd], [rl], [d], a, transm=True, pao=True)
    c = (k - 1) / a
    w = rl / rd
    u = y / np.linalg.norm(y)
    lam1 = c + w * y @ y
    pa = (np.eye(k) - w * np.outer(y, y) / lam1) / c
    wa = np.sqrt(k - 1) * ((np.eye(k) - np.outer(u, u)) / np.sqrt(c) + np.outer(u, u) / np.sqrt(lam1))
    wbar = pa @ (w * y * d)
    np.testing.assert_allclose(r["pao"], pa, atol=1e-13)
    np.testing.assert_allclose(r["trans"], wa, atol=1e-13)
    np.testing.assert_allclose(r["transm"], wbar, atol=1e-13)


def _kalman_check(X, H, R, d, a, trans, wbar_included=True):
    k = X.shape[1]
    Y = H @ X
    Pb = a * X @ X.T / (k - 1)
    K = Pb @ H.T @ np.linalg.inv(H @ Pb @ H.T + R)
    wbar = trans.mean(axis=1) if wbar_included else None
    return K, Y, Pb


@pytest.mark.parametrize("n,p,k", [(6, 4, 10), (12, 30, 8), (30, 25, 20)])
def test_kalman_identities(n, p, k):
    rng = np.random.default_rng(n * p + k)
    X = rng.standard_normal((n, k))
    X -= X.mean(axis=1, keepdims=True)
    H = rng.standard_normal((p, n))
    r = rng.uniform(0.3, 2.0, p)
    d = rng.standard_normal(p)
    a = 1.1
    Y = H @ X
    out = oracle.letkf_core(Y, r, np.ones(p), d, a, transm=True)
    Wa, wbar = out["trans"], out["transm"]
    Pb = a * X @ X.T / (k - 1)
    K = Pb @ H.T @ np.linalg.inv(H @ Pb @ H.T + np.diag(r))
    np.testing.assert_allclose(X @ wbar, K @ d, rtol=1e-9, atol=1e-10)
    Xa = X @ Wa
    np.testing.assert_allclose(Xa @ Xa.T / (k - 1), (np.eye(n) - K @ H) @ Pb, rtol=1e-9, atol=1e-10)
    # Wa symmetric; zero-mean perturbations stay zero-mean: Wa 1 = sqrt(a) 1
    np.testing.assert_allclose(Wa, Wa.T, atol=1e-13)
    np.testing.assert_allclose(Wa @ np.ones(k), np.sqrt(a) * np.ones(k), atol=1e-12)


def test_kalman_identity_catches_dropped_mean_term():
    """Planted mistake (on the CPU, in the oracle's output): T without w-bar."""
    rng = np.random.default_rng(9)
    n, p, k = 10, 12, 8
    X = rng.standard_normal((n, k))
    X -= X.mean(axis=1, keepdims=True)
    H = rng.standard_normal((p, n))
    r = np.ones(p)
    d = rng.standard_normal(p)
    Y = H @ X
    T = oracle.letkf_core(Y, r, np.ones(p), d, 1.0)["trans"]
    Pb = X @ X.T / (k - 1)
    K = Pb @ H.T @ np.linalg.inv(H @ Pb @ H.T + np.eye(p))
    ok = np.allclose(X @ T.mean(axis=1), K @ d, atol=1e-10)
    Tbad = T - T.mean(axis=1, keepdims=True) + 1.0 / k  # mean term removed
    bad = np.allclose(X @ (Tbad @ np.ones(k) / k), K @ d, atol=1e-6)
    assert ok and not bad


def test_rdiag_wloc_equivalence():
    """rdiag_wloc with rdiag' = rdiag/rho equals the plain form (common_letkf.f90:101-113)."""
    h, rd, rl, dp = synth.local_problem(40, 12, seed=4)
    a = oracle.letkf_core(h, rd, rl, dp, 1.0)
    b = oracle.letkf_core(h, rd / rl, rl, dp, 1.0, rdiag_wloc=True)
    np.testing.assert_allclose(a["trans"], b["trans"], rtol=1e-12, atol=1e-13)
    assert abs(a["parm_infl"] - b["parm_infl"]) < 1e-12


def test_minfl_clamps_before_use():
    h, rd, rl, dp = synth.local_problem(30, 10, seed=8)
    a = oracle.letkf_core(h, rd, rl, dp, 0.9, minfl=1.2)
    b = oracle.letkf_core(h, rd, rl, dp, 1.2)
    np.testing.assert_array_equal(a["trans"], b["trans"])
    assert a["parm_infl"] == b["parm_infl"]
