"""GPU parity of the local solve against the reference letkf_core.

The device path (lekf_core_f64 / lekf_core_batch through the C ABI) is
compared with the reference Fortran letkf_core (oracle/_ref, compiled from
common/common_letkf.f90 + common_mtx.f90 + EISPACK + reference BLAS) or, where
that library is absent, with the oracle's C restatement (bit-identical to it,
tests/test_oracle_core.py).  Tolerance: 1e-10 relative (the plain FP64 bar),
invariant to eigenvector sign and order because T = V f(L) V^T.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import core, synth

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _impl():
    return "ref" if oracle.ref_available() else "c"


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [(0, 20), (1, 20), (4, 20), (21, 20), (200, 20), (57, 3), (13, 5), (100, 16), (100, 31),
         (100, 32), (200, 40), (1000, 50), (300, 63), (200, 64), (150, 80), (120, 95),
         # p < k: C has k - p equal eigenvalues (k-1)/alpha (clustered spectrum)
         (0, 40), (1, 33), (5, 50), (30, 64), (10, 95),
         # large k: block Gram + eigensolver in global memory
         (200, 96), (300, 128), (150, 200), (400, 256), (1, 130), (0, 256), (60, 256)]


@pytest.mark.parametrize("p,k", CASES)
def test_core_matches_reference(gpu, p, k):
    h, rd, rl, dp = synth.local_problem(p, k, seed=1000 * p + k)
    ref = oracle.letkf_core(h, rd, rl, dp, 1.05, impl=_impl())
    tr, parm = core.letkf_core(k, max(p, 1), p, h if p else np.zeros((1, k)), rd if p else [1.0],
                               rl if p else [1.0], dp if p else [0.0], 1.05)
    assert _rel(tr, ref["trans"]) < RTOL
    assert abs(parm - ref["parm_infl"]) <= RTOL * abs(ref["parm_infl"])


@pytest.mark.parametrize("opts", [dict(transm=True), dict(pao=True), dict(transm=True, pao=True),
                                  dict(rdiag_wloc=True), dict(minfl=1.2),
                                  dict(transm=True, pao=True, rdiag_wloc=True, minfl=1.1)])
@pytest.mark.parametrize("p,k", [(0, 20), (37, 20), (250, 50), (180, 140)])
def test_core_options(gpu, opts, p, k):
    h, rd, rl, dp = synth.local_problem(p, k, seed=77 + p + k)
    if opts.get("rdiag_wloc"):
        rd = rd / np.maximum(rl, 1e-3) if p else rd
    ref = oracle.letkf_core(h, rd, rl, dp, 1.02, impl=_impl(), **opts)
    out = core.letkf_core(k, max(p, 1), p, h if p else np.zeros((1, k)), rd if p else [1.0],
                          rl if p else [1.0], dp if p else [0.0], 1.02, **opts)
    tr, parm = out[0], out[1]
    rest = list(out[2:])
    assert _rel(tr, ref["trans"]) < RTOL
    assert abs(parm - ref["parm_infl"]) <= RTOL * abs(ref["parm_infl"])
    if opts.get("transm"):
        tm = rest.pop(0)
        np.testing.assert_allclose(tm, ref["transm"], rtol=0, atol=RTOL * max(1.0, np.abs(ref["transm"]).max()))
    if opts.get("pao"):
        pa = rest.pop(0)
        assert _rel(pa, ref["pao"]) < RTOL


def test_leading_dimension(gpu):
    """nobs > nobsl: only the first nobsl rows of hdxb are read (common_letkf.f90:55,117)."""
    p, k, nobs = 30, 20, 500
    h, rd, rl, dp = synth.local_problem(nobs, k, seed=5)
    ref = oracle.letkf_core(h[:p], rd[:p], rl[:p], dp[:p], 1.0, impl=_impl())
    tr, parm = core.letkf_core(k, nobs, p, h, rd, rl, dp, 1.0)
    assert _rel(tr, ref["trans"]) < RTOL


def test_bad_sizes_raise(gpu):
    import letkf_amd
    with pytest.raises(letkf_amd.LetkfError) as e:
        core.letkf_core(20, 5, 6, np.zeros((5, 20)), np.ones(5), np.ones(5), np.zeros(5), 1.0)
    assert e.value.code == 99


@pytest.mark.parametrize("k", [20, 50, 160])
def test_batch_matches_single_and_reference(gpu, k):
    """Many uneven local problems in one launch; each equals its lone solve bit for bit."""
    from letkf_amd.device import DeviceArray, sync
    ps = [0, 1, 3, 17, 64, 150, 333, 5, 200, 90]
    nobs = 2000
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((nobs, k))
    Y -= Y.mean(axis=1, keepdims=True)
    R = rng.uniform(0.5, 2.0, nobs)
    D = rng.standard_normal(nobs)
    idx, rho, off = [], [], [0]
    for p in ps:
        sel = rng.choice(nobs, p, replace=False)
        idx.append(sel)
        rho.append(rng.uniform(0.01, 1.0, p))
        off.append(off[-1] + p)
    t = DeviceArray.from_numpy
    idx_all = np.concatenate(idx).astype(np.int32)
    rho_all = np.concatenate(rho)
    dY, dR, dD = t(Y), t(R), t(D)
    dOff = t(np.array(off), np.int64)
    dIdx, dRho = t(idx_all, np.int32), t(rho_all)
    npt = len(ps)
    infl = t(np.full(npt, 1.03))
    infl_out = DeviceArray.zeros(npt)
    trans = DeviceArray.zeros(npt * k * k)
    job = core.make_job(k=k, npt=npt, hdxf=dY, rdiag=dR, dep=dD, loc_off=dOff, loc_idx=dIdx,
                        loc_rho=dRho, infl_in=infl, infl_out=infl_out, trans=trans)
    core.letkf_core_batch([job])
    sync()
    tb = trans.numpy().reshape(npt, k, k)
    io = infl_out.numpy()
    for b, p in enumerate(ps):
        sel = idx[b]
        ref = oracle.letkf_core(Y[sel], R[sel], rho[b], D[sel], 1.03, impl=_impl())
        tb_b = tb[b].T  # stored column-major
        assert _rel(tb_b, ref["trans"]) < RTOL, (b, p)
        assert abs(io[b] - ref["parm_infl"]) <= RTOL * abs(ref["parm_infl"])
        # the same problem alone gives the same bits
        one = core.letkf_core(k, max(p, 1), p, Y[sel] if p else np.zeros((1, k)),
                              R[sel] if p else [1.0], rho[b] if p else [1.0],
                              D[sel] if p else [0.0], 1.03)[0]
        assert np.array_equal(one, tb_b), (b, p)


@pytest.mark.parametrize("k", [20, 50, 140])
def test_rank_deficient_point_flagged(gpu, k):
    """lambda_min < lambda_max * sqrt(eps) (common/common_mtx.f90:66-74): one obs with
    rdiag 1e-9 puts the top eigenvalue of (k-1)/rho I + Y^T R^-1 Y about 1e9 times above
    the others (condition past 1/sqrt(eps) = 6.7e7).  The C ABI returns
    LEKF_WARN_RANKDEF (the strict wrapper raises it) with finite outputs, where the
    reference truncates the rank and divides by the zeroed eigenvalues; the same
    problem with that obs's rdiag restored returns LEKF_OK."""
    from letkf_amd import _lib
    h, rd, rl, dp = synth.local_problem(30, k, seed=5 + k)
    bad = rd.copy()
    bad[3] = 1.0e-9
    with pytest.raises(_lib.LetkfError) as e:
        core.letkf_core(k, 30, 30, h, bad, rl, dp, 1.05, strict=True)
    assert e.value.code == _lib.LEKF_WARN_RANKDEF
    tr, parm = core.letkf_core(k, 30, 30, h, bad, rl, dp, 1.05)
    assert np.isfinite(tr).all() and np.isfinite(parm)
    tr0, _ = core.letkf_core(k, 30, 30, h, rd, rl, dp, 1.05, strict=True)
    assert np.isfinite(tr0).all()
