"""Large ensembles (LEKF_KSPLIT < k <= LEKF_KMAX) through the das-level path.

For k > 95 the Gram is formed by the block kernel and the eigensolver works
in the point's workspace record in global memory.  Checks:

* Lorenz-96, nx = 40, k = 200 (the largest L96 size of the served stream):
  every point against the reference letkf_core + apply (oracle/_ref, built
  from the reference's own sources) on the GPU's local lists, which are the
  closed-form 21-obs windows (pinned bit-exact by the L96 golden test);
* a SPEEDY-type global grid at k = 256 (reduced 32 x 16 x 7 grid, 100 to 500
  local obs per point, so p is above and below k): a sample of points
  against the reference on the GPU's lists (selection does not depend on k
  and is pinned bit-exact by the SPEEDY-sized test);
* the same grid in FP32 and split-BF16x3: error <= 1e-4 of the background
  spread against the reference.
Tolerance of the FP64 path: 1e-10 of the analysis perturbation scale.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def _mean_member_order(a):
    s = a[:, 0].copy()
    for m in range(1, a.shape[1]):
        s = s + a[:, m]
    return s / a.shape[1]


def test_l96_k200(gpu):
    p = synth.l96(nx=40, ne=200, seed=31)
    nx, ne = p["nx"], p["ne"]
    h = das.L96Letkf(nx=nx, ne=ne, xlocal=3.0, infl_mode="fixed", infl_value=1.05)
    h.set_ensemble(p["xb"].reshape(nx, 1, ne, 1, order="F"))
    h.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h.run()
    a3, _ = h.analysis()
    xa = a3[:, 0, :, 0]
    assert np.all(h.nobsl() == 21)
    hm = _mean_member_order(p["hxf"])
    hd = p["hxf"] - hm[:, None]
    dep = p["y"] - hm
    lists = [h.local(b) for b in range(nx)]
    off = np.cumsum([0] + [l[0].size for l in lists])
    xb = np.ascontiguousarray(p["xb"]).reshape(nx, 1, ne)
    ref, infl = oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                                       np.concatenate([l[1] for l in lists]), hd, p["err"] ** 2,
                                       dep, xb, np.full(nx, 1.05), fixed_infl=1.05)
    ref = ref[:, 0, :]
    dev = ref - ref.mean(axis=1, keepdims=True)
    assert np.abs(xa - ref).max() / np.abs(dev).max() < RTOL


@pytest.fixture(scope="module")
def big():
    p = synth.global_problem(nlon=32, lat=synth.gaussian_lats(16), sig=synth.SPEEDY_SIG, ne=256,
                             nstation=420, seed=17)
    p.update(sigma_obs=600.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.0)
    sobs, _, _ = oracle.set_letkf_obs_global(p["obs"], p["lon"], p["lat"], p["gross_error"])
    return p, sobs


def _handle(p, prec):
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                        sigma_obs=p["sigma_obs"], sigma_obsv=p["sigma_obsv"],
                        gross_error=p["gross_error"], prec=prec, infl_mode="fixed",
                        infl_value=1.0)
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    h.run()
    h.sync()
    return h


def _reference(p, sobs, h, sample):
    """Reference analysis of u, v, T at the sampled points on the GPU's lists."""
    nij = h.nij
    lists = [h.local(int(b)) for b in sample]
    off = np.cumsum([0] + [l[0].size for l in lists])
    xb = np.stack([p["gues3d"][int(b) % nij, int(b) // nij, :, :3].T for b in sample])
    ref, _ = oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                                    np.concatenate([l[1] for l in lists]),
                                    np.ascontiguousarray(sobs["hdxf"].T), sobs["err"] ** 2,
                                    sobs["dep"], np.ascontiguousarray(xb), np.ones(len(sample)),
                                    fixed_infl=1.0)
    return ref, np.array([l[0].size for l in lists])


def _sample(h):
    rng = np.random.default_rng(5)
    return rng.choice(h.nij * h.nlev, 24, replace=False)


def test_global_k256_matches_reference(gpu, big):
    p, sobs = big
    h = _handle(p, "f64")
    sample = _sample(h)
    ref, nl = _reference(p, sobs, h, sample)
    assert nl.min() < 256 < nl.max()            # both p < k and p > k are exercised
    a3, _ = h.analysis()
    nij = h.nij
    for s, b in enumerate(sample):
        got = a3[int(b) % nij, int(b) // nij, :, :3].T     # (3, ne)
        dev = ref[s] - ref[s].mean(axis=1, keepdims=True)
        for v in range(3):
            err = np.abs(got[v] - ref[s, v]).max() / max(np.abs(dev[v]).max(), 1e-300)
            assert err < RTOL, (int(b), v, err)
    assert h.flags() == (0, 0)


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_global_k256_mixed_error_budget(gpu, big, prec):
    p, sobs = big
    h = _handle(p, prec)
    sample = _sample(h)
    ref, _ = _reference(p, sobs, h, sample)
    a3, _ = h.analysis()
    nij = h.nij
    worst = 0.0
    for s, b in enumerate(sample):
        ij, lev = int(b) % nij, int(b) // nij
        got = a3[ij, lev, :, :3].T
        spread = p["gues3d"][ij, lev, :, :3].std(axis=0, ddof=1)
        worst = max(worst, float((np.abs(got - ref[s]).max(axis=1) / spread).max()))
    assert worst <= 1e-4, worst


def test_vector_path_transm_and_inflation(gpu):
    """Jobs without trans/pao take the path that never forms V: w-bar (transm)
    and the adaptive inflation against the reference letkf_core."""
    from letkf_amd import core
    from letkf_amd.device import DeviceArray, sync
    k = 200
    ps = [0, 3, 150, 260, 90, 400]
    nobs = 1500
    rng = np.random.default_rng(8)
    Y = rng.standard_normal((nobs, k))
    Y -= Y.mean(axis=1, keepdims=True)
    R = rng.uniform(0.5, 2.0, nobs)
    D = rng.standard_normal(nobs)
    idx, rho, off = [], [], [0]
    for p in ps:
        idx.append(rng.choice(nobs, p, replace=False))
        rho.append(rng.uniform(0.01, 1.0, p))
        off.append(off[-1] + p)
    t = DeviceArray.from_numpy
    npt = len(ps)
    infl_out = DeviceArray.zeros(npt)
    transm = DeviceArray.zeros(npt * k)
    dY, dR, dD = t(Y), t(R), t(D)
    dOff = t(np.array(off), np.int64)
    dIdx = t(np.concatenate(idx).astype(np.int32), np.int32)
    dRho, dInfl = t(np.concatenate(rho)), t(np.full(npt, 1.04))
    job = core.make_job(k=k, npt=npt, hdxf=dY, rdiag=dR, dep=dD, loc_off=dOff, loc_idx=dIdx,
                        loc_rho=dRho, infl_in=dInfl, infl_out=infl_out, transm=transm)
    core.letkf_core_batch([job])
    sync()
    tm = transm.numpy().reshape(npt, k)
    io = infl_out.numpy()
    for b, p in enumerate(ps):
        sel = idx[b]
        ref = oracle.letkf_core(Y[sel], R[sel], rho[b], D[sel], 1.04, transm=True,
                                impl="ref" if oracle.ref_available() else "c")
        scale = max(1.0, float(np.abs(ref["transm"]).max()))
        assert np.abs(tm[b] - ref["transm"]).max() <= RTOL * scale, (b, p)
        assert abs(io[b] - ref["parm_infl"]) <= RTOL * abs(ref["parm_infl"]), (b, p)
