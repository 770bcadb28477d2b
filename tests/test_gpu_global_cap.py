"""Global grid with the nearest-first cap (nobs_max), as the GFS-sized config uses.

Lists: the oracle's uncapped obs_local (speedy/letkf/letkf_tools.f90:272-573;
bit-exact to the GPU's uncapped lists, test_gpu_das.py) capped nearest first:
the reference's cap (scale/letkf/letkf_tools.f90:1041-1156) with the key -rho,
so a list keeps its nobs_max entries of largest rho (ties: the earlier entry)
in the reference's storage-slot order (oracle.cap_nearest_first, pinned in
tests/test_oracle_regional.py).  The kept set is also checked against an
independent sort by (-rho, list position).  The GPU lists must equal the
oracle's in content and order, and the capped lists must hold u, v, T and q
obs (a key by err^2/rho would keep only the q obs, whose error is 1e-4).
Analysis: the reference letkf_core + apply (oracle/_ref) on the capped lists,
for the 3-D u, v, T and q rows and, with the level-1 transform, ps, within
1e-10 of the analysis perturbation scale; q above level 4 is the background
rebuilt from its mean and perturbations, bit for bit (the SPEEDY rules,
speedy/letkf/letkf_tools.f90:172-197).
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
CAP = 40


@pytest.fixture(scope="module")
def prob():
    p = synth.global_problem(nlon=64, lat=synth.gaussian_lats(32), sig=synth.SPEEDY_SIG, ne=24,
                             nstation=1200, seed=11)
    p.update(sigma_obs=700.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.0)
    return p


def _handle(p, cap, list_budget=None):
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                        sigma_obs=p["sigma_obs"], sigma_obsv=p["sigma_obsv"],
                        gross_error=p["gross_error"], infl_mode="fixed", infl_value=1.0,
                        nobs_max=cap)
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    return h


def _oracle_lists(p):
    o = p["obs"]
    sobs, grd, perm = oracle.set_letkf_obs_global(o, p["lon"], p["lat"], p["gross_error"])
    nlon, nlat, nlev = p["lon"].size, p["lat"].size, p["sig"].size
    nij = nlon * nlat
    lon1 = np.tile(p["lon"], nlat)
    lat1 = np.repeat(p["lat"], nlon)
    dz = p["sigma_obs"] * np.sqrt(10.0 / 3.0) * 2.0
    dzv = p["sigma_obsv"] * np.sqrt(10.0 / 3.0) * 2.0
    dlat = dz / oracle.PI_REF / oracle.RE_REF * 180.0
    g2 = p["gues2d"][:, :, 0]
    psm = g2[:, 0].copy()
    for m in range(1, p["ne"]):
        psm = psm + g2[:, m]
    psm = psm / p["ne"]
    lists = []
    for b in range(nij * nlev):
        lev, ij = divmod(b, nij)
        dlon = dlat / np.cos(oracle.PI_REF * lat1[ij] / 180.0)
        idx, rho, _ = oracle.obs_local_global(lon1[ij], lat1[ij], dlon, dlat,
                                              np.log(psm[ij] * p["sig"][lev]), lev + 1, p["lon"],
                                              p["lat"], grd, sobs, dz, dzv, p["sigma_obs"],
                                              p["sigma_obsv"])
        if idx.size > CAP:
            keep = np.lexsort((np.arange(idx.size), -rho))[:CAP]  # independent: a full sort
            cidx, crho = oracle.cap_nearest_first(idx, rho, CAP)
            np.testing.assert_array_equal(np.sort(cidx), np.sort(idx[keep]))
            idx, rho = cidx, crho
        lists.append((idx, rho))
    return sobs, lists


def test_global_cap_lists_and_analysis(gpu, prob):
    p = prob
    sobs, lists = _oracle_lists(p)
    h = _handle(p, CAP)
    h.run()
    h.sync()
    nl = h.nobsl()
    nij, nlev = h.nij, h.nlev
    assert max(len(l[0]) for l in lists) == CAP        # the cap binds somewhere
    rng = np.random.default_rng(2)
    ncapped = 0
    nall4 = 0
    for b in rng.choice(nij * nlev, 200, replace=False):
        idx, rho = h.local(b)
        ridx, rrho = lists[b]
        np.testing.assert_array_equal(idx, ridx)  # content and order
        np.testing.assert_allclose(rho, rrho, rtol=1e-13)
        lev, ij = divmod(int(b), nij)
        assert nl[ij, lev] == ridx.size
        if ridx.size == CAP:
            ncapped += 1
            kinds = set(np.round(sobs["elm"][idx]).astype(int).tolist())
            assert len(kinds) >= 3, kinds  # never a one-element list
            nall4 += {oracle.ID_U, oracle.ID_V, oracle.ID_T, oracle.ID_Q} <= kinds
    # u, v, T and q at one station and level share rho; a list misses one
    # only where the gross-error check rejected it at the nearest stations
    assert ncapped > 20 and nall4 >= 0.8 * ncapped, (nall4, ncapped)
    a3, a2 = h.analysis()
    off = np.cumsum([0] + [l[0].size for l in lists])
    ne = p["ne"]
    # the point's rows: u, v, T, q, and at level 1 also ps (speedy/letkf/
    # letkf_tools.f90:180-197: the 2-D variable takes the level-1 transform)
    xb = np.zeros((nlev * nij, 5, ne))
    xb[:, :4] = np.ascontiguousarray(p["gues3d"].transpose(1, 0, 3, 2)).reshape(nlev * nij, 4, ne)
    xb[:nij, 4] = p["gues2d"][:, :, 0]
    xa, _ = oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                                   np.concatenate([l[1] for l in lists]),
                                   np.ascontiguousarray(sobs["hdxf"].T), sobs["err"] ** 2,
                                   sobs["dep"], xb, np.ones(nlev * nij), fixed_infl=1.0)
    ref = xa[:, :4].reshape(nlev, nij, 4, ne).transpose(1, 0, 3, 2)
    dev = ref - ref.mean(axis=2, keepdims=True)
    upd = np.ones((nij, nlev, ne, 4), dtype=bool)
    upd[:, 4:, :, 3] = False  # q at levels 5.. is not analysed (letkf_tools.f90:172-177)
    for v in range(4):
        err = np.abs(a3[..., v] - ref[..., v])[upd[..., v]].max() / np.abs(dev[..., v]).max()
        assert err < 1e-10, (v, err)
    # the frozen q rows are the background rebuilt from its mean and
    # perturbations, mean summed in member order (common_speedy.f90:335-370)
    xq = p["gues3d"][:, 4:, :, 3]
    mq = xq[:, :, 0].copy()
    for m in range(1, ne):
        mq = mq + xq[:, :, m]
    mq = mq / ne
    np.testing.assert_array_equal(a3[:, 4:, :, 3], mq[:, :, None] + (xq - mq[:, :, None]))
    ps_ref = xa[:nij, 4]
    ps_dev = ps_ref - ps_ref.mean(axis=1, keepdims=True)
    assert np.abs(a2[:, :, 0] - ps_ref).max() / np.abs(ps_dev).max() < 1e-10


def test_global_cap_large_cap_equals_uncapped(gpu, prob):
    """A cap above every list length changes nothing."""
    p = prob
    h1 = _handle(p, 0)
    h1.run()
    h2 = _handle(p, 100000)
    h2.run()
    a1, b1 = h1.analysis()
    a2, b2 = h2.analysis()
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
