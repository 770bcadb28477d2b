"""Driver options after the local solves: RTPP, RTPS, the q spread limit and
the regional boundary buffer (scale/letkf/letkf_tools.f90:215-268).

The GPU applies them as one pass over each point's analysis rows.  The check
recomputes sampled points the reference's way: the reference letkf_core
(oracle/_ref, with transm and pao) on the GPU's own local lists, then the
oracle's weight-space relaxation (orc_relax_point, pinned in
test_oracle_relax.py).  Tolerance 1e-10 of each row's background spread.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def _impl():
    return "ref" if oracle.ref_available() else "c"


def _compare(xa_gpu, xb, ref_rows, rows):
    for r in rows:
        scale = max(float(xb[r].std(ddof=1)), 1e-300)
        err = float(np.abs(xa_gpu[r] - ref_rows[r]).max()) / scale
        assert err < RTOL, (r, err)


# ---------------------------------------------------------------- regional
@pytest.fixture(scope="module")
def radar():
    p = synth.radar(sub=(24, 20, 8), ne=50, seed=3)
    p["nobs_max"] = 40
    perm, _ = oracle.obs_bins_regional(p["ri"], p["rj"], p["nx"], p["ny"])
    so = dict(hd=np.asarray(p["hdxf"])[perm], err2=np.asarray(p["err"])[perm] ** 2,
              dep=np.asarray(p["dep"])[perm])
    return p, so


def _regional(p, relax):
    h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                          ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"], sigma_v=p["sigma_v"],
                          nobs_max=p["nobs_max"], infl_value=1.0)
    h.set_ensemble(p["gues3d"])
    h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
    h.set_relax(**relax)
    h.run()
    h.sync()
    return h


def _beta(ij, p, width):
    if width <= 0:
        return 1.0
    i, j = ij % p["nx"] + 1, ij // p["nx"] + 1
    d = min(min(i, p["nx"] + 1 - i) * p["dx"], min(j, p["ny"] + 1 - j) * p["dy"]) / width
    return max(d, 0.0) if d < 1.0 else 1.0


@pytest.mark.parametrize("relax", [
    dict(mode="rtpp", alpha=0.5, boundary_width=3000.0),
    dict(mode="rtps", alpha=0.8, q_sprd_max=0.15, iq=4),
    dict(mode="none", q_sprd_max=0.15, iq=4, boundary_width=2500.0),
])
def test_regional_relaxation(gpu, radar, relax):
    p, so = radar
    h = _regional(p, relax)
    a3, _ = h.analysis()
    nij = p["nx"] * p["ny"]
    rng = np.random.default_rng(4)
    mode = {"none": 0, "rtpp": 1, "rtps": 2}[relax["mode"]]
    width = relax.get("boundary_width", 0.0)
    hit_beta = hit_q = False
    for b in rng.choice(nij * p["nlev"], 40, replace=False):
        lev, ij = divmod(int(b), nij)
        idx, rho = h.local(int(b))
        core = oracle.letkf_core(so["hd"][idx], so["err2"][idx], rho, so["dep"][idx], 1.0,
                                 transm=True, pao=True, impl=_impl())
        xb = p["gues3d"][ij, lev, :, :].T                       # (nv3d, ne)
        beta = _beta(ij, p, width)
        hit_beta |= beta < 1.0
        ref, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], mode,
                                    relax.get("alpha", 0.0), beta, relax.get("q_sprd_max", 0.0),
                                    relax.get("iq", -1))
        if relax.get("q_sprd_max", 0.0) > 0:
            plain, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], mode,
                                          relax.get("alpha", 0.0), beta)
            q = plain[relax["iq"]]
            hit_q |= q.std(ddof=1) / q.mean() > relax["q_sprd_max"]
        _compare(a3[ij, lev, :, :].T, xb, ref, range(p["nv3d"]))
    assert hit_beta or width == 0.0
    assert hit_q or relax.get("q_sprd_max", 0.0) == 0.0


# ---------------------------------------------------------------- global
def _global(p, relax, ne):
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=ne, nv3d=4, nv2d=1,
                        sigma_obs=p["sigma_obs"], sigma_obsv=p["sigma_obsv"],
                        gross_error=p["gross_error"], infl_mode="fixed", infl_value=1.0)
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    h.set_relax(**relax)
    h.run()
    h.sync()
    return h


@pytest.mark.parametrize("ne,relax", [
    (20, dict(mode="rtps", alpha=0.9, q_sprd_max=0.05, iq=3)),
    (20, dict(mode="rtpp", alpha=0.3)),
    (130, dict(mode="rtps", alpha=0.6)),
])
def test_global_relaxation(gpu, ne, relax):
    p = synth.global_problem(nlon=32, lat=synth.gaussian_lats(16), sig=synth.SPEEDY_SIG, ne=ne,
                             nstation=300, seed=21)
    p.update(sigma_obs=700.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.0)
    sobs, _, _ = oracle.set_letkf_obs_global(p["obs"], p["lon"], p["lat"], p["gross_error"])
    h = _global(p, relax, ne)
    a3, a2 = h.analysis()
    nij = h.nij
    mode = {"none": 0, "rtpp": 1, "rtps": 2}[relax["mode"]]
    rng = np.random.default_rng(6)
    hd = np.ascontiguousarray(sobs["hdxf"].T)
    for b in rng.choice(nij * 4, 24, replace=False):          # levels 0..3: q is updated
        lev, ij = divmod(int(b), nij)
        idx, rho = h.local(int(b))
        core = oracle.letkf_core(hd[idx], sobs["err"][idx] ** 2, rho, sobs["dep"][idx], 1.0,
                                 transm=True, pao=True, impl=_impl())
        xb = p["gues3d"][ij, lev, :, :].T                       # (4, ne)
        got = a3[ij, lev, :, :].T
        if lev == 0:
            xb = np.vstack([xb, p["gues2d"][ij, :, 0]])
            got = np.vstack([got, a2[ij, :, 0]])
        ref, _ = oracle.relax_point(xb, core["trans"], core["transm"], core["pao"], mode,
                                    relax.get("alpha", 0.0), 1.0, relax.get("q_sprd_max", 0.0),
                                    relax.get("iq", -1))
        _compare(got, xb, ref, range(xb.shape[0]))


def test_relax_arguments_checked(gpu):
    p = synth.l96(nx=40, ne=20)
    h = das.L96Letkf(nx=40, ne=20)
    import letkf_amd
    with pytest.raises(letkf_amd.LetkfError):
        h.set_relax(mode="rtps", alpha=-1.0)
    with pytest.raises(letkf_amd.LetkfError):
        h.set_relax(mode="none", boundary_width=1000.0)   # boundary buffer: regional only
    with pytest.raises(letkf_amd.LetkfError):
        h.set_relax(mode="none", q_sprd_max=0.1, iq=3)     # L96 has one 3-D row
    h.set_relax(mode="rtpp", alpha=0.0)                    # RTPP with 0 = plain analysis
    h.set_ensemble(p["xb"].reshape(40, 1, 20, 1, order="F"))
    h.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h.run()
    a, _ = h.analysis()
    h2 = das.L96Letkf(nx=40, ne=20)
    h2.set_ensemble(p["xb"].reshape(40, 1, 20, 1, order="F"))
    h2.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h2.run()
    b, _ = h2.analysis()
    dev = b - b.mean(axis=2, keepdims=True)
    assert np.abs(a - b).max() <= 1e-12 * np.abs(dev).max()
