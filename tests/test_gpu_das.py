"""GPU parity of the das_letkf-level path (selection + solves + apply).

* Lorenz-96 step (config 1) against the golden fixture written by the
  reference Fortran letkf_core + apply;
* a reduced global sigma-level patch against the golden fixture;
* the full SPEEDY-sized config (96x48x7, k=20, ~12k obs) against the oracle's
  das_letkf restatement run on the same seeded inputs.
Selection (obs index, local lists, counts) must be bit-exact; the analysis
ensemble within 1e-10 of the analysis perturbation scale of each variable.
"""
import os

import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-10


def _scaled_err(a, ref, ens_axis, var_axis=None):
    """max |a - ref| / max |ref - mean(ref)| per variable."""
    dev = ref - ref.mean(axis=ens_axis, keepdims=True)
    if var_axis is None:
        return np.max(np.abs(a - ref)) / np.max(np.abs(dev))
    out = 0.0
    for v in range(ref.shape[var_axis]):
        sl = [slice(None)] * ref.ndim
        sl[var_axis] = v
        sl = tuple(sl)
        out = max(out, np.max(np.abs(a[sl] - ref[sl])) / max(np.max(np.abs(dev[sl])), 1e-300))
    return out


def test_l96_step_golden(gpu):
    z = np.load(os.path.join(G, "l96_step.npz"), allow_pickle=False)
    nx, ne = z["xb"].shape
    h = das.L96Letkf(nx=nx, ne=ne, xlocal=3.0, infl_mode="fixed", infl_value=1.05)
    h.set_ensemble(z["xb"].reshape(nx, 1, ne, 1, order="F"))
    h.set_obs(z["gidx"], z["y"], z["err"], z["hxf"])
    h.run()
    a3, _ = h.analysis()
    xa = a3[:, 0, :, 0]
    assert _scaled_err(xa, z["xa"], 1) < RTOL
    nl = h.nobsl()
    assert np.all(nl == 21)
    for b in (0, 17, 39):
        idx, rho = h.local(b)
        off = z["loc_off"]
        np.testing.assert_array_equal(idx, z["loc_idx"][off[b]:off[b + 1]])
        np.testing.assert_allclose(rho, z["loc_rho"][off[b]:off[b + 1]], rtol=1e-15)
    infl, _ = h.infl()
    assert np.all(infl == 1.05)


def _global_handle(p, prec="f64"):
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4,
                        nv2d=p["gues2d"].shape[2], ips=0, iq=3, q_top_lev=5,
                        sigma_obs=p["sigma_obs"], sigma_obsv=p["sigma_obsv"],
                        gross_error=p["gross_error"], prec=prec, infl_value=p["infl"])
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    return h


def _check_global(h, ref, p):
    perm, grd = h.obs_index()
    np.testing.assert_array_equal(perm, ref["perm"])
    np.testing.assert_array_equal(grd.T[:-1], ref["nobsgrd"][:-1])
    h.run()
    h.sync()
    np.testing.assert_array_equal(h.nobsl(), ref["nobsl"])
    a3, a2 = h.analysis()
    assert _scaled_err(a3, ref["anal3d"], 2, 3) < RTOL
    assert _scaled_err(a2, ref["anal2d"], 1, 2) < RTOL
    i3, _ = h.infl()
    np.testing.assert_allclose(i3, ref["infl3d"], rtol=1e-12)
    assert h.flags() == (0, 0)


def test_global_patch_golden(gpu):
    z = np.load(os.path.join(G, "global_patch.npz"), allow_pickle=False)
    sh, sv, ge, infl = z["sigma"]
    p = dict(lon=z["lon"], lat=z["lat"], sig=z["sig"], ne=z["gues3d"].shape[2],
             gues3d=np.asfortranarray(z["gues3d"]), gues2d=np.asfortranarray(z["gues2d"]),
             sigma_obs=sh, sigma_obsv=sv, gross_error=ge, infl=infl,
             obs=dict(elm=z["elm"], lon=z["olon"], lat=z["olat"], lev=z["olev"], dat=z["dat"],
                      err=z["err"], hxf=z["hxf"]))
    ref = dict(perm=z["perm"], nobsgrd=z["nobsgrd"], nobsl=z["nobsl"], anal3d=z["anal3d"],
               anal2d=z["anal2d"], infl3d=z["infl3d"])
    _check_global(_global_handle(p), ref, p)


@pytest.fixture(scope="module")
def speedy():
    p = synth.speedy()
    ref = oracle.das_global_problem(p)
    return p, ref


def test_speedy_full_config(gpu, speedy):
    p, ref = speedy
    h = _global_handle(p)
    _check_global(h, ref, p)
    # selection lists of a sample of points equal the oracle's obs_local in order
    lon1 = np.tile(p["lon"], p["nlat"])
    lat1 = np.repeat(p["lat"], p["nlon"])
    dz = p["sigma_obs"] * np.sqrt(10.0 / 3.0) * 2.0
    dzv = p["sigma_obsv"] * np.sqrt(10.0 / 3.0) * 2.0
    dlat = dz / oracle.PI_REF / oracle.RE_REF * 180.0
    ps = p["gues2d"][:, :, 0].mean(axis=1)
    rng = np.random.default_rng(0)
    for b in rng.choice(p["nij"] * p["nlev"], 40, replace=False):
        lev, ij = divmod(int(b), p["nij"])
        dlon = dlat / np.cos(oracle.PI_REF * lat1[ij] / 180.0)
        ps_m = p["gues2d"][ij, 0, 0]
        for m in range(1, p["ne"]):
            ps_m = ps_m + p["gues2d"][ij, m, 0]
        ps_m = ps_m / p["ne"]
        idx_r, rho_r, _ = oracle.obs_local_global(lon1[ij], lat1[ij], dlon, dlat,
                                                  np.log(ps_m * p["sig"][lev]), lev + 1, p["lon"],
                                                  p["lat"], ref["nobsgrd"], ref["sobs"], dz, dzv,
                                                  p["sigma_obs"], p["sigma_obsv"])
        idx, rho = h.local(b)
        np.testing.assert_array_equal(idx, idx_r)
        np.testing.assert_allclose(rho, rho_r, rtol=1e-13)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-4), ("bf16x3", 1e-4)])
def test_speedy_mixed_error_budget(gpu, speedy, prec, tol):
    """Mixed forms: analysis error <= 1e-4 of the background spread vs the FP64 oracle."""
    p, ref = speedy
    h = _global_handle(p, prec=prec)
    h.run()
    a3, _ = h.analysis()
    g = p["gues3d"]
    spread = g.std(axis=2, ddof=1)  # (nij, nlev, nv3d)
    err = np.abs(a3 - ref["anal3d"]).max(axis=2)
    rel = (err / np.maximum(spread, 1e-300))
    assert np.nanmax(rel) <= tol, np.nanmax(rel)
