"""The oracle against the committed golden fixtures (CPU).

The fixtures were written by tests/golden/make_golden.py from the reference
Fortran (oracle/_ref).  The C restatement must reproduce them bit for bit.
"""
import os

import numpy as np

import oracle

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_core_cases_bit_identical():
    z = _load("core_cases.npz")
    parm0 = float(z["parm0"][0])
    for c, (p, k) in enumerate(z["cases"]):
        h, rd, rl, dp = z[f"c{c}_h"], z[f"c{c}_rd"], z[f"c{c}_rl"], z[f"c{c}_dp"]
        for o, (tm, pa, wl, mf) in enumerate(z["opts"]):
            r = oracle.letkf_core(h, rd / rl if wl else rd, rl, dp, parm0, transm=bool(tm),
                                  pao=bool(pa), rdiag_wloc=bool(wl),
                                  minfl=None if mf < 0 else float(mf))
            key = f"c{c}_o{o}"
            assert np.array_equal(r["trans"], z[key + "_trans"]), (p, k, o)
            assert r["parm_infl"] == z[key + "_parm"][0]
            if tm:
                assert np.array_equal(r["transm"], z[key + "_transm"])
            if pa:
                assert np.array_equal(r["pao"], z[key + "_pao"])


def test_eigen_cases_bit_identical():
    z = _load("eigen_cases.npz")
    for i in range(4):
        ev, vec, _, st = oracle.mtx_eigen(z[f"a{i}"])
        assert st == 0
        assert np.array_equal(ev, z[f"ev{i}"]) and np.array_equal(vec, z[f"vec{i}"])


def test_l96_selection_and_step():
    z = _load("l96_step.npz")
    nx = z["xb"].shape[0]
    off = z["loc_off"]
    for ix in range(nx):
        idx, rho = oracle.obs_local_l96(nx, z["gidx"], ix, 3.0)
        np.testing.assert_array_equal(idx, z["loc_idx"][off[ix]:off[ix + 1]])
        np.testing.assert_array_equal(rho, z["loc_rho"][off[ix]:off[ix + 1]])
    # assimilation never increases the (inflated) variance: var_a <= alpha var_b per point
    xb, xa = z["xb"], z["xa"]
    var_b = xb.var(axis=1, ddof=1)
    var_a = xa.var(axis=1, ddof=1)
    assert np.all(var_a <= 1.05 * var_b * (1 + 1e-12))
    assert np.all(var_a < var_b)


def test_global_patch_restatement():
    z = _load("global_patch.npz")
    sh, sv, ge, infl = z["sigma"]
    nlon, nlat, nlev = z["lon"].size, z["lat"].size, z["sig"].size
    p = dict(lon=z["lon"], lat=z["lat"], sig=z["sig"], nlon=nlon, nlat=nlat, nlev=nlev,
             nij=nlon * nlat, gues3d=z["gues3d"], gues2d=z["gues2d"], sigma_obs=sh,
             sigma_obsv=sv, gross_error=ge, infl=infl,
             obs=dict(elm=z["elm"], lon=z["olon"], lat=z["olat"], lev=z["olev"], dat=z["dat"],
                      err=z["err"], hxf=z["hxf"]))
    r = oracle.das_global_problem(p)
    np.testing.assert_array_equal(r["perm"], z["perm"])
    np.testing.assert_array_equal(r["nobsl"], z["nobsl"])
    np.testing.assert_array_equal(r["anal3d"], z["anal3d"])
    np.testing.assert_array_equal(r["anal2d"], z["anal2d"])
    np.testing.assert_array_equal(r["infl3d"], z["infl3d"])


def test_global_patch_is_the_reference_fortran():
    """The fixture's analysis is the reference Fortran letkf_core + apply
    (oracle/_ref) over the oracle's obs_local lists, bit for bit, as
    tests/golden/make_golden.py writes it."""
    import sys
    if not oracle.ref_available():
        import pytest
        pytest.skip("reference build absent")
    sys.path.insert(0, G)
    import make_golden
    z = _load("global_patch.npz")
    sh, sv, ge, infl = z["sigma"]
    nlon, nlat, nlev = z["lon"].size, z["lat"].size, z["sig"].size
    p = dict(lon=z["lon"], lat=z["lat"], sig=z["sig"], nlon=nlon, nlat=nlat, nlev=nlev,
             nij=nlon * nlat, ne=z["gues3d"].shape[2], gues3d=z["gues3d"], gues2d=z["gues2d"],
             sigma_obs=sh, sigma_obsv=sv, gross_error=ge, infl=infl,
             obs=dict(elm=z["elm"], lon=z["olon"], lat=z["olat"], lev=z["olev"], dat=z["dat"],
                      err=z["err"], hxf=z["hxf"]))
    r = oracle.das_global_problem(p)
    a3, a2, i3 = make_golden.patch_by_reference(p, r)
    np.testing.assert_array_equal(a3, z["anal3d"])
    np.testing.assert_array_equal(a2, z["anal2d"])
    np.testing.assert_array_equal(i3, z["infl3d"])
