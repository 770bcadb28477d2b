"""The mixed forms on long capped lists against the reference letkf_core.

FP32 and split-BF16x3 Grams (Y^T R^-1 Y on MFMA, FP32 accumulation) with the
eigendecomposition, W and the apply in FP64, at the two ensemble sizes the
split Gram + register eigen kernels serve: k = 40 (the GFS-sized global grid,
16 levels, nearest-first cap 1000) and k = 50 (the radar grid, 40 x 40 x 60,
cap 1000 per bucket).  Reference: letkf_core + apply of the reference
Fortran (oracle/_ref, FP64) on the same local lists.  Error: max over
members of |xa - xa_ref| / the background spread (standard deviation over
members) of each point and variable, the north-star budget 1e-4.  The plain
FP64 form runs as the control, within 1e-10 of the perturbation scale.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
BUDGET = 1e-4


def _compare(a, ref, xb, prec, label):
    """a, ref, xb: (point, member, variable)."""
    spread = np.maximum(xb.std(axis=1, ddof=1), 1e-300)
    err = (np.abs(a - ref).max(axis=1) / spread).max()
    dev = np.abs(ref - ref.mean(axis=1, keepdims=True)).max(axis=(0, 1))
    rel = (np.abs(a - ref).max(axis=(0, 1)) / dev).max()
    print(f"{label} {prec}: max err / background spread {err:.2e}, "
          f"/ perturbation scale {rel:.2e}", flush=True)
    if prec == "f64":
        assert rel < 1e-10
    else:
        assert err <= BUDGET


@needs_ref
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_mixed_gfs_k40(gpu, prec):
    p = synth.gfs(ne=40, nlev=16)
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=40, nv3d=4, nv2d=1,
                        iq=p["iq"], q_top_lev=p["q_top_lev"], sigma_obs=p["sigma_obs"],
                        sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"],
                        infl_mode="fixed", infl_value=1.01, nobs_max=1000, prec=prec)
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    h.run()
    h.sync()
    a3, _ = h.analysis()
    nl = h.nobsl()
    nij, nlev = h.nij, h.nlev
    full = np.flatnonzero(nl.ravel() == 1000)
    assert full.size > 500
    rng = np.random.default_rng(40)
    pick = np.concatenate([rng.choice(full, 150, replace=False),
                           rng.choice(nij * nlev, 50, replace=False)])
    sobs, _, _ = oracle.set_letkf_obs_global(o, p["lon"], p["lat"], p["gross_error"])
    idxs, rhos, pts = [], [], []
    for f in pick:
        ij, lev = divmod(int(f), nlev)
        idx, rho = h.local(lev * nij + ij)
        idxs.append(idx)
        rhos.append(rho)
        pts.append((ij, lev))
    h.close()
    off = np.cumsum([0] + [i.size for i in idxs])
    xb = np.stack([p["gues3d"][ij, lev].T for ij, lev in pts])  # (point, variable, member)
    xa, _ = oracle.ref_solve_apply(off, np.concatenate(idxs), np.concatenate(rhos),
                                   np.ascontiguousarray(sobs["hdxf"].T), sobs["err"] ** 2,
                                   sobs["dep"], xb, np.full(len(pts), 1.01), fixed_infl=1.01)
    got = np.stack([a3[ij, lev] for ij, lev in pts])
    _compare(got, xa.transpose(0, 2, 1), xb.transpose(0, 2, 1), prec, "GFS-sized k=40")


@pytest.fixture(scope="module")
def radar40():
    return synth.radar(sub=(40, 40, 60), ne=50, seed=21)


@needs_ref
@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_mixed_radar_k50(gpu, radar40, prec):
    p = radar40
    h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                          ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"],
                          sigma_v=p["sigma_v"], nobs_max=1000, infl_mode="fixed",
                          infl_value=1.0, prec=prec)
    h.set_ensemble(p["gues3d"])
    h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
    h.run()
    h.sync()
    a3, _ = h.analysis()
    nl = h.nobsl()
    nij, nlev = h.nij, h.nlev
    long_ = np.flatnonzero(nl.ravel() >= 1500)
    assert long_.size > 500
    rng = np.random.default_rng(50)
    pick = np.concatenate([rng.choice(long_, 150, replace=False),
                           rng.choice(nij * nlev, 50, replace=False)])
    perm, _ = oracle.obs_bins_regional(p["ri"], p["rj"], p["nx"], p["ny"])
    idxs, rhos, pts = [], [], []
    for f in pick:
        ij, lev = divmod(int(f), nlev)
        idx, rho = h.local(lev * nij + ij)
        idxs.append(idx)
        rhos.append(rho)
        pts.append((ij, lev))
    h.close()
    off = np.cumsum([0] + [i.size for i in idxs])
    xb = np.stack([p["gues3d"][ij, lev].T for ij, lev in pts])
    xa, _ = oracle.ref_solve_apply(off, np.concatenate(idxs), np.concatenate(rhos),
                                   np.asarray(p["hdxf"])[perm], np.asarray(p["err"])[perm] ** 2,
                                   np.asarray(p["dep"])[perm], xb, np.ones(len(pts)),
                                   fixed_infl=1.0)
    got = np.stack([a3[ij, lev] for ij, lev in pts])
    _compare(got, xa.transpose(0, 2, 1), xb.transpose(0, 2, 1), prec, "radar k=50")
