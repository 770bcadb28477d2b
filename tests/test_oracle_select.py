"""Pins of the oracle's selection and index build (CPU only).

* obs binning (speedy/letkf/letkf_obs.f90:243-313) equals a stable numpy
  lexsort by (lat band, lon cell, input order) with polar obs dropped;
* global obs_local (speedy/letkf/letkf_tools.f90:272-573) equals a brute-force
  scan of every obs with the same cutoffs wherever the reference's lat/lon
  rectangle covers the cutoff cap (mid latitudes), and is a subset of it
  near the poles, where the reference rectangle clips the cap;
* wrap-around rectangles at lon 0/360 agree with brute force;
* Lorenz-96 selection has p = 21 at every point for sigma = 3 (closed form).
"""
import numpy as np
import pytest

import oracle
from letkf_amd import synth

LAT = synth.SPEEDY_LAT
LON = 3.75 * np.arange(96)


def _brute_dist(olon, olat, glon, glat):
    pi = oracle.PI_REF
    lo1, lo2 = olon * pi * (1.0 / 180.0), glon * pi * (1.0 / 180.0)
    la1, la2 = olat * pi * (1.0 / 180.0), glat * pi * (1.0 / 180.0)
    c = np.sin(la1) * np.sin(la2) + np.cos(la1) * np.cos(la2) * np.cos(lo2 - lo1)
    return np.arccos(np.clip(c, -1.0, 1.0)) * oracle.RE_REF


def _obs(n, seed):
    rng = np.random.default_rng(seed)
    olat = np.rad2deg(np.arcsin(rng.uniform(-1, 1, n)))
    olon = rng.uniform(0, 360, n)
    return olon, olat


def test_bins_equal_stable_lexsort():
    olon, olat = _obs(5000, 1)
    olat[:5] = [90.0, -89.0, 87.159, -87.159, 87.2]  # edges and polar obs
    perm, grd = oracle.obs_bins_global(olon, olat, LON, LAT)
    band = np.searchsorted(LAT, olat, side="right") - 1
    ok = (band >= 0) & (band <= LAT.size - 2)
    cell = np.minimum(np.searchsorted(LON, olon, side="right") - 1, LON.size - 1)
    idx = np.flatnonzero(ok)
    order = idx[np.lexsort((idx, cell[idx], band[idx]))]
    np.testing.assert_array_equal(perm, order)
    # cumulative counts
    counts = np.zeros((LAT.size, LON.size), dtype=int)
    np.add.at(counts, (band[ok], cell[ok]), 1)
    cum = np.cumsum(counts[:-1].ravel()).reshape(LAT.size - 1, LON.size)
    np.testing.assert_array_equal(grd[:-1], cum)


def _setup(nobs=4000, seed=2):
    olon, olat = _obs(nobs, seed)
    perm, grd = oracle.obs_bins_global(olon, olat, LON, LAT)
    rng = np.random.default_rng(seed + 1)
    lv = rng.choice(synth.SPEEDY_SIG, perm.size) * 1.0e5
    obs = dict(elm=np.full(perm.size, float(oracle.ID_T)), lon=olon[perm], lat=olat[perm], lev=lv,
               dat=np.full(perm.size, 1.0e5), err=np.ones(perm.size))
    return obs, grd


@pytest.mark.parametrize("glat_i,glon", [(24, 100.0), (10, 0.0), (30, 356.25), (5, 180.0),
                                         (40, 3.75), (20, 357.0)])
def test_global_select_equals_brute_force(glat_i, glon):
    obs, grd = _setup()
    sigma, sigmav = 500.0e3, 0.4
    dz = sigma * np.sqrt(10.0 / 3.0) * 2.0
    dzv = sigmav * np.sqrt(10.0 / 3.0) * 2.0
    dlat = dz / oracle.PI_REF / oracle.RE_REF * 180.0
    glat = LAT[glat_i]
    dlon = dlat / np.cos(oracle.PI_REF * glat / 180.0)
    logp = np.log(1.0e5 * 0.51)
    idx, rl, rd = oracle.obs_local_global(glon, glat, dlon, dlat, logp, 2, LON, LAT, grd, obs, dz,
                                          dzv, sigma, sigmav)
    dist = _brute_dist(obs["lon"], obs["lat"], glon, glat)
    dlev = np.abs(np.log(obs["lev"]) - logp)
    sel = np.flatnonzero((dist <= dz) & (dlev <= dzv))
    if abs(glat) < 60.0:
        np.testing.assert_array_equal(np.sort(idx), sel)
    else:
        assert set(idx) <= set(sel)
    rho = np.exp(-0.5 * ((dist[idx] / sigma) ** 2 + (dlev[idx] / sigmav) ** 2))
    np.testing.assert_allclose(rl, rho, rtol=1e-14)
    np.testing.assert_array_equal(rd, obs["err"][idx] ** 2)


def test_rect_candidates_wrap_cases():
    obs, grd = _setup(3000, 5)
    nobs = obs["lon"].size
    for glon, dlon in [(0.0, 20.0), (355.0, 20.0), (10.0, 400.0), (180.0, 190.0), (1.0, 181.0)]:
        c = oracle.rect_candidates(glon, 0.0, dlon, 10.0, LON, LAT, grd, nobs)
        assert len(set(c)) == len(c)  # no obs scanned twice
        lo, hi = glon - dlon, glon + dlon
        if hi - lo >= 360.0:
            want_lon = np.ones(nobs, bool)
        else:
            x = obs["lon"]
            want_lon = ((x >= lo) & (x <= hi)) | ((x + 360 >= lo) & (x + 360 <= hi)) | (
                (x - 360 >= lo) & (x - 360 <= hi))
        inlat = np.abs(obs["lat"]) <= 10.0 - 1e-9
        core_set = np.flatnonzero(want_lon & inlat)
        # candidates cover every obs strictly inside the lat/lon box (bins are a superset)
        assert set(core_set) <= set(c)


def test_l96_closed_form():
    nx = 40
    for ix in (0, 5, 39):
        idx, rl = oracle.obs_local_l96(nx, np.arange(nx), ix, 3.0)
        assert idx.size == 21
        dd = np.minimum(np.abs(idx - ix), nx - np.abs(idx - ix))
        assert dd.max() == 10
        np.testing.assert_allclose(rl, np.exp(-0.5 * (dd / 3.0) ** 2), rtol=1e-15)
