"""Pins of the oracle's SCALE radar operator (orc_phys2ijkz, orc_itpl_radar,
orc_calc_ref_vr, orc_radar_xtoy; scale/common/common_obs_scale.f90:338-476,
609-960,1298-1325 and phys2ijkz; scale/obs/obsope_tools.f90:353-380).

common_obs_scale.f90 needs the SCALE library modules, so it is not built
here.  What fixes the restatement instead:

* com_gamma and com_distll_1 of common/common.f90, built from the reference
  (oracle/_ref), against the oracle's Gamma values and distance;
* phys2ijkz on height columns linear in the level index (rk in closed form)
  and its range rules; itpl_3d reproducing fields linear in (i, j, k);
* calc_ref_vr: methods 1 and 2 (separate code) give the same rain-only
  reflectivity; method 3's power laws (doubling a mixing ratio or the air
  density scales Z by 2^b); continuity of the melting terms as snow or
  graupel vanish; the radial velocity's projection on the cardinal
  directions and the vertical; the method-1 terminal velocity in closed form;
* the dBZ conversion, the low-reflectivity value and the qc codes.
"""
import math

import numpy as np
import pytest

import oracle

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
F = oracle.RADAR_FIELDS


def _pt(**kw):
    d = dict(u=0.0, v=0.0, w=0.0, t=280.0, p=9.0e4, q=5e-3, qc=0.0, qr=0.0, qi=0.0, qs=0.0,
             qg=0.0)
    d.update(kw)
    return d


@needs_ref
def test_gamma_and_distance_match_reference():
    ref = oracle._lib_ref()
    for x, g in zip((4.8, 4.25, 4.5), oracle.radar_gammas()):
        assert abs(ref.ref_com_gamma(x) - g) <= 4 * np.spacing(g)
    rng = np.random.default_rng(1)
    for _ in range(50):
        a, b = rng.uniform(0, 360, 2), rng.uniform(-80, 80, 2)
        assert oracle.distll_1(a[0], b[0], a[1], b[1]) == ref.ref_com_distll_1(a[0], b[0], a[1], b[1])


def _heights(nx, ny, nlev, z0=100.0, dz=250.0):
    ij = np.arange(nx * ny)
    return np.asfortranarray(z0 + dz * np.arange(nlev)[None, :] + 0.0 * ij[:, None])


def test_phys2ijkz_linear_columns_and_rules():
    nx, ny, nlev = 6, 5, 10
    h = _heights(nx, ny, nlev)
    for z in (100.0, 101.0, 377.5, 1234.0, 2350.0):
        qc, rk = oracle.phys2ijkz(h, nx, ny, 3.3, 2.7, z)
        assert qc == 0
        assert abs(rk - (1.0 + (z - 100.0) / 250.0)) < 1e-13
    assert oracle.phys2ijkz(h, nx, ny, 3.3, 2.7, 2350.1)[0] == oracle.IQC["out_vhi"]
    assert oracle.phys2ijkz(h, nx, ny, 3.3, 2.7, 99.0)[0] == oracle.IQC["out_vlo"]
    for ri, rj in ((1.0, 2.0), (0.5, 2.0), (6.01, 2.0), (3.0, 1.0), (3.0, 5.5)):
        assert oracle.phys2ijkz(h, nx, ny, ri, rj, 500.0)[0] == oracle.IQC["out_h"]
    assert oracle.phys2ijkz(h, nx, ny, 6.0, 5.0, 500.0)[0] == 0
    # an invalid lowest height in one corner column lifts the lowest usable level
    h2 = h.copy(order="F")
    h2[(3 - 1) + nx * (3 - 1), 0] = -999.0
    qc, _ = oracle.phys2ijkz(h2, nx, ny, 3.5, 3.5, 200.0)
    assert qc == oracle.IQC["out_vlo"]
    qc, rk = oracle.phys2ijkz(h2, nx, ny, 3.5, 3.5, 400.0)
    assert qc == 0 and rk > 2.0


def test_itpl_reproduces_linear_fields():
    nx, ny, nlev = 7, 6, 5
    i = np.arange(nx)[:, None, None] + 1.0
    j = np.arange(ny)[None, :, None] + 1.0
    k = np.arange(nlev)[None, None, :] + 1.0
    f = 3.0 + 0.5 * i - 1.25 * j + 2.0 * k
    var = np.asfortranarray(f.reshape(nx * ny, nlev, order="F"))
    rng = np.random.default_rng(2)
    for _ in range(40):
        ri, rj, rk = rng.uniform(1.01, nx), rng.uniform(1.01, ny), rng.uniform(1.0, nlev)
        want = 3.0 + 0.5 * ri - 1.25 * rj + 2.0 * rk
        assert abs(oracle.itpl_radar(var, nx, ri, rj, rk) - want) < 1e-12
    assert abs(oracle.itpl_radar(var, nx, 2.5, 3.5, 1.0) - (3.0 + 1.25 - 4.375 + 2.0)) < 1e-13


def test_zero_hydrometeors_give_no_echo():
    for m in (1, 2, 3):
        ref, vr = oracle.calc_ref_vr(_pt(u=3.0, v=4.0), 0.0, 0.0, method=m)
        assert ref == 0.0 and vr == 4.0


def test_rain_only_methods_1_and_2_agree():
    for q in (1e-5, 3e-4, 2e-3):
        for t, p in ((290.0, 1.0e5), (270.0, 7.0e4)):
            r1, _ = oracle.calc_ref_vr(_pt(qr=q, t=t, p=p), 0.0, 0.0, method=1)
            r2, _ = oracle.calc_ref_vr(_pt(qr=q, t=t, p=p), 0.0, 0.0, method=2)
            assert r1 > 0.0 and abs(r1 - r2) <= 1e-14 * r1


@pytest.mark.parametrize("species,expo", [("qr", np.float32(1.84)), ("qs", np.float32(1.66)),
                                          ("qg", 1.5)])
def test_method3_power_laws(species, expo):
    base = _pt(**{species: 4e-4})
    r1, _ = oracle.calc_ref_vr(base, 0.0, 0.0, method=3)
    r2, _ = oracle.calc_ref_vr(dict(base, **{species: 8e-4}), 0.0, 0.0, method=3)
    assert abs(r2 / r1 - 2.0 ** float(expo)) < 1e-12 * 2.0 ** float(expo)
    # air density: doubling p at fixed T doubles rho
    r3, _ = oracle.calc_ref_vr(dict(base, p=2 * base["p"]), 0.0, 0.0, method=3)
    assert abs(r3 / r1 - 2.0 ** float(expo)) < 1e-12 * 2.0 ** float(expo)


def test_method3_melting_terms_vanish_continuously():
    rain, _ = oracle.calc_ref_vr(_pt(qr=1e-3), 0.0, 0.0, method=3)
    for sp in ("qs", "qg"):
        prev = None
        for tiny in (1e-12, 1e-15, 1e-18):
            r, _ = oracle.calc_ref_vr(_pt(qr=1e-3, **{sp: tiny}), 0.0, 0.0, method=3)
            d = abs(r - rain) / rain
            assert prev is None or d < prev
            prev = d
        assert prev < 1e-4
    # equal rain and snow: both melting fractions are at their maximum 1/2
    both, _ = oracle.calc_ref_vr(_pt(qr=1e-3, qs=1e-3), 0.0, 0.0, method=3)
    assert both > 0.0


def test_radial_velocity_projection():
    w = dict(u=3.0, v=-4.0, w=1.5)
    for az, el, want in ((0.0, 0.0, -4.0), (90.0, 0.0, 3.0), (180.0, 0.0, 4.0),
                         (270.0, 0.0, -3.0)):
        _, vr = oracle.calc_ref_vr(_pt(**w), az, el, method=3)
        assert abs(vr - want) < 1e-9
    _, vr = oracle.calc_ref_vr(_pt(**w), 37.0, 90.0, method=3)
    assert abs(vr - 1.5) < 1e-9        # pi is 3.1415926535 in the reference
    # method 1 terminal velocity at p = p0: wt = 5.4 qt^(1/8)
    qt = 2e-3
    _, vr = oracle.calc_ref_vr(_pt(w=1.5, qr=qt, p=1.0e5), 0.0, 90.0, method=1, use_wt=True)
    assert abs(vr - (1.5 - 5.4 * qt ** 0.125)) < 1e-8


def test_unknown_method_rejected():
    with pytest.raises(ValueError):
        oracle.calc_ref_vr(_pt(qr=1e-3), 0.0, 0.0, method=4)


def _column_problem(ne=3):
    nx, ny, nlev = 5, 4, 8
    h = _heights(nx, ny, nlev)
    rng = np.random.default_rng(4)
    x = np.zeros((nx * ny, nlev, ne, 11), order="F")
    for q, name in enumerate(F):
        base = dict(u=5.0, v=-2.0, w=0.5, t=280.0, p=8.0e4, q=5e-3, qc=1e-4, qr=1e-3, qi=0.0,
                    qs=2e-4, qg=1e-4)[name]
        x[..., q] = base * (1.0 + 0.1 * rng.standard_normal((nx * ny, nlev, ne)) * (name != "qi"))
    return nx, ny, h, x


def test_radar_xtoy_values_and_codes():
    nx, ny, h, x = _column_problem()
    obs = dict(elm=np.array([4001, 4002, 4001, 4001, 4002, 4001, 2819]),
               lon=np.array([140.0, 140.0, 140.2, 140.1, 140.0, 139.5, 140.0]),
               lat=np.array([35.1, 35.1, 35.0, 35.0, 35.0, 35.2, 35.1]),
               z=np.array([800.0, 800.0, 1.2e5, 500.0, 2000.0, 800.0, 800.0]),
               ri=np.array([3.2, 3.2, 3.2, 0.9, 2.5, 3.0, 3.2]),
               rj=np.array([2.4, 2.4, 2.4, 2.0, 2.5, 2.0, 2.4]))
    radar = dict(lon=np.array([140.0, 140.0, 140.0, 140.0, 140.0, 139.5, 140.0]),
                 lat=np.array([35.0, 35.0, 35.0, 35.0, 35.0, 35.2, 35.0]),
                 z=np.full(7, 50.0))
    hx, qc, rk = oracle.radar_xtoy(obs, radar, x, h, nx, ny, zmax=99.0e3)
    assert np.all(qc[0] == 0) and np.all(qc[1] == 0)
    assert np.all(qc[2] == oracle.IQC["radar_vhi"])
    assert np.all(qc[3] == oracle.IQC["out_h"])
    assert np.all(qc[5] == oracle.IQC["out_h"])       # the obs sits at the radar
    assert np.all(qc[6] == oracle.IQC["otype"])
    assert abs(rk[0] - (1.0 + (800.0 - 100.0) / 250.0)) < 1e-13
    # reflectivity is 10 log10 of calc_ref_vr at the interpolated point, and
    # an obs due north of the radar sees v cos(elev) + w sin(elev)
    for m in range(x.shape[2]):
        pt = {f: oracle.itpl_radar(np.asfortranarray(x[:, :, m, q]), nx, 3.2, 2.4, rk[0])
              for q, f in enumerate(F)}
        elev = math.atan2(750.0, oracle.distll_1(140.0, 35.1, 140.0, 35.0))  # radians
        ref, vr = oracle.calc_ref_vr(pt, 0.0, elev * 180.0 / 3.1415926535, method=3)
        assert abs(hx[0, m] - 10.0 * math.log10(ref)) < 1e-12
        assert abs(hx[1, m] - vr) < 1e-12
        assert abs(vr - (pt["v"] * math.cos(elev) + pt["w"] * math.sin(elev))) < 1e-9
    # an obs with no hydrometeors: the low-reflectivity value, qc good
    x0 = x.copy(order="F")
    x0[..., 7:] = 0.0
    hx0, qc0, _ = oracle.radar_xtoy(obs, radar, x0, h, nx, ny, min_ref_dbz=5.0, low_ref_shift=-3.0)
    assert np.all(qc0[0] == 0) and np.all(hx0[0] == 2.0)
