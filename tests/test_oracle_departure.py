"""Pins of the oracle's SCALE departure / QC step (orc_obs_departure_scale;
scale/letkf/letkf_obs.f90:320-343,476-508,560-645, common_nml.f90:348-371,
596-598).

letkf_obs.f90 needs the SCALE library and MPI, so it is not built here.  Each
rule is pinned by known answers on both sides of its threshold, with small
integers and powers of ten that make the arithmetic exact.
"""
import numpy as np
import pytest

import oracle

REF, VR = oracle.ID_RADAR_REF, oracle.ID_RADAR_VR
Q = oracle.IQC


def one(elm, dat, err, hx, qc=None, **cfg):
    d, e, dep, hd, q = oracle.obs_departure_scale(
        [elm], [dat], [err], np.atleast_2d(np.asarray(hx, dtype=float)),
        None if qc is None else np.atleast_2d(qc), **cfg)
    return d[0], e[0], dep[0], hd[0], int(q[0])


def test_reflectivity_to_dbz():
    for z, dbz in ((1.0, 0.0), (10.0, 10.0), (100.0, 20.0), (1.0e3, 30.0), (1.0e5, 50.0)):
        assert one(REF, z, 5.0, [dbz, dbz + 20.0])[0] == dbz
    # below MIN_RADAR_REF = 10^(min_ref_dbz/10): MIN_RADAR_REF_DBZ + LOW_REF_SHIFT
    assert one(REF, 3.0, 5.0, [5.0, 25.0], min_radar_ref_dbz=5.0, low_ref_shift=-2.0)[0] == 3.0
    assert one(REF, 3.17, 5.0, [5.0, 25.0], min_radar_ref_dbz=5.0)[0] == 10.0 * np.log10(3.17)
    assert one(REF, 0.5, 5.0, [0.0, 20.0])[0] == 0.0                # 0 dBZ floor, no shift
    # outside [0, 1e10): undef, rejected as a bad obs before any member test
    for z in (-1.0, 1.0e10, 5.0e12):
        d, _, dep, hd, q = one(REF, z, 5.0, [20.0, 30.0])
        assert d == oracle.UNDEF and q == Q["obs_bad"] and dep == 0.0
        assert np.array_equal(hd, [20.0, 30.0])
    assert one(REF, 9.99e9, 5.0, [99.0, 100.0])[4] == 0


def test_member_echo_rule():
    thr = 15.0 + 1e-6
    above = np.nextafter(thr, np.inf)
    # obs with echo (30 dBZ): needs MIN_RADAR_REF_MEMBER_OBSREF members above 15 dBZ
    assert one(REF, 1.0e3, 100.0, [thr, thr, 0.0])[4] == Q["ref_mem"]      # strict >
    assert one(REF, 1.0e3, 100.0, [above, thr, 0.0])[4] == 0
    assert one(REF, 1.0e3, 100.0, [above, thr, 0.0], min_radar_ref_member_obsref=2)[4] == Q["ref_mem"]
    assert one(REF, 1.0e3, 100.0, [above, above, 0.0], min_radar_ref_member_obsref=2)[4] == 0
    # obs without echo (10 dBZ) uses MIN_RADAR_REF_MEMBER instead
    assert one(REF, 10.0, 100.0, [0.0, 0.0, 0.0])[4] == Q["ref_mem"]
    assert one(REF, 10.0, 100.0, [0.0, 0.0, 0.0], min_radar_ref_member=0)[4] == 0
    assert one(REF, 10.0, 100.0, [above, 0.0, 0.0], min_radar_ref_member_obsref=3)[4] == 0
    # the obs value itself is compared with the same threshold
    z_at = 10.0 ** (thr / 10.0)
    assert one(REF, z_at * (1 - 1e-12), 100.0, [0.0, 0.0], min_radar_ref_member=0)[4] == 0
    assert one(REF, z_at * (1 + 1e-9), 100.0, [0.0, 0.0], min_radar_ref_member=0)[4] == Q["ref_mem"]
    # a threshold below MIN_RADAR_REF_DBZ is raised to it
    assert one(REF, 1.0e3, 100.0, [12.0, 0.0], radar_ref_thres_dbz=5.0,
               min_radar_ref_dbz=11.0)[4] == 0
    assert one(REF, 1.0e3, 100.0, [10.5, 0.0], radar_ref_thres_dbz=5.0,
               min_radar_ref_dbz=11.0)[4] == Q["ref_mem"]


def test_use_flags_and_error_override():
    assert one(REF, 1.0e3, 5.0, [30.0, 31.0], use_radar_ref=False)[4] == Q["otype"]
    assert one(VR, 1.0, 3.0, [1.0, 2.0], use_radar_vr=False)[4] == Q["otype"]
    assert one(VR, 1.0, 3.0, [1.0, 2.0], use_radar_ref=False)[4] == 0
    assert one(REF, 1.0e3, 1.0, [30.0, 31.0], use_obserr_radar_ref=True)[1] == 5.0
    assert one(VR, 1.0, 1.0, [1.0, 2.0], use_obserr_radar_vr=True, obserr_radar_vr=2.5)[1] == 2.5
    assert one(VR, 1.0, 1.0, [1.0, 2.0], use_obserr_radar_ref=True)[1] == 1.0
    # the error is replaced even where the obs is rejected
    assert one(REF, -5.0, 1.0, [30.0, 31.0], use_obserr_radar_ref=True)[1:5:3] == (5.0, Q["obs_bad"])


def test_member_qc_maximum_ends_the_step():
    d, e, dep, hd, q = one(VR, 1.0, 3.0, [4.0, 8.0], qc=[0, 21])
    assert q == 21 and dep == 0.0 and np.array_equal(hd, [4.0, 8.0])
    assert one(VR, 1.0, 3.0, [4.0, 8.0], qc=[22, 19])[4] == 22
    assert one(VR, 1.0, 3.0, [4.0, 8.0], qc=[0, 0])[4] == 0


def test_mean_perturbations_and_departure():
    hx = [1.0, 2.0, 4.0, 9.0]                       # mean 4
    d, e, dep, hd, q = one(2819, 7.0, 1.0, hx)
    assert q == 0 and dep == 3.0
    assert np.array_equal(hd, [-3.0, -2.0, 0.0, 5.0])
    rng = np.random.default_rng(3)
    h = rng.standard_normal((50, 37)) * 4.0 + 20.0
    y = rng.standard_normal(50) + 20.0
    d, e, dep, hd, q = oracle.obs_departure_scale(np.full(50, VR), y, np.full(50, 10.0), h)
    assert np.all(q == 0)
    m = h.mean(axis=1)
    np.testing.assert_allclose(dep, y - m, rtol=0, atol=1e-13)
    np.testing.assert_allclose(hd, h - m[:, None], rtol=0, atol=1e-13)
    np.testing.assert_allclose(hd.sum(axis=1), 0.0, atol=1e-12)


@pytest.mark.parametrize("elm,key", [(oracle.ID_RAIN, "gross_error_rain"), (REF, "gross_error_radar_ref"),
                                     (VR, "gross_error_radar_vr"), (oracle.ID_RADAR_PRH, "gross_error_radar_prh"),
                                     (oracle.ID_TCLON, "gross_error_tcx"), (oracle.ID_TCLAT, "gross_error_tcy"),
                                     (oracle.ID_TCMIP, "gross_error_tcp")])
def test_gross_error_factor_of_each_element(elm, key):
    # departure 4 (H(x) mean 0, obs 4 dBZ = 10^0.4 mm^6 m^-3 for reflectivity), error 1
    dat = 10.0 ** 0.4 if elm == REF else 4.0
    y = one(elm, dat, 1.0, [0.0, 0.0], min_radar_ref_member=0)[0]
    hx = [y - 4.0, y - 4.0]
    kw = dict(min_radar_ref_member=0, min_radar_ref_member_obsref=0)
    assert one(elm, dat, 1.0, hx, gross_error=4.0, **kw)[4] == 0           # |d| = g err: kept
    assert one(elm, dat, 1.0, hx, gross_error=3.9, **kw)[4] == Q["gross_err"]
    assert one(elm, dat, 1.0, hx, gross_error=3.9, **{key: 4.5}, **kw)[4] == 0
    assert one(elm, dat, 1.0, hx, gross_error=9.0, **{key: 3.5}, **kw)[4] == Q["gross_err"]
    # other elements keep GROSS_ERROR whatever this element's factor is
    assert one(2819, 4.0, 1.0, [0.0, 0.0], gross_error=3.9, **{key: 9.0})[4] == Q["gross_err"]


def test_h08_rejected():
    with pytest.raises(ValueError):
        one(oracle.ID_H08IR, 250.0, 1.0, [250.0, 251.0])
