"""GPU parity of SCALE's departure and QC step (lekf_obs_departure_scale)
against the oracle's restatement (orc_obs_departure_scale, pinned in
test_oracle_departure.py).

Obs of every element the step treats apart (reflectivity, Doppler velocity,
rain, TC vitals, a conventional element), reflectivities below the minimum,
negative and beyond 1e10, members near the echo threshold, operator qc set on
some members; k = 70 (more than one wave of obs per member row).  qc, errors,
perturbations and the departures of non-reflectivity obs must be
bit-identical; a reflectivity's dBZ value and departure within 1e-12 dBZ (the
device log10 may differ from the host's in the last bit).
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das
import letkf_amd

pytestmark = pytest.mark.gpu
NE = 70
REF, VR = oracle.ID_RADAR_REF, oracle.ID_RADAR_VR


def _obs(n=3000, seed=11):
    rng = np.random.default_rng(seed)
    elm = rng.choice([REF, REF, REF, VR, VR, oracle.ID_RAIN, oracle.ID_TCLON, oracle.ID_RADAR_PRH,
                      2819, 3073], n).astype(float)
    dat = rng.standard_normal(n) * 3.0
    err = rng.uniform(0.5, 3.0, n)
    ref = elm == REF
    nref = int(ref.sum())
    z = 10.0 ** rng.uniform(-1.0, 5.5, nref)
    z[rng.random(nref) < 0.05] = -1.0
    z[rng.random(nref) < 0.03] = 2.0e10
    dat[ref] = z
    hx = dat[:, None] + rng.standard_normal((n, NE)) * 2.0
    hx[ref] = rng.choice([0.0, 10.0, 20.0, 35.0], (nref, 1)) + rng.standard_normal((nref, NE)) * 8.0
    hx[ref] = np.maximum(hx[ref], 0.0)
    edge = np.flatnonzero(ref)[:40]
    hx[edge, :3] = 15.0 + 1e-6                     # exactly at the threshold: no echo
    qc = np.zeros((n, NE), dtype=np.int32)
    bad = rng.random(n) < 0.08
    qc[bad, rng.integers(0, NE, bad.sum())] = rng.choice([19, 20, 21, 22], bad.sum())
    return elm, dat, err, hx, qc


CONFIGS = [
    {},
    dict(use_obserr_radar_ref=1, use_obserr_radar_vr=1, min_radar_ref_member=5,
         min_radar_ref_member_obsref=20, radar_ref_thres_dbz=12.0, min_radar_ref_dbz=5.0,
         low_ref_shift=-3.0, gross_error=2.5, gross_error_radar_ref=1.5, gross_error_radar_vr=4.0,
         gross_error_rain=3.0, gross_error_tcx=0.8),
    dict(use_radar_vr=0, radar_ref_thres_dbz=3.0, min_radar_ref_dbz=8.0, gross_error=10.0),
]


@pytest.mark.parametrize("cfg", range(len(CONFIGS)))
def test_departure_matches_oracle(gpu, cfg):
    elm, dat, err, hx, qcm = _obs(seed=11 + cfg)
    h = das.RegionalLetkf(nx=4, ny=4, zlev=np.arange(3) * 300.0, dx=1000.0, dy=1000.0, ne=NE,
                          nv3d=1, sigma_h=3000.0, sigma_v=1000.0)
    print(f"departure config {cfg}: {elm.size} obs, k={NE}", flush=True)
    o = dict(CONFIGS[cfg])
    y, e, dep, hd, qc = h.obs_departure(elm, dat, err, hx, qcm, **o)
    ry, re, rdep, rhd, rqc = oracle.obs_departure_scale(elm, dat, err, hx, qcm,
                                                        **{k: bool(v) if k.startswith("use_") else v
                                                           for k, v in o.items()})
    h.close()
    np.testing.assert_array_equal(qc, rqc)
    np.testing.assert_array_equal(e, re)
    np.testing.assert_array_equal(hd, rhd)
    conv = (elm == REF) & (ry != oracle.UNDEF)
    np.testing.assert_array_equal(y[~conv], ry[~conv])
    np.testing.assert_array_equal(dep[~conv], rdep[~conv])
    assert np.max(np.abs(y[conv] - ry[conv])) <= 1e-12
    assert np.max(np.abs(dep[conv] - rdep[conv])) <= 1e-12
    # the cases the data were built to reach
    for code in ("good", "gross_err", "obs_bad") + (("ref_mem",) if cfg != 2 else ("otype",)):
        assert np.any(qc == oracle.IQC[code]), code
    assert np.any(qc == 21) and np.any(qc == 22)
    good = qc == 0
    np.testing.assert_allclose(hd[good].sum(axis=1), 0.0, atol=1e-9)


def test_departure_arguments_checked(gpu):
    h = das.RegionalLetkf(nx=4, ny=4, zlev=np.arange(3) * 300.0, dx=1000.0, dy=1000.0, ne=NE,
                          nv3d=1, sigma_h=3000.0, sigma_v=1000.0)
    hx = np.zeros((2, NE))
    with pytest.raises(letkf_amd.LetkfError):
        h.obs_departure([oracle.ID_H08IR, REF], [250.0, 10.0], [1.0, 1.0], hx)
    with pytest.raises(ValueError):
        h.obs_departure([VR, REF], [1.0, 10.0], [1.0, 1.0], np.zeros((2, NE - 1)))
    with pytest.raises(TypeError):
        h.obs_departure([VR, REF], [1.0, 10.0], [1.0, 1.0], hx, gross=3.0)
    y, e, dep, hd, qc = h.obs_departure([VR], [1.0], [1.0], np.zeros((1, NE)))
    assert qc[0] == 0 and dep[0] == 1.0
    h.close()
