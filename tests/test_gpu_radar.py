"""GPU parity of SCALE's radar obs operator (lekf_obsope_radar) on a regional
handle's resident ensemble, against the oracle's restatement
(orc_radar_xtoy: obsope's checks, phys2ijkz, itpl_3d, Trans_XtoY_radar and
calc_ref_vr; pinned in test_oracle_radar.py).

Grid 24 x 20 x 12 with terrain-following heights (some lowest levels
invalid), k = 70 members (more than one wave of lanes), the 11 fields stored
in a permuted order among 12 state variables.  Obs of both radar elements
and one other element, inside and outside the grid, above the top, below
the lowest valid level, above RADAR_ZMAX and at a radar's position.
qc and rk must be bit-identical; H(x) within 1e-9 (dBZ or m/s), 99% of it
within 1e-10: the device pow, log10, acos and trigonometric functions may
differ from the host's in the last bit, and two steps of the reference's
formulas amplify that (see the test).
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das

pytestmark = pytest.mark.gpu
NX, NY, NLEV, NE = 24, 20, 12, 70
VAR = (5, 0, 11, 3, 7, 1, 9, 2, 10, 4, 8)  # state variable of u, v, w, T, p, q, qc, qr, qi, qs, qg


def _problem(seed=7):
    rng = np.random.default_rng(seed)
    nij = NX * NY
    terrain = 200.0 * rng.random(nij)
    hgt = terrain[:, None] + 300.0 * np.arange(NLEV)[None, :] + 20.0 * rng.random((nij, NLEV))
    hgt[rng.choice(nij, 25, replace=False), 0] = -999.0      # invalid lowest levels
    hgt = np.asfortranarray(hgt)
    g3 = np.zeros((nij, NLEV, NE, 12), order="F")
    z = hgt[:, :, None]
    fields = dict(u=8.0 * rng.standard_normal((nij, NLEV, NE)),
                  v=8.0 * rng.standard_normal((nij, NLEV, NE)),
                  w=rng.standard_normal((nij, NLEV, NE)),
                  t=285.0 - 0.0065 * np.maximum(z, 0) + rng.standard_normal((nij, NLEV, NE)),
                  p=1.0e5 * np.exp(-np.maximum(z, 0) / 8000.0) * (1 + 0.001 * rng.standard_normal((nij, NLEV, NE))),
                  q=5e-3 * rng.random((nij, NLEV, NE)))
    for name in ("qc", "qr", "qi", "qs", "qg"):
        a = 1e-3 * rng.lognormal(0.0, 1.0, (nij, NLEV, NE))
        a[rng.random((nij, NLEV, NE)) < 0.3] = 0.0
        fields[name] = a
    for q, name in enumerate(oracle.RADAR_FIELDS):
        g3[..., VAR[q]] = fields[name]
    g3[..., 6] = rng.standard_normal((nij, NLEV, NE))       # a state variable the operator ignores
    n = 400
    ri = rng.uniform(0.8, NX + 0.3, n)
    rj = rng.uniform(0.8, NY + 0.3, n)
    zo = rng.uniform(-100.0, 4500.0, n)
    elm = rng.choice([oracle.ID_RADAR_REF, oracle.ID_RADAR_VR], n).astype(float)
    elm[:6] = 2819.0
    lon = 139.0 + 0.01 * ri
    lat = 35.0 + 0.009 * rj
    rlon = np.where(np.arange(n) % 2 == 0, 139.05, 139.2)
    rlat = np.where(np.arange(n) % 2 == 0, 35.05, 35.15)
    rz = np.full(n, 80.0)
    lon[10], lat[10], rlon[10], rlat[10] = 139.1, 35.1, 139.1, 35.1       # at the radar
    ri[10], rj[10] = 10.0, 11.1
    zo[11:14] = [4100.0, 4200.0, 3950.0]                                   # above zmax
    obs = dict(elm=elm, lon=lon, lat=lat, z=zo, ri=ri, rj=rj)
    radar = dict(lon=rlon, lat=rlat, z=rz)
    xo = np.asfortranarray(np.stack([g3[..., VAR[q]] for q in range(11)], axis=3))
    return hgt, g3, xo, obs, radar


@pytest.fixture(scope="module")
def radar_case():
    return _problem()


@pytest.mark.parametrize("method,use_wt", [(3, False), (3, True), (2, True), (1, True)])
def test_radar_operator_matches_oracle(gpu, radar_case, method, use_wt):
    hgt, g3, xo, obs, radar = radar_case
    h = das.RegionalLetkf(nx=NX, ny=NY, zlev=np.arange(NLEV) * 300.0, dx=1000.0, dy=1000.0,
                          ne=NE, nv3d=12, sigma_h=3000.0, sigma_v=1000.0)
    h.set_ensemble(g3)
    h.set_heights(hgt)
    print(f"radar operator method {method} wt {use_wt}: {obs['elm'].size} obs, k={NE}", flush=True)
    hx, qc, rk = h.radar_operator(obs["elm"], obs["lon"], obs["lat"], obs["z"], obs["ri"],
                                  obs["rj"], radar["lon"], radar["lat"], radar["z"], var=VAR,
                                  method=method, use_terminal_velocity=use_wt, zmax=4000.0)
    ref_hx, ref_qc, ref_rk = oracle.radar_xtoy(obs, radar, xo, hgt, NX, NY, method=method,
                                               use_wt=use_wt, zmax=4000.0)
    np.testing.assert_array_equal(qc, ref_qc)
    np.testing.assert_array_equal(rk, ref_rk)
    good = ref_qc == 0
    assert good.sum() > 100 * NE
    for code in ("otype", "radar_vhi", "out_h", "out_vhi", "out_vlo"):
        assert np.any(ref_qc == oracle.IQC[code]), code
    np.testing.assert_array_equal(hx[~good], ref_hx[~good])
    # The reference's formulas amplify last-bit differences of the device
    # functions in two places: com_distll_1 takes acos of a cosine within
    # ~1e-6 of 1 at radar ranges (the elevation angle, hence the radial
    # velocity), and method 3 forms the rain share 1 - Fs - Fg from cube roots,
    # which cancels when rain, snow and graupel are nearly equal.  Nearly all
    # values agree to 1e-10 and every one to 1e-9 (dBZ or m/s).
    d = np.abs(hx - ref_hx)[good]
    assert d.max() < 1e-9
    assert np.mean(d < 1e-10) > 0.99
    # both elements took real values
    assert np.any(good & (obs["elm"][:, None] == oracle.ID_RADAR_VR))
    assert np.any(good & (obs["elm"][:, None] == oracle.ID_RADAR_REF) & (ref_hx > 0.0))


def test_radar_operator_arguments_checked(gpu, radar_case):
    hgt, g3, _, obs, radar = radar_case
    import letkf_amd
    h = das.RegionalLetkf(nx=NX, ny=NY, zlev=np.arange(NLEV) * 300.0, dx=1000.0, dy=1000.0,
                          ne=NE, nv3d=12, sigma_h=3000.0, sigma_v=1000.0)
    h.set_ensemble(g3)
    args = [obs[k][:5] for k in ("elm", "lon", "lat", "z", "ri", "rj")] + \
        [radar[k][:5] for k in ("lon", "lat", "z")]
    with pytest.raises(letkf_amd.LetkfError):
        h.radar_operator(*args, var=VAR)                  # no heights yet
    h.set_heights(hgt)
    with pytest.raises(letkf_amd.LetkfError):
        h.radar_operator(*args, var=VAR, method=4)
    with pytest.raises(letkf_amd.LetkfError):
        h.radar_operator(*args, var=(12,) + VAR[1:])      # no such state variable
    hx, qc, rk = h.radar_operator(*args, var=VAR)
    assert hx.shape == (5, NE) and qc.shape == (5, NE) and rk.shape == (5,)


def test_resident_radar_obs_equal_host_chain(gpu, radar_case):
    """set_obs_radar (operator, departure/QC and obs index on the device) gives the
    same qc, local lists and analysis, bit for bit, as the host-staged chain
    radar_operator -> obs_departure -> set_obs."""
    hgt, g3, _, obs, radar = radar_case
    rng = np.random.default_rng(5)
    elm = obs["elm"]
    n = elm.size
    ref, vr = elm == oracle.ID_RADAR_REF, elm == oracle.ID_RADAR_VR
    dat = np.where(ref, 10.0 ** (rng.uniform(-5.0, 45.0, n) / 10.0), rng.normal(0.0, 8.0, n))
    err = np.where(ref, 5.0, np.where(vr, 3.0, 1.0))
    bucket = np.where(ref, 0, np.where(vr, 1, 2)).astype(np.int32)
    kw = dict(var=VAR, method=3, zmax=4000.0)
    qkw = dict(min_radar_ref_member=3, gross_error=4.0)

    def handle():
        h = das.RegionalLetkf(nx=NX, ny=NY, zlev=np.arange(NLEV) * 300.0, dx=1000.0, dy=1000.0,
                              ne=NE, nv3d=12, sigma_h=3000.0, sigma_v=1000.0, nobs_max=50)
        h.set_ensemble(g3)
        h.set_heights(hgt)
        return h

    print(f"resident radar obs: {n} obs, k={NE}", flush=True)
    h1 = handle()
    hx, qcm, _ = h1.radar_operator(elm, obs["lon"], obs["lat"], obs["z"], obs["ri"], obs["rj"],
                                   radar["lon"], radar["lat"], radar["z"], **kw)
    _, e, dep, hd, qc = h1.obs_departure(elm, dat, err, hx, qcm, **qkw)
    h1.set_obs(obs["ri"], obs["rj"], obs["z"], e, dep, bucket, hd, (qc == 0).astype(np.int32))
    h1.run()
    a1, _ = h1.analysis()
    n1 = h1.nobsl()
    h1.close()
    h2 = handle()
    qc2 = h2.set_obs_radar(elm, obs["lon"], obs["lat"], obs["z"], obs["ri"], obs["rj"],
                           radar["lon"], radar["lat"], radar["z"], dat, err, bucket, **kw, **qkw)
    h2.run()
    a2, _ = h2.analysis()
    n2 = h2.nobsl()
    h2.close()
    np.testing.assert_array_equal(qc2, qc)
    assert (qc == 0).sum() > 50 and np.any(qc == oracle.IQC["gross_err"])
    np.testing.assert_array_equal(n2, n1)
    assert n1.max() > 0
    np.testing.assert_array_equal(a2, a1)
    assert not np.array_equal(a1, g3)                     # the obs did change the ensemble
