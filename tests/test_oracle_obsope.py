"""Pins of the oracle's SPEEDY obs operator (phys2ijk, itpl_2d/3d, obsope rules).

The restatement (oracle/lekf_oracle.c: orc_phys2ij, orc_itpl_*, orc_trans_xtoy)
follows speedy/common/common_obs_speedy.f90:30-282 and
speedy/letkf/obsope.f90:56-89.  Pinned here against the reference's own
phys2ijk and Trans_XtoY (built in oracle/_ref, with obsope's range rules in
the binding) bit for bit, on the SPEEDY-sized obs set with its edge rows, and
by what the mathematics and the reference's rules fix: bi-/trilinear
interpolation reproduces fields linear in the grid indices; the longitude wrap
uses column 1; latitudes on the table map to integer rj; the vertical and type
rejection rules.  The oracle covers u, v, T, q and ps; RH and rain obs (which
Trans_XtoY also maps, via calc_rh and the rain field) are rejected by the
oracle and the GPU operator alike.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import synth

LAT = synth.SPEEDY_LAT
SIG = synth.SPEEDY_SIG
NLON, NLAT, NLEV = 96, LAT.size, SIG.size


def _fields(u_fn, t_fn, ps=1.0e5):
    i = np.arange(1, NLON + 1)[None, None, :]
    j = np.arange(1, NLAT + 1)[None, :, None]
    k = np.arange(1, NLEV + 1)[:, None, None]
    v3d = np.zeros((4, NLEV, NLAT, NLON))
    v3d[0] = u_fn(i, j, k)
    v3d[1] = u_fn(i, j, k)
    v3d[2] = t_fn(i, j, k)
    v3d[3] = 1.0e-3
    return v3d, np.full((NLAT, NLON), ps)


def test_phys2ij_longitude_cases():
    assert oracle.phys2ij(0.0, 10.0, NLON, LAT)[1] == NLON + 1
    assert oracle.phys2ij(360.0, 10.0, NLON, LAT)[1] == NLON + 1
    assert oracle.phys2ij(180.0, 10.0, NLON, LAT)[1] == NLON / 2 + 1
    assert oracle.phys2ij(-1.0, 10.0, NLON, LAT)[0] == 1
    assert oracle.phys2ij(361.0, 10.0, NLON, LAT)[0] == 1


def test_phys2ij_latitude_table_and_poles():
    for jj in range(1, NLAT):
        st, _, rj = oracle.phys2ij(100.0, LAT[jj], NLON, LAT)
        assert st == 0 and rj == jj + 1
    mid = 0.5 * (LAT[3] + LAT[4])
    assert abs(oracle.phys2ij(100.0, mid, NLON, LAT)[2] - 4.5) < 1e-14
    assert oracle.phys2ij(100.0, LAT[0] - 0.1, NLON, LAT)[0] == 1   # poleward of the table
    assert oracle.phys2ij(100.0, LAT[-1] + 0.1, NLON, LAT)[0] == 1


def test_interpolation_reproduces_linear_fields():
    rng = np.random.default_rng(3)
    v3d, ps = _fields(lambda i, j, k: 2.0 + 0.5 * i - 0.25 * j + 0 * k,
                      lambda i, j, k: 300.0 - 7.0 * k + 0 * i + 0 * j)
    for _ in range(200):
        lon = rng.uniform(5.0, 350.0)
        lat = rng.uniform(LAT[1], LAT[-2])
        lev = 1.0e5 * rng.uniform(SIG[-1] + 0.01, SIG[0] - 0.01)
        st, y, ri, rj, rk = oracle.trans_xtoy(oracle.ID_U, lon, lat, lev, 0.0, LAT, SIG, v3d, ps)
        assert st == 0
        assert abs(y - (2.0 + 0.5 * ri - 0.25 * rj)) < 1e-12
        st, y, ri, rj, rk = oracle.trans_xtoy(oracle.ID_T, lon, lat, lev, 0.0, LAT, SIG, v3d, ps)
        assert st == 0
        assert abs(y - (300.0 - 7.0 * rk)) < 1e-10
        # rk is linear in ln p between the levels
        k = int(np.ceil(rk))
        lp = np.log(1.0e5 * SIG)
        want = (k - 1) + (np.log(lev) - lp[k - 2]) / (lp[k - 1] - lp[k - 2])
        assert abs(rk - want) < 1e-12


def test_longitude_wrap_uses_column_one():
    v3d, ps = _fields(lambda i, j, k: i + 0.0 * j + 0.0 * k, lambda i, j, k: 0 * i + 0 * j + 0 * k)
    lon = 360.0 - 0.25 * 360.0 / NLON                    # ri = nlon + 0.75
    st, y, ri, rj, rk = oracle.trans_xtoy(oracle.ID_U, lon, 20.0, 5.0e4, 0.0, LAT, SIG, v3d, ps)
    assert st == 0 and abs(ri - (NLON + 0.75)) < 1e-12
    assert abs(y - (0.25 * NLON + 0.75 * 1.0)) < 1e-10


def test_vertical_and_type_rules():
    v3d, ps = _fields(lambda i, j, k: 1.0 + 0 * i + 0 * j + 0 * k,
                      lambda i, j, k: 250.0 + 0 * i + 0 * j + 0 * k)
    below = 1.0e5 * (SIG[0] + 0.01)   # under the lowest level
    above = 1.0e5 * SIG[-1] * 0.5     # over the top level
    st, y, _, _, rk = oracle.trans_xtoy(oracle.ID_U, 30.0, 10.0, below, 0.0, LAT, SIG, v3d, ps)
    assert st == 0 and rk == 1.00001 and abs(y - 1.0) < 1e-14   # winds: lowest level
    assert oracle.trans_xtoy(oracle.ID_T, 30.0, 10.0, below, 0.0, LAT, SIG, v3d, ps)[0] == 1
    assert oracle.trans_xtoy(oracle.ID_U, 30.0, 10.0, above, 0.0, LAT, SIG, v3d, ps)[0] == 1
    st, y, _, _, rk = oracle.trans_xtoy(oracle.ID_PS, 30.0, 10.0, 0.0, 1.0e5, LAT, SIG, v3d, ps)
    assert st == 0 and rk == 0.0 and y == 1.0e5
    assert oracle.trans_xtoy(oracle.ID_PS, 30.0, 10.0, 0.0, -200.0, LAT, SIG, v3d, ps)[0] == 1
    assert oracle.trans_xtoy(oracle.ID_RH, 30.0, 10.0, 5.0e4, 0.0, LAT, SIG, v3d, ps)[0] == 1


def _speedy_obs_with_edges(p, n=600, seed=1):
    o = p["obs"]
    rng = np.random.default_rng(seed)
    sel = rng.choice(o["elm"].size, n, replace=False)
    rows = [tuple(float(np.asarray(o[k])[i]) for k in ("elm", "lon", "lat", "lev", "dat"))
            for i in sel]
    lat0, latn = p["lat"][0], p["lat"][-1]
    rows += [(oracle.ID_U, 0.0, 10.0, 5.0e4, 1.0), (oracle.ID_V, 360.0, -20.0, 3.0e4, 1.0),
             (oracle.ID_T, 359.5, 30.0, 7.0e4, 1.0), (oracle.ID_T, 100.0, lat0 - 0.5, 5.0e4, 1.0),
             (oracle.ID_Q, 100.0, latn + 0.5, 5.0e4, 1.0), (oracle.ID_U, 50.0, 0.0, 1.2e5, 1.0),
             (oracle.ID_T, 50.0, 0.0, 1.2e5, 1.0), (oracle.ID_T, 50.0, 0.0, 1.0e3, 1.0),
             (oracle.ID_PS, 50.0, 0.0, 0.0, -500.0), (oracle.ID_PS, 3.0, 40.0, 0.0, 1.0e5),
             (oracle.ID_U, -2.0, 0.0, 5.0e4, 1.0), (oracle.ID_V, 361.0, 5.0, 5.0e4, 1.0),
             (oracle.ID_Q, 180.0, LAT[5], 1.0e5 * SIG[3], 1.0), (oracle.ID_PS, 0.0, LAT[0], 0.0, 9.0e4)]
    return rows


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_oracle_equals_reference_obs_operator():
    """orc_trans_xtoy against the reference's phys2ijk + obsope rules + Trans_XtoY
    on three members of the SPEEDY-sized problem: the same obs rejected, and
    ri, rj, rk and H(x) bit for bit."""
    p = synth.speedy()
    rows = _speedy_obs_with_edges(p)
    nlon, nlat, nlev = p["lon"].size, p["lat"].size, p["sig"].size
    assert np.array_equal(p["lat"], LAT) and np.array_equal(p["sig"], SIG)
    used = 0
    for m in (0, 7, 19):
        v3d = np.ascontiguousarray(p["gues3d"][:, :, m, :].transpose(2, 1, 0)).reshape(4, nlev, nlat,
                                                                                      nlon)
        ps = p["gues2d"][:, m, 0].reshape(nlat, nlon)
        v2d = np.stack([ps, np.zeros_like(ps)])
        for r in rows:
            a = oracle.trans_xtoy(*r, LAT, SIG, v3d, ps)
            b = oracle.ref_obsope1(*r, v3d, v2d)
            assert (a[0] == 0) == (b[0] == 0), r
            if b[0] == 0:
                used += 1
                assert a[1:] == b[1:], (r, a, b)
    assert used > 1500

