"""SPEEDY file formats (letkf_amd.speedy_io) against the reference's own file
routines, compiled from speedy/common/common_speedy.f90 and
common_obs_speedy.f90 into oracle/_ref:

* grid files: our writer and write_grd produce the same bytes, and each side
  reads the other's file to the same float64 values (float32-rounded);
* obs files: write_obs2 and ours produce the same bytes from analysis-unit
  records of every element type (the hPa/percent conversions in single
  precision), read_obs2 / read_obs and ours read them to the same values,
  get_nobs and ours count the same records;
* byte order: the reference build writes big-endian (-byteswapio), which is
  the native-endian file with every 4-byte word reversed, record markers
  included; checked on known bytes.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import speedy_io as sio

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
ELEMS = (sio.ID_U, sio.ID_V, sio.ID_T, sio.ID_Q, sio.ID_RH, sio.ID_PS)


def _fields(seed=0):
    rng = np.random.default_rng(seed)
    v3 = np.asfortranarray(rng.standard_normal((96, 48, 7, 4)) * [10.0, 10.0, 30.0, 1e-3] +
                           [0.0, 0.0, 250.0, 5e-3])
    v2 = np.asfortranarray(np.stack([1.0e5 + 800.0 * rng.standard_normal((96, 48)),
                                     np.abs(rng.standard_normal((96, 48)))], axis=2))
    return v3, v2


def _obs(n=600, seed=1):
    """Analysis-unit records of every element (Pa levels, Pa ps, fractional rh)."""
    rng = np.random.default_rng(seed)
    elem = rng.choice(ELEMS, n).astype(np.float64)
    e = elem.astype(int)
    lev = rng.uniform(1000.0, 101000.0, n)
    dat = rng.standard_normal(n) * 10.0
    err = rng.uniform(0.1, 3.0, n)
    ps = e == sio.ID_PS
    dat[ps] = 1.0e5 + 500.0 * rng.standard_normal(ps.sum())
    err[ps] = 100.0
    lev[ps] = dat[ps]
    rh = e == sio.ID_RH
    dat[rh] = rng.uniform(0.0, 1.0, rh.sum())
    err[rh] = 0.1
    return dict(elem=elem, rlon=rng.uniform(0.0, 360.0, n), rlat=rng.uniform(-90.0, 90.0, n),
                rlev=lev, odat=dat, oerr=err, ohx=dat + rng.standard_normal(n),
                oqc=rng.integers(0, 2, n).astype(np.int32))


def _swap_words(b):
    return np.frombuffer(b, dtype=np.uint8).reshape(-1, 4)[:, ::-1].tobytes()


@needs_ref
def test_reference_grid_is_speedy_t30():
    assert oracle.ref_speedy_grid() == (sio.NLON, sio.NLAT, sio.NLEV, sio.NV3D, sio.NV2D)


@needs_ref
def test_grid_files_match_reference(tmp_path):
    v3, v2 = _fields()
    ours, ref = tmp_path / "ours.grd", tmp_path / "ref.grd"
    sio.write_grd(ours, v3, v2, big_endian=False)
    oracle.ref_write_grd(str(ref), v3, v2)
    assert ours.read_bytes() == ref.read_bytes()
    r3, r2 = oracle.ref_read_grd(str(ours))
    np.testing.assert_array_equal(r3, v3.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(r2, v2.astype(np.float32).astype(np.float64))
    o3, o2 = sio.read_grd(ref, big_endian=False)
    np.testing.assert_array_equal(o3, r3)
    np.testing.assert_array_equal(o2, r2)
    assert o3.shape == (96, 48, 7, 4) and o2.shape == (96, 48, 2)


@needs_ref
def test_obs2_files_match_reference(tmp_path):
    obs = _obs()
    ours, ref = tmp_path / "ours.dat", tmp_path / "ref.dat"
    sio.write_obs2(ours, obs, big_endian=False)
    oracle.ref_write_obs2(str(ref), obs)
    assert ours.read_bytes() == ref.read_bytes()
    assert oracle.ref_get_nobs(str(ref), 8) == sio.count_obs(ref, 8) == obs["elem"].size
    r = oracle.ref_read_obs2(str(ref), obs["elem"].size)
    o = sio.read_obs(ref, nval=8, big_endian=False)
    for k in sio.OBS_KEYS + ("ohx", "oqc"):
        np.testing.assert_array_equal(o[k], r[k], err_msg=k)
    # the file round trip changes values only by the single-precision steps
    np.testing.assert_allclose(o["rlev"], obs["rlev"], rtol=1e-6)
    np.testing.assert_array_equal(o["oqc"], obs["oqc"])


@needs_ref
def test_obs6_read_matches_reference(tmp_path):
    rng = np.random.default_rng(2)
    n = 300
    wk = np.stack([rng.choice(ELEMS, n), rng.uniform(0, 360, n), rng.uniform(-90, 90, n),
                   rng.uniform(10, 1010, n), rng.uniform(0, 100, n), rng.uniform(0.5, 5, n)],
                  axis=1).astype(np.float32)
    path = tmp_path / "obsin.dat"
    sio.write_records(path, wk, big_endian=False)
    assert oracle.ref_get_nobs(str(path), 6) == sio.count_obs(path, 6) == n
    r = oracle.ref_read_obs(str(path), n)
    o = sio.read_obs(path, nval=6, big_endian=False)
    for k in sio.OBS_KEYS:
        np.testing.assert_array_equal(o[k], r[k], err_msg=k)
    assert "ohx" not in o


def test_big_endian_is_word_swap(tmp_path):
    v3, v2 = _fields(3)
    le, be = tmp_path / "le.grd", tmp_path / "be.grd"
    sio.write_grd(le, v3, v2, big_endian=False)
    sio.write_grd(be, v3, v2, big_endian=True)
    assert be.read_bytes() == _swap_words(le.read_bytes())
    b3, b2 = sio.read_grd(be)
    np.testing.assert_array_equal(b3, v3.astype(np.float32))
    obs = _obs(50, seed=4)
    sio.write_obs2(le.with_suffix(".dat"), obs, big_endian=False)
    sio.write_obs2(be.with_suffix(".dat"), obs, big_endian=True)
    assert be.with_suffix(".dat").read_bytes() == _swap_words(le.with_suffix(".dat").read_bytes())


def test_known_bytes_and_units(tmp_path):
    """One u obs at 500 hPa, one ps obs of 1000 hPa +- 1 hPa, one rh obs of 50 %:
    big-endian bytes of the first record and the units after reading."""
    wk = np.array([[sio.ID_U, 10.0, 20.0, 500.0, 1.5, 1.0],
                   [sio.ID_PS, 30.0, -40.0, 1000.0, 1000.0, 1.0],
                   [sio.ID_RH, 50.0, 60.0, 850.0, 50.0, 10.0]], dtype=np.float32)
    path = tmp_path / "obs.dat"
    sio.write_records(path, wk)
    raw = path.read_bytes()
    assert len(raw) == 3 * 32
    assert raw[:4] == b"\x00\x00\x00\x18"                 # record length 24
    assert raw[8:12] == b"\x41\x20\x00\x00"               # 10.0f
    assert raw[28:32] == b"\x00\x00\x00\x18"
    o = sio.read_obs(path, nval=6)
    np.testing.assert_array_equal(o["rlev"], [50000.0, 1000.0, 85000.0])   # ps level kept
    np.testing.assert_array_equal(o["odat"][:2], [1.5, 100000.0])
    assert o["odat"][2] == np.float64(np.float32(50.0) * np.float32(0.01))
    assert o["oerr"][1] == 100.0


def test_rejects_bad_files(tmp_path):
    path = tmp_path / "bad.dat"
    path.write_bytes(b"\x00" * 33)
    with pytest.raises(ValueError):
        sio.read_obs(path)
    sio.write_records(path, np.zeros((2, 6), np.float32))
    with pytest.raises(ValueError):
        sio.read_obs(path, nval=8)                       # 6-value records read as 8
    with pytest.raises(ValueError):
        sio.read_grd(path)
    assert sio.count_obs(tmp_path / "missing.dat") == 0


def test_grads_descriptor(tmp_path):
    """The .ctl describes the records write_grd lays down: 96 longitudes from 0
    by 3.75, the T30 latitude table, the 7 sigma levels, 4 x 7 + 2 records."""
    from letkf_amd import speedy_cycle as sc
    path = tmp_path / "yyyymmddhh.ctl"
    sio.write_ctl(path, sc.LAT, sc.SIG)
    lines = path.read_text().splitlines()
    kv = {ln.split()[0]: ln.split()[1:] for ln in lines}
    assert kv["OPTIONS"] == ["template", "big_endian"]
    assert kv["XDEF"] == ["96", "LINEAR", "0.0", "3.75"]
    lat = np.array(kv["YDEF"][2:], dtype=float)
    np.testing.assert_array_equal(lat, sc.LAT)
    np.testing.assert_array_equal(np.array(kv["ZDEF"][2:], dtype=float), sc.SIG)
    i = lines.index("VARS 6")
    rows = [ln.split()[:2] for ln in lines[i + 1:i + 7]]
    assert [r[0] for r in rows] == ["U", "V", "T", "Q", "PS", "RAIN"]
    nrec = sum(max(int(r[1]), 1) for r in rows)
    assert nrec == sio.NV3D * sio.NLEV + sio.NV2D
    assert lines[-1] == "ENDVARS"
