"""SPEEDY file formats of the reference's LETKF cycle.

* Grid files (``read_grd4`` / ``write_grd4``,
  speedy/common/common_speedy.f90:160-287): unformatted direct access, one
  record per (variable, level) of nlon x nlat float32 with longitude fastest;
  records run over the 3-D variables (u, v, T, q) level by level, then the
  2-D variables (ps, rain): nlon=96, nlat=48, nlev=7, nv3d=4, nv2d=2
  (speedy/common/common_speedy.f90:18-22).  The reference build writes them big-endian
  (``-byteswapio``, speedy/letkf/make_letkf.sh:6).
* Observation files (``read_obs`` / ``read_obs2``,
  speedy/common/common_obs_speedy.f90:400-602): unformatted sequential, one
  record of 6 (obs) or 8 (obs + H(x) + qc, written by obsope) float32 per
  observation, each record framed by 4-byte length markers.  On reading,
  pressure levels of u, v, T, q and rh are converted hPa -> Pa and ps
  values and errors hPa -> Pa, rh values and errors percent -> fraction, in
  single precision as the reference does.

* GrADS descriptors (``write_ctl``) of the grid files, laid out as
  speedy/common/yyyymmddhh.ctl.

All arrays are numpy; nothing here touches the GPU.
"""
from __future__ import annotations

import os

import numpy as np

ID_U, ID_V, ID_T, ID_Q, ID_RH, ID_PS = 2819, 2820, 3073, 3330, 3331, 14593

NLON, NLAT, NLEV, NV3D, NV2D = 96, 48, 7, 4, 2

OBS_KEYS = ("elem", "rlon", "rlat", "rlev", "odat", "oerr")


def _f4(big_endian):
    return np.dtype(">f4" if big_endian else "<f4")


def _i4(big_endian):
    return np.dtype(">i4" if big_endian else "<i4")


def read_grd(path, nlon=NLON, nlat=NLAT, nlev=NLEV, nv3d=NV3D, nv2d=NV2D, big_endian=True):
    """(v3d (nlon, nlat, nlev, nv3d), v2d (nlon, nlat, nv2d)) as float32, Fortran order."""
    nrec = nv3d * nlev + nv2d
    a = np.fromfile(path, dtype=_f4(big_endian), count=nrec * nlon * nlat)
    if a.size != nrec * nlon * nlat:
        raise ValueError(f"{path}: {a.size} values, expected {nrec * nlon * nlat}")
    recs = a.astype(np.float32).reshape(nrec, nlat, nlon)       # record = (j, i) with i fastest
    v3d = np.asfortranarray(recs[:nv3d * nlev].reshape(nv3d, nlev, nlat, nlon).transpose(3, 2, 1, 0))
    v2d = np.asfortranarray(recs[nv3d * nlev:].reshape(nv2d, nlat, nlon).transpose(2, 1, 0))
    return v3d, v2d


def write_grd(path, v3d, v2d, big_endian=True):
    """Write v3d (nlon, nlat, nlev, nv3d) and v2d (nlon, nlat, nv2d), rounded to float32
    (REAL(v, r_sngl) in write_grd)."""
    v3d = np.asarray(v3d)
    v2d = np.asarray(v2d)
    nlon, nlat, nlev, nv3d = v3d.shape
    recs3 = v3d.astype(np.float32).transpose(3, 2, 1, 0).reshape(nv3d * nlev, nlat, nlon)
    recs2 = v2d.astype(np.float32).transpose(2, 1, 0).reshape(v2d.shape[2], nlat, nlon)
    np.concatenate([recs3, recs2]).astype(_f4(big_endian)).tofile(path)


def _nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return np.trunc(x + np.copysign(np.float32(0.5), x)).astype(np.int64)


def decode_obs(wk):
    """Records (n, 6|8) float32 in file units -> dict of float64 arrays in the
    analysis units, with read_obs / read_obs2's single-precision conversions
    (speedy/common/common_obs_speedy.f90:470-500)."""
    wk = np.array(wk, dtype=np.float32)
    nval = wk.shape[1]
    elem = _nint(wk[:, 0])
    h2p, pct = np.float32(100.0), np.float32(0.01)
    lev_hpa = np.isin(elem, [ID_U, ID_V, ID_T, ID_Q, ID_RH])
    wk[lev_hpa, 3] = wk[lev_hpa, 3] * h2p
    ps = elem == ID_PS
    wk[ps, 4] = wk[ps, 4] * h2p
    wk[ps, 5] = wk[ps, 5] * h2p
    rh = elem == ID_RH
    wk[rh, 4] = wk[rh, 4] * pct
    wk[rh, 5] = wk[rh, 5] * pct
    out = {k: wk[:, i].astype(np.float64) for i, k in enumerate(OBS_KEYS)}
    if nval >= 8:
        out["ohx"] = wk[:, 6].astype(np.float64)
        out["oqc"] = _nint(wk[:, 7]).astype(np.int32)
    return out


def encode_obs2(obs):
    """dict (elem, rlon, rlat, rlev, odat, oerr, ohx, oqc) in analysis units ->
    (n, 8) float32 records in file units, as write_obs2 forms them
    (speedy/common/common_obs_speedy.f90:569-598): each value rounded to single
    precision first, then scaled in single precision."""
    n = np.asarray(obs["elem"]).size
    wk = np.empty((n, 8), dtype=np.float32)
    for i, k in enumerate(OBS_KEYS + ("ohx",)):
        wk[:, i] = np.asarray(obs[k], dtype=np.float64).astype(np.float32)
    wk[:, 7] = np.asarray(obs["oqc"]).astype(np.float32)
    elem = _nint(wk[:, 0])
    p2h, pct = np.float32(0.01), np.float32(100.0)
    lev_hpa = np.isin(elem, [ID_U, ID_V, ID_T, ID_Q, ID_RH])
    wk[lev_hpa, 3] = wk[lev_hpa, 3] * p2h
    ps = elem == ID_PS
    wk[ps, 4] = wk[ps, 4] * p2h
    wk[ps, 5] = wk[ps, 5] * p2h
    rh = elem == ID_RH
    wk[rh, 4] = wk[rh, 4] * pct
    wk[rh, 5] = wk[rh, 5] * pct
    return wk


def read_records(path, nval, big_endian=True):
    """The raw (n, nval) float32 records of an unformatted sequential obs file."""
    raw = np.fromfile(path, dtype=np.uint8)
    f4, i4 = _f4(big_endian), _i4(big_endian)
    rec = 4 + 4 * nval + 4
    if raw.size % rec:
        raise ValueError(f"{path}: {raw.size} bytes is not a whole number of {nval}-value records")
    n = raw.size // rec
    r = raw.reshape(n, rec)
    lead = r[:, :4].copy().view(i4).ravel()
    tail = r[:, -4:].copy().view(i4).ravel()
    if np.any(lead != 4 * nval) or np.any(tail != 4 * nval):
        raise ValueError(f"{path}: record markers are not {4 * nval}")
    return r[:, 4:-4].copy().view(f4).astype(np.float32).reshape(n, nval)


def read_obs(path, nval=8, big_endian=True):
    """read_obs (nval 6) / read_obs2 (nval 8): dict of float64 arrays (elem,
    rlon, rlat, rlev, odat, oerr[, ohx, oqc]) in the analysis units."""
    return decode_obs(read_records(path, nval, big_endian))


def write_records(path, records, big_endian=True):
    """Write (n, 6|8) float32 records as an unformatted sequential file."""
    wk = np.asarray(records, dtype=np.float32)
    n, nval = wk.shape
    f4, i4 = _f4(big_endian), _i4(big_endian)
    mark = np.full((n, 1), 4 * nval, dtype=i4).view(np.uint8).reshape(n, 4)
    body = wk.astype(f4).view(np.uint8).reshape(n, 4 * nval)
    np.concatenate([mark, body, mark], axis=1).tofile(path)


def write_obs2(path, obs, big_endian=True):
    """write_obs2: obs dict in analysis units -> 8-value records."""
    write_records(path, encode_obs2(obs), big_endian)


def count_obs(path, nval=8):
    """get_nobs: records in the file (0 when it does not exist)."""
    if not os.path.exists(path):
        return 0
    return os.path.getsize(path) // (4 * nval + 8)


SPEEDY_VARS = (("U", 7, "U-wind [m/s]"), ("V", 7, "V-wind [m/s]"), ("T", 7, "Temperature [K]"),
               ("Q", 7, "Specific Humidity [kg/kg]"), ("PS", 0, "Surface Pressure [Pa]"),
               ("RAIN", 0, "Precipitation [mm/6hr]"))


def write_ctl(path, lat, sig, dset="^%y4%m2%d2%h2.grd", title="SPEEDY MODEL OUTPUT",
              nlon=NLON, tdef=(10000, "00Z01JAN1982", "6HR"), variables=SPEEDY_VARS,
              big_endian=True):
    """GrADS descriptor of the grid files (the layout of speedy/common/yyyymmddhh.ctl):
    longitudes 0, 360/nlon, ..., the given latitude table and sigma levels, one
    record per (variable, level) in the order read_grd / write_grd use."""
    opts = "template big_endian" if big_endian else "template"
    lines = [f"DSET {dset}", f"TITLE {title}", "UNDEF -9.99E33", f"OPTIONS {opts}",
             f"XDEF {nlon} LINEAR 0.0 {360.0 / nlon:.2f}",
             f"YDEF {len(lat)} LEVELS " + " ".join(f"{v:7.3f}" for v in lat),
             f"ZDEF {len(sig)} LEVELS " + " ".join(f"{v:5.3f}" for v in sig),
             f"TDEF {tdef[0]} LINEAR {tdef[1]} {tdef[2]}", f"VARS {len(variables)}"]
    lines += [f"{name} {nlev} 99 {desc}" for name, nlev, desc in variables]
    lines.append("ENDVARS")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
