"""The oracle: CPU restatement of the reference LETKF local analysis.

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or the timed CPU baseline).  The product package ``letkf_amd`` never
imports it.

Two legs:

* ``impl="c"``   -- ``liblekf_oracle.so``, our C restatement (lekf_oracle.c),
  each function citing the reference file:line it follows.
* ``impl="ref"`` -- ``_ref/libletkf_ref.so``, the reference Fortran
  (common/common_letkf.f90, common/common_mtx.f90, common/netlib.f,
  common/netlibblas.f) compiled with amdflang from /root/reference by
  oracle/Makefile, reached through our BIND(C) shim ref_bind.f90.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_long

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_C_LIB = os.path.join(_HERE, "liblekf_oracle.so")
_REF_LIB = os.path.join(_HERE, "_ref", "libletkf_ref.so")

_c = None
_ref = None

_dp = POINTER(c_double)
_ip = POINTER(c_int)
_lp = POINTER(c_long)

# speedy/common/common_obs_speedy.f90:17-24
ID_U, ID_V, ID_T, ID_Q, ID_RH, ID_PS, ID_RAIN = 2819, 2820, 3073, 3330, 3331, 14593, 9999
PI_REF = 3.1415926535  # common/common.f90:28
RE_REF = 6371.3e3  # common/common.f90:35


def _lib_c():
    global _c
    if _c is None:
        if not os.path.exists(_C_LIB):
            raise RuntimeError(f"oracle library missing: {_C_LIB} (run make -C oracle)")
        _c = ctypes.CDLL(_C_LIB)
        _c.orc_letkf_core.restype = c_int
        _c.orc_letkf_core.argtypes = [c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _dp, c_int, _dp]
        _c.orc_mtx_eigen.restype = c_int
        _c.orc_mtx_eigen.argtypes = [c_int, _dp, _dp, _dp, _ip]
        _c.orc_l96_rk4.restype = None
        _c.orc_l96_rk4.argtypes = [c_int, c_int, c_double, c_double, _dp]
        _c.orc_relax_point.restype = None
        _c.orc_relax_point.argtypes = [c_int, c_int, _dp, _dp, _dp, _dp, c_int, c_double,
                                       c_double, c_double, c_int, _dp, _dp]
        _c.orc_distll_1.restype = c_double
        _c.orc_distll_1.argtypes = [c_double] * 4
        _c.orc_obs_local_l96.restype = c_int
        _c.orc_obs_local_l96.argtypes = [c_int, c_int, _ip, c_int, c_double, _ip, _dp]
        _c.orc_obs_bins_global.restype = c_int
        _c.orc_obs_bins_global.argtypes = [c_int, _dp, _dp, c_int, c_int, _dp, _dp, _ip, _ip]
        _c.orc_rect_candidates.restype = c_int
        _c.orc_rect_candidates.argtypes = [c_double, c_double, c_double, c_double, c_int, c_int,
                                           _dp, _dp, _ip, _ip]
        _c.orc_obs_local_global.restype = c_int
        _c.orc_obs_local_global.argtypes = (
            [c_double] * 5 + [c_int, c_int, c_int, _dp, _dp, _ip] + [_dp] * 6 +
            [c_double] * 4 + [_dp, _ip, _ip, _dp, _dp])
        _c.orc_obs_bins_regional.restype = c_int
        _c.orc_obs_bins_regional.argtypes = [c_int, _dp, _dp, _ip, c_int, c_int, _ip, _ip]
        _c.orc_obs_local_regional.restype = c_int
        _c.orc_obs_local_regional.argtypes = (
            [c_double] * 3 + [c_int, c_int, _ip] + [_dp] * 4 + [_ip, c_int] + [c_double] * 6 +
            [c_int, _ip, _ip, _dp, _dp, _ip, _ip, _dp, _dp])
        _c.orc_cap_stream.restype = c_int
        _c.orc_cap_stream.argtypes = [c_int, _dp, c_int, _ip, _ip]
        _c.orc_das_global.restype = c_int
        _c.orc_das_global.argtypes = (
            [c_int] * 5 + [_dp] * 4 + [_dp, _dp, _dp, c_int, _dp, _dp, c_int, c_int, _dp, _dp,
                                       c_int, _ip] + [_dp] * 8 + [c_double, c_double, _dp, c_int,
                                                                  c_int, _ip])
        _c.orc_phys2ij.restype = c_int
        _c.orc_phys2ij.argtypes = [c_double, c_double, c_int, c_int, _dp, _dp, _dp]
        _c.orc_itpl_2d.restype = c_double
        _c.orc_itpl_2d.argtypes = [_dp, c_int, c_double, c_double]
        _c.orc_itpl_3d.restype = c_double
        _c.orc_itpl_3d.argtypes = [_dp, c_int, c_int, c_double, c_double, c_double]
        _c.orc_trans_xtoy.restype = c_int
        _c.orc_trans_xtoy.argtypes = ([c_int] + [c_double] * 4 + [c_int] * 3 + [_dp] * 4 +
                                      [_dp] * 4)
    return _c


def ref_available() -> bool:
    return os.path.exists(_REF_LIB)


def _lib_ref():
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_LIB):
            raise RuntimeError(f"reference library missing: {_REF_LIB} (make -C oracle ref)")
        _ref = ctypes.CDLL(_REF_LIB)
        _ref.ref_letkf_core.restype = None
        _ref.ref_letkf_core.argtypes = [c_int, c_int, c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        c_int, _dp, c_int, c_int, _dp, c_int]
        _ref.ref_mtx_eigen.restype = None
        _ref.ref_mtx_eigen.argtypes = [c_int, _dp, _dp, _dp, _ip]
        _ref.ref_l96_rk4.restype = None
        _ref.ref_l96_rk4.argtypes = [c_int, c_double, c_double, _dp, _dp]
        _ref.ref_l96_nx.restype = c_int
        _ref.ref_l96_nx.argtypes = []
        _ref.ref_solve_apply.restype = c_int
        _ref.ref_solve_apply.argtypes = [c_int, c_int, c_int, c_int, _lp, _ip, _dp, _dp, _dp,
                                         _dp, c_int, _dp, _dp, _dp, c_double]
        _ref.ref_max_threads.restype = c_int
        _ref.ref_max_threads.argtypes = []
        _ref.ref_com_gamma.restype = c_double
        _ref.ref_com_gamma.argtypes = [c_double]
        _ref.ref_com_distll_1.restype = c_double
        _ref.ref_com_distll_1.argtypes = [c_double] * 4
        _ref.ref_spd_grid.restype = None
        _ref.ref_spd_grid.argtypes = [_ip]
        for f in (_ref.ref_spd_read_grd, _ref.ref_spd_write_grd):
            f.restype = None
            f.argtypes = [ctypes.c_char_p, c_int, _dp, _dp]
        _ref.ref_spd_obsope1.restype = None
        _ref.ref_spd_obsope1.argtypes = [_dp, _dp] + [c_double] * 5 + [_dp] * 4 + [_ip]
        _ref.ref_spd_get_nobs.restype = c_int
        _ref.ref_spd_get_nobs.argtypes = [ctypes.c_char_p, c_int, c_int]
        _ref.ref_spd_read_obs.restype = None
        _ref.ref_spd_read_obs.argtypes = [ctypes.c_char_p, c_int, c_int] + [_dp] * 6
        for f in (_ref.ref_spd_read_obs2, _ref.ref_spd_write_obs2):
            f.restype = None
            f.argtypes = [ctypes.c_char_p, c_int, c_int] + [_dp] * 7 + [_ip]
    return _ref


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def letkf_core(hdxb, rdiag, rloc, dep, parm_infl, *, transm=False, pao=False, rdiag_wloc=False,
               minfl=None, impl="c"):
    """letkf_core (common/common_letkf.f90:50-243) on one local problem.

    hdxb: (nobsl, ne) array of obs-space perturbations (row = obs).
    Returns dict(trans, transm, pao, parm_infl, status).
    """
    hdxb = np.asarray(hdxb, dtype=np.float64)
    p, ne = hdxb.shape
    ld = max(p, 1)
    h = np.zeros((ne, ld), dtype=np.float64)  # column-major (ld, ne) == row-major (ne, ld)
    h[:, :p] = hdxb.T
    rd = np.ones(ld)
    rl = np.ones(ld)
    dp = np.zeros(ld)
    rd[:p] = rdiag
    rl[:p] = rloc
    dp[:p] = dep
    trans = np.zeros((ne, ne))
    tm = np.zeros(ne)
    pa = np.zeros((ne, ne))
    parm = np.array([parm_infl], dtype=np.float64)
    mf = np.array([minfl if minfl is not None else 0.0])
    if impl == "c":
        st = _lib_c().orc_letkf_core(ne, ld, p, _d(h), _d(rd), _d(rl), _d(dp), _d(parm),
                                     _d(trans), _d(tm) if transm else None,
                                     _d(pa) if pao else None, int(bool(rdiag_wloc)),
                                     _d(mf) if minfl is not None else None)
    elif impl == "ref":
        _lib_ref().ref_letkf_core(ne, ld, p, _d(h), _d(rd), _d(rl), _d(dp), _d(parm), _d(trans),
                                  _d(tm), int(bool(transm)), _d(pa), int(bool(pao)),
                                  int(bool(rdiag_wloc)), _d(mf), int(minfl is not None))
        st = 0
    else:
        raise ValueError(impl)
    # column-major (ne, ne) buffers: transpose the row-major view back
    return dict(trans=trans.T.copy(), transm=tm if transm else None,
                pao=pa.T.copy() if pao else None, parm_infl=float(parm[0]), status=int(st))


def l96_rk4(x, kt, dt=0.005, force=8.0, impl="c"):
    """kt RK4 steps of Lorenz-96 (lorenz96/model/lorenz96.f90:93-128,290-309).

    impl "c": the restatement, any nx; "ref": the reference model itself
    (oracle/_ref, its nx = 40 only)."""
    x = _f64(x).copy()
    if impl == "c":
        _lib_c().orc_l96_rk4(x.size, int(kt), float(dt), float(force), _d(x))
        return x
    lib = _lib_ref()
    if x.size != lib.ref_l96_nx():
        raise ValueError("the reference model is compiled for nx = %d" % lib.ref_l96_nx())
    out = np.zeros_like(x)
    lib.ref_l96_rk4(int(kt), float(dt), float(force), _d(x), _d(out))
    return out


RELAX_NONE, RELAX_RTPP, RELAX_RTPS = 0, 1, 2


def relax_point(xb, trans, transm, pa, mode, alpha, beta=1.0, q_sprd_max=0.0, q_row=-1):
    """Relaxed analysis of one point (scale/letkf/letkf_tools.f90:215-268).

    xb: (nrow, k) background rows; trans = Wa, transm = w-bar, pa = Pa from
    letkf_core(..., transm=True, pao=True).  Returns (xa (nrow, k), infl (nrow)).
    """
    xb = _f64(np.atleast_2d(xb))
    nrow, k = xb.shape
    xa = np.zeros_like(xb)
    infl = np.zeros(nrow)
    _lib_c().orc_relax_point(k, nrow, _d(xb), _d(_f64(np.asarray(trans).T)), _d(_f64(transm)),
                             _d(_f64(np.asarray(pa).T)), int(mode), float(alpha), float(beta),
                             float(q_sprd_max), int(q_row), _d(xa), _d(infl))
    return xa, infl


def mtx_eigen(a, impl="c"):
    """mtx_eigen (common/common_mtx.f90:41-99): (eival desc, eivec columns, nrank, status)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    af = _f64(a.T)  # column-major
    ev = np.zeros(n)
    vec = np.zeros((n, n))
    nr = np.zeros(1, dtype=np.int32)
    if impl == "c":
        st = _lib_c().orc_mtx_eigen(n, _d(af), _d(ev), _d(vec), _i(nr))
    else:
        _lib_ref().ref_mtx_eigen(n, _d(af), _d(ev), _d(vec), _i(nr))
        st = 0
    return ev, vec.T.copy(), int(nr[0]), int(st)


def distll_1(alon, alat, blon, blat):
    return _lib_c().orc_distll_1(alon, alat, blon, blat)


def obs_local_l96(nx, obs_gidx, ix, xlocal):
    """lorenz96/letkf/letkf.f90:216-254 selection: (idx, rloc)."""
    g = np.ascontiguousarray(obs_gidx, dtype=np.int32)
    ny = g.size
    idx = np.zeros(max(ny, 1), dtype=np.int32)
    rl = np.zeros(max(ny, 1))
    n = _lib_c().orc_obs_local_l96(nx, ny, _i(g), ix, xlocal, _i(idx), _d(rl))
    return idx[:n].copy(), rl[:n].copy()


def obs_bins_global(olon, olat, lon, lat):
    """set_letkf_obs binning (speedy/letkf/letkf_obs.f90:243-313): (perm, nobsgrd[nlat, nlon])."""
    olon = _f64(olon)
    olat = _f64(olat)
    lon = _f64(lon)
    lat = _f64(lat)
    nobs = olon.size
    perm = np.zeros(max(nobs, 1), dtype=np.int32)
    grd = np.zeros((lat.size, lon.size), dtype=np.int32)  # column-major (nlon, nlat)
    n = _lib_c().orc_obs_bins_global(nobs, _d(olon), _d(olat), lon.size, lat.size, _d(lon), _d(lat),
                                     _i(perm), _i(grd))
    return perm[:n].copy(), grd


def set_letkf_obs_global(obs, lon, lat, gross_error, qc=None):
    """set_letkf_obs (speedy/letkf/letkf_obs.f90:146-313) on raw obs records.

    obs: dict elm, lon, lat, lev, dat, err (nobs) and hxf (nobs, ne) = H(x) per member.
    Ensemble mean of H(x) in member order (:160-163), perturbations (:164-166),
    departure (:167), gross-error QC (:168-170), then the stable lat-band /
    lon-cell binning (obs_bins_global).  Returns (sorted obs dict with dep and
    hdxf (ne, nobs) C-order, nobsgrd (nlat, nlon), perm).
    """
    hx = np.asarray(obs["hxf"], dtype=np.float64)
    ne = hx.shape[1]
    m = hx[:, 0].copy()
    for i in range(1, ne):
        m = m + hx[:, i]
    m = m / float(ne)
    dep = np.asarray(obs["dat"], dtype=np.float64) - m
    keep = ~(np.abs(dep) > gross_error * np.asarray(obs["err"]))
    if qc is not None:
        keep &= np.asarray(qc) == 1
    kidx = np.flatnonzero(keep)
    pk, grd = obs_bins_global(np.asarray(obs["lon"])[kidx], np.asarray(obs["lat"])[kidx], lon, lat)
    perm = kidx[pk]
    s = {k: np.asarray(obs[k], dtype=np.float64)[perm] for k in ("elm", "lon", "lat", "lev", "dat",
                                                                 "err")}
    s["dep"] = dep[perm]
    s["hdxf"] = np.ascontiguousarray((hx - m[:, None])[perm].T)
    return s, grd, perm


def rect_candidates(lon1, lat1, dlon_zero, dlat_zero, lon, lat, nobsgrd, nobs):
    lon = _f64(lon)
    lat = _f64(lat)
    g = np.ascontiguousarray(nobsgrd, dtype=np.int32)
    use = np.zeros(max(nobs, 1) * 2, dtype=np.int32)
    n = _lib_c().orc_rect_candidates(lon1, lat1, dlon_zero, dlat_zero, lon.size, lat.size, _d(lon),
                                     _d(lat), _i(g), _i(use))
    return use[:n].copy()


def obs_local_global(lon1, lat1, dlon_zero, dlat_zero, logpfm, ilev, lon, lat, nobsgrd, obs,
                     dist_zero, dist_zerov, sigma_obs, sigma_obsv, var_local_row=None):
    """obs_local (speedy/letkf/letkf_tools.f90:272-573): (idx, rloc, rdiag)."""
    lon = _f64(lon)
    lat = _f64(lat)
    g = np.ascontiguousarray(nobsgrd, dtype=np.int32)
    nobs = obs["elm"].size
    vl = _f64(np.ones(7) if var_local_row is None else var_local_row)
    use = np.zeros(max(nobs, 1) * 2, dtype=np.int32)
    idx = np.zeros(max(nobs, 1), dtype=np.int32)
    rl = np.zeros(max(nobs, 1))
    rd = np.zeros(max(nobs, 1))
    arrs = [_f64(obs[k]) for k in ("elm", "lon", "lat", "lev", "dat", "err")]
    n = _lib_c().orc_obs_local_global(lon1, lat1, dlon_zero, dlat_zero, logpfm, ilev, lon.size,
                                      lat.size, _d(lon), _d(lat), _i(g), *[_d(a) for a in arrs],
                                      dist_zero, dist_zerov, sigma_obs, sigma_obsv, _d(vl), _i(use),
                                      _i(idx), _d(rl), _d(rd))
    return idx[:n].copy(), rl[:n].copy(), rd[:n].copy()


DIST_ZERO_FAC_SCALE = 3.651483717  # scale/letkf/letkf_obs.f90:29


def obs_bins_regional(ri, rj, nx, ny, qc=None):
    """Regional obs index (scale/letkf/letkf_obs.f90:740-830): (perm, nobsgrd (ny, nx+1))."""
    ri = _f64(ri)
    rj = _f64(rj)
    q = np.ascontiguousarray(qc, dtype=np.int32) if qc is not None else None
    perm = np.zeros(max(ri.size, 1), dtype=np.int32)
    grd = np.zeros((ny, nx + 1), dtype=np.int32)  # column-major (0:nx, ny)
    n = _lib_c().orc_obs_bins_regional(ri.size, _d(ri), _d(rj), _i(q) if q is not None else None,
                                       nx, ny, _i(perm), _i(grd))
    return perm[:n].copy(), grd


def obs_local_regional(ri, rj, rz, nx, ny, nobsgrd, obs, dx, dy, sigma_h, sigma_v, nobs_max,
                       nbkt):
    """Regional obs_local with the per-bucket cap (scale/letkf/letkf_tools.f90:853-1178).

    obs: sorted arrays ri, rj, z, err (float) and bucket (int).  Returns (idx, rloc, rdiag)
    in the reference's order (scan order uncapped; bucket then slot order capped).
    """
    dlon_zero = sigma_h * DIST_ZERO_FAC_SCALE / dx
    dlat_zero = sigma_h * DIST_ZERO_FAC_SCALE / dy
    ori, orj, oz, oe = (_f64(obs[k]) for k in ("ri", "rj", "z", "err"))
    ob = np.ascontiguousarray(obs["bucket"], dtype=np.int32)
    n = ori.size
    cap = max(nobs_max, 1)
    isort = np.zeros(nbkt * cap, dtype=np.int32)
    sobs = np.zeros(nbkt * cap, dtype=np.int32)
    srd = np.zeros(nbkt * cap)
    srl = np.zeros(nbkt * cap)
    cnt = np.zeros(nbkt, dtype=np.int32)
    idx = np.zeros(max(n, 1), dtype=np.int32)
    rl = np.zeros(max(n, 1))
    rd = np.zeros(max(n, 1))
    g = np.ascontiguousarray(nobsgrd, dtype=np.int32)
    m = _lib_c().orc_obs_local_regional(ri, rj, rz, nx, ny, _i(g), _d(ori), _d(orj), _d(oz), _d(oe),
                                        _i(ob), nbkt, dx, dy, sigma_h, sigma_v, dlon_zero,
                                        dlat_zero, nobs_max, _i(isort), _i(sobs), _d(srd), _d(srl),
                                        _i(cnt), _i(idx), _d(rl), _d(rd))
    return idx[:m].copy(), rl[:m].copy(), rd[:m].copy()


def cap_stream(key, cap):
    """The reference's per-bucket cap over one candidate stream in scan order
    (scale/letkf/letkf_tools.f90:1041-1156): key = priority (smaller first).
    Returns the candidate index of each storage slot, in slot order (the
    reference's output order)."""
    key = _f64(key)
    cap = int(cap)
    isort = np.zeros(max(cap, 1), dtype=np.int32)
    src = np.zeros(max(cap, 1), dtype=np.int32)
    m = _lib_c().orc_cap_stream(key.size, _d(key), cap, _i(isort), _i(src))
    return src[:m].copy()


def cap_nearest_first(idx, rho, cap):
    """The global grid's nobs_max cap: the reference cap with key = -rho, so the
    cap entries of largest rho (ties: the earlier list entry) are kept, in
    storage-slot order.  Returns (idx, rho) capped."""
    idx = np.asarray(idx)
    rho = _f64(rho)
    if cap <= 0 or idx.size <= cap:
        return idx.copy(), rho.copy()
    keep = cap_stream(-rho, cap)
    return idx[keep].copy(), rho[keep].copy()


def das_global(prob, *, q_top_lev=5, iq=3, var_local_row=None, return_nobsl=False):
    """das_letkf (speedy/letkf/letkf_tools.f90:45-267) on a global problem dict.

    prob keys (numpy, reference column-major shapes expressed C-order reversed):
      gues3d (nv3d, ne, nlev, nij), gues2d (nv2d, ne, nij), lon1, lat1 (nij), sig (nlev),
      lon (nlon), lat (nlat), nobsgrd (nlat, nlon), obs dict with sorted arrays and
      hdxf (ne, nobs), infl3d (nlev, nij), infl2d (nv2d, nij), sigma_obs, sigma_obsv, ips.
    Returns (anal3d, anal2d, infl3d, infl2d[, nobsl (nlev, nij)]).
    """
    g3 = _f64(prob["gues3d"]).copy()
    g2 = _f64(prob["gues2d"]).copy()
    nv3d, ne, nlev, nij = g3.shape
    nv2d = g2.shape[0]
    a3 = np.zeros_like(g3)
    a2 = np.zeros_like(g2)
    i3 = _f64(prob["infl3d"]).copy()
    i2 = _f64(prob["infl2d"]).copy()
    obs = prob["obs"]
    nobs = obs["elm"].size
    hd = _f64(obs["hdxf"])
    vl = _f64(np.ones(7) if var_local_row is None else var_local_row)
    nl = np.zeros((nlev, nij), dtype=np.int32)
    lon = _f64(prob["lon"])
    lat = _f64(prob["lat"])
    arrs = [_f64(obs[k]) for k in ("elm", "lon", "lat", "lev", "dat", "err", "dep")]
    grd = np.ascontiguousarray(prob["nobsgrd"], dtype=np.int32)
    st = _lib_c().orc_das_global(nij, nlev, ne, nv3d, nv2d, _d(g3), _d(g2), _d(a3), _d(a2),
                                 _d(_f64(prob["lon1"])), _d(_f64(prob["lat1"])),
                                 _d(_f64(prob["sig"])), int(prob.get("ips", 0)), _d(i3), _d(i2),
                                 lon.size, lat.size, _d(lon), _d(lat), nobs, _i(grd),
                                 *[_d(a) for a in arrs], _d(hd), float(prob["sigma_obs"]),
                                 float(prob["sigma_obsv"]), _d(vl), int(q_top_lev), int(iq),
                                 _i(nl))
    if st != 0:
        raise RuntimeError(f"oracle das_global status {st}")
    out = (a3, a2, i3, i2)
    return out + (nl,) if return_nobsl else out


def das_global_problem(p, *, q_top_lev=5, iq=3, infl3d=None):
    """set_letkf_obs + das_letkf on a letkf_amd.synth global problem dict.

    Returns dict(anal3d (nij,nlev,ne,nv3d), anal2d (nij,ne,nv2d), infl3d (nij,nlev),
    nobsl (nij,nlev), perm, nobsgrd (nlat,nlon), sobs) in the reference layouts.
    """
    s, grd, perm = set_letkf_obs_global(p["obs"], p["lon"], p["lat"], p["gross_error"])
    nij, nlev = p["nij"], p["nlev"]
    nv2d = p["gues2d"].shape[2]
    i3 = np.full((nlev, nij), p["infl"]) if infl3d is None else np.ascontiguousarray(infl3d.T)
    prob = dict(gues3d=np.ascontiguousarray(p["gues3d"].transpose(3, 2, 1, 0)),
                gues2d=np.ascontiguousarray(p["gues2d"].transpose(2, 1, 0)),
                lon1=np.tile(p["lon"], p["nlat"]), lat1=np.repeat(p["lat"], p["nlon"]),
                sig=p["sig"], lon=p["lon"], lat=p["lat"], nobsgrd=grd, obs=s, infl3d=i3,
                infl2d=np.full((nv2d, nij), p["infl"]), sigma_obs=p["sigma_obs"],
                sigma_obsv=p["sigma_obsv"], ips=p.get("ips", 0))
    a3, a2, o3, o2, nl = das_global(prob, q_top_lev=q_top_lev, iq=iq, return_nobsl=True)
    return dict(anal3d=np.asfortranarray(a3.transpose(3, 2, 1, 0)),
                anal2d=np.asfortranarray(a2.transpose(2, 1, 0)), infl3d=o3.T.copy(),
                nobsl=nl.T.copy(), perm=perm, nobsgrd=grd, sobs=s)


def phys2ij(rlon, rlat, nlon, lat):
    """phys2ijk lon/lat part (speedy/common/common_obs_speedy.f90:156-175): (status, ri, rj)."""
    lat = _f64(lat)
    ri = np.zeros(1)
    rj = np.zeros(1)
    st = _lib_c().orc_phys2ij(rlon, rlat, nlon, lat.size, _d(lat), _d(ri), _d(rj))
    return st, float(ri[0]), float(rj[0])


def trans_xtoy(elm, rlon, rlat, rlev, odat, lat, sig, v3d, ps):
    """Trans_XtoY + phys2ijk + obsope range rules for one member.

    v3d: (4, nlev, nlat, nlon) C-order (== column-major nlon x nlat x nlev x 4), ps (nlat, nlon).
    Returns (status, yobs, ri, rj, rk).
    """
    lat = _f64(lat)
    sig = _f64(sig)
    v3d = _f64(v3d)
    ps = _f64(ps)
    nlon = ps.shape[1]
    y = np.zeros(1)
    ri = np.zeros(1)
    rj = np.zeros(1)
    rk = np.zeros(1)
    st = _lib_c().orc_trans_xtoy(int(elm), rlon, rlat, rlev, odat, nlon, lat.size, sig.size,
                                 _d(lat), _d(sig), _d(v3d), _d(ps), _d(y), _d(ri), _d(rj), _d(rk))
    return st, float(y[0]), float(ri[0]), float(rj[0]), float(rk[0])


def ref_obsope1(elm, rlon, rlat, rlev, odat, v3d, v2d):
    """The reference's phys2ijk + obsope range rules + Trans_XtoY for one obs
    (speedy/common/common_obs_speedy.f90:30-282, speedy/letkf/obsope.f90:56-89) on
    the T30 grid.  v3d: (nv3d, nlev, nlat, nlon) C-order, v2d: (nv2d, nlat, nlon).
    Returns (status, yobs, ri, rj, rk); status 0 = used."""
    out = np.zeros(4)
    st = np.zeros(1, dtype=np.int32)
    v3 = _f64(v3d)
    v2 = _f64(v2d)
    _lib_ref().ref_spd_obsope1(_d(v3), _d(v2), float(elm), float(rlon), float(rlat), float(rlev),
                               float(odat), _d(out[0:]), _d(out[1:]), _d(out[2:]), _d(out[3:]),
                               _i(st))
    return int(st[0]), float(out[3]), float(out[0]), float(out[1]), float(out[2])


def ref_solve_apply(loc_off, loc_idx, loc_rho, hdxf, rdiag, dep, xb, infl, fixed_infl=0.0):
    """Reference letkf_core + apply over packed local problems (OpenMP over points).

    hdxf: (nobs, ne) row = obs; xb: (npt, nv, ne).  Returns (xa, infl_out).
    """
    lib = _lib_ref()
    npt = loc_off.size - 1
    nobs, ne = hdxf.shape
    nv = xb.shape[1]
    off = np.ascontiguousarray(loc_off, dtype=np.int64)
    idx = np.ascontiguousarray(loc_idx, dtype=np.int32)
    rho = _f64(loc_rho)
    h = _f64(np.asarray(hdxf).T)  # (ne, nobs) C-order == column-major (nobs, ne)
    xbf = _f64(xb)
    xa = np.zeros_like(xbf)
    inf = _f64(infl).copy()
    maxp = int(np.max(np.diff(off))) if npt else 1
    lib.ref_solve_apply(npt, ne, nobs, max(maxp, 1), off.ctypes.data_as(_lp), _i(idx), _d(rho),
                        _d(h), _d(_f64(rdiag)), _d(_f64(dep)), nv, _d(xbf), _d(xa), _d(inf),
                        float(fixed_infl))
    return xa, inf


def ref_threads() -> int:
    return int(_lib_ref().ref_max_threads())


# ---------------------------------------------------------------------------
# SPEEDY files through the reference's own routines (native byte order)
# ---------------------------------------------------------------------------
def ref_speedy_grid():
    """(nlon, nlat, nlev, nv3d, nv2d) of speedy/common/common_speedy.f90:18-22."""
    d = np.zeros(5, dtype=np.int32)
    _lib_ref().ref_spd_grid(_i(d))
    return tuple(int(v) for v in d)


def ref_read_grd(path):
    """read_grd (speedy/common/common_speedy.f90:163-194): (v3d, v2d) in float64."""
    nlon, nlat, nlev, nv3d, nv2d = ref_speedy_grid()
    v3 = np.zeros((nlon, nlat, nlev, nv3d), order="F")
    v2 = np.zeros((nlon, nlat, nv2d), order="F")
    b = os.fsencode(path)
    _lib_ref().ref_spd_read_grd(b, len(b), _d(v3), _d(v2))
    return v3, v2


def ref_write_grd(path, v3d, v2d):
    """write_grd (speedy/common/common_speedy.f90:226-257)."""
    v3 = np.asfortranarray(v3d, dtype=np.float64)
    v2 = np.asfortranarray(v2d, dtype=np.float64)
    g = ref_speedy_grid()
    if v3.shape != g[:4] or v2.shape != g[:2] + g[4:]:
        raise ValueError("grid shape differs from the reference's SPEEDY grid")
    b = os.fsencode(path)
    _lib_ref().ref_spd_write_grd(b, len(b), _d(v3), _d(v2))


def ref_get_nobs(path, nrec=8):
    """get_nobs (speedy/common/common_obs_speedy.f90:400-456): the record count."""
    b = os.fsencode(path)
    return int(_lib_ref().ref_spd_get_nobs(b, len(b), int(nrec)))


_OBS_KEYS = ("elem", "rlon", "rlat", "rlev", "odat", "oerr")


def ref_read_obs(path, nn):
    """read_obs (speedy/common/common_obs_speedy.f90:458-502), 6-value records."""
    out = {k: np.zeros(nn) for k in _OBS_KEYS}
    b = os.fsencode(path)
    _lib_ref().ref_spd_read_obs(b, len(b), int(nn), *[_d(out[k]) for k in _OBS_KEYS])
    return out


def ref_read_obs2(path, nn):
    """read_obs2 (speedy/common/common_obs_speedy.f90:504-552), 8-value records."""
    out = {k: np.zeros(nn) for k in _OBS_KEYS + ("ohx",)}
    out["oqc"] = np.zeros(nn, dtype=np.int32)
    b = os.fsencode(path)
    _lib_ref().ref_spd_read_obs2(b, len(b), int(nn), *[_d(out[k]) for k in _OBS_KEYS + ("ohx",)],
                                 _i(out["oqc"]))
    return out


def ref_write_obs2(path, obs):
    """write_obs2 (speedy/common/common_obs_speedy.f90:554-602); obs in the
    in-memory units (Pa levels, Pa ps, fractional rh)."""
    arr = {k: _f64(obs[k]) for k in _OBS_KEYS + ("ohx",)}
    qc = np.ascontiguousarray(obs["oqc"], dtype=np.int32)
    b = os.fsencode(path)
    _lib_ref().ref_spd_write_obs2(b, len(b), int(qc.size), *[_d(arr[k]) for k in _OBS_KEYS + ("ohx",)],
                                  _i(qc))


# ---------------------------------------------------------------------------
# SCALE radar obs operator (orc_phys2ijkz, orc_itpl_radar, orc_calc_ref_vr,
# orc_radar_xtoy)
# ---------------------------------------------------------------------------
RADAR_FIELDS = ("u", "v", "w", "t", "p", "q", "qc", "qr", "qi", "qs", "qg")
ID_RADAR_REF, ID_RADAR_VR = 4001, 4002
IQC = dict(good=0, gross_err=5, ref_low=11, ref_mem=12, radar_vhi=19, out_vhi=20, out_vlo=21,
           out_h=22, obs_bad=50, otype=90)


def radar_gammas():
    """Gamma(4.8), Gamma(4.25), Gamma(4.5): com_gamma of 4 + b, 4 + d and 4.5 in
    calc_ref_vr (scale/common/common_obs_scale.f90)."""
    import math
    return np.array([math.gamma(4.8), math.gamma(4.25), math.gamma(4.5)])


def _radar_lib():
    lib = _lib_c()
    if not getattr(lib, "_radar_bound", False):
        lib.orc_phys2ijkz.restype = c_int
        lib.orc_phys2ijkz.argtypes = [_dp, c_int, c_int, c_int, c_double, c_double, c_double, _dp]
        lib.orc_itpl_radar.restype = c_double
        lib.orc_itpl_radar.argtypes = [_dp, c_long, c_int, c_double, c_double, c_double]
        lib.orc_calc_ref_vr.restype = c_int
        lib.orc_calc_ref_vr.argtypes = [c_double] * 13 + [c_int, c_int, _dp, _dp, _dp]
        lib.orc_radar_xtoy.restype = c_int
        lib.orc_radar_xtoy.argtypes = ([c_int] + [c_double] * 8 + [_dp, c_long, _dp, c_int, c_int,
                                        c_int, c_int, c_int, _dp, _dp, _dp, _dp])
        lib._radar_bound = True
    return lib


def phys2ijkz(hgt, nx, ny, ri, rj, rlev):
    """(qc, rk) of SCALE's phys2ijkz; hgt (nij, nlev) column-major."""
    h = np.asfortranarray(hgt, dtype=np.float64)
    rk = np.zeros(1)
    qc = _radar_lib().orc_phys2ijkz(_d(h), int(nx), int(ny), int(h.shape[1]), float(ri), float(rj),
                                    float(rlev), _d(rk))
    return int(qc), float(rk[0])


def itpl_radar(var, nx, ri, rj, rk):
    """Trilinear value of var (nij, nlev) at 1-based (ri, rj, rk)."""
    v = np.asfortranarray(var, dtype=np.float64)
    return float(_radar_lib().orc_itpl_radar(_d(v), v.shape[0], int(nx), float(ri), float(rj),
                                             float(rk)))


def calc_ref_vr(fields, az, elev, method=3, use_wt=False, gam=None):
    """(ref, vr) of calc_ref_vr from a dict of the point values of RADAR_FIELDS."""
    g = radar_gammas() if gam is None else np.ascontiguousarray(gam, dtype=np.float64)
    ref = np.zeros(1)
    vr = np.zeros(1)
    f = fields
    st = _radar_lib().orc_calc_ref_vr(f["q"], f["qc"], f["qr"], f["qi"], f["qs"], f["qg"], f["u"],
                                      f["v"], f["w"], f["t"], f["p"], float(az), float(elev),
                                      int(method), int(bool(use_wt)), _d(g), _d(ref), _d(vr))
    if st != 0:
        raise ValueError(f"unknown reflectivity method {method}")
    return float(ref[0]), float(vr[0])


def radar_xtoy(obs, radar, x3, hgt, nx, ny, method=3, use_wt=False, min_ref_dbz=0.0,
               low_ref_shift=0.0, zmax=99.0e3, gam=None):
    """SCALE's radar obsope on every (obs, member).

    obs: dict elm, lon, lat, z, ri, rj (nobs); radar: dict lon, lat, z (nobs);
    x3: (nij, nlev, ne, 11) fields in RADAR_FIELDS order, column-major;
    hgt (nij, nlev).  Returns hxf (nobs, ne), qc (nobs, ne), rk (nobs)."""
    lib = _radar_lib()
    x = np.asfortranarray(x3, dtype=np.float64)
    h = np.asfortranarray(hgt, dtype=np.float64)
    nij, nlev, ne, nf = x.shape
    assert nf == 11 and nij == nx * ny and h.shape == (nij, nlev)
    g = radar_gammas() if gam is None else np.ascontiguousarray(gam, dtype=np.float64)
    opt = np.array([min_ref_dbz, low_ref_shift, zmax])
    n = np.asarray(obs["elm"]).size
    hxf = np.zeros((n, ne))
    qc = np.zeros((n, ne), dtype=np.int32)
    rkv = np.zeros(n)
    y = np.zeros(1)
    rk = np.zeros(1)
    members = [np.asfortranarray(x[:, :, m, :]) for m in range(ne)]
    for i in range(n):
        for m in range(ne):
            xm = members[m]
            q = lib.orc_radar_xtoy(int(obs["elm"][i]), float(obs["lon"][i]), float(obs["lat"][i]),
                                   float(obs["z"][i]), float(obs["ri"][i]), float(obs["rj"][i]),
                                   float(radar["lon"][i]), float(radar["lat"][i]),
                                   float(radar["z"][i]), _d(xm), nij * nlev, _d(h), int(nx), int(ny),
                                   int(nlev), int(method), int(bool(use_wt)), _d(opt), _d(g), _d(y),
                                   _d(rk))
            hxf[i, m] = y[0]
            qc[i, m] = q
        rkv[i] = rk[0]
    return hxf, qc, rkv


# ---------------------------------------------------------------------------
# SCALE departure / QC step of set_letkf_obs (orc_obs_departure_scale)
# ---------------------------------------------------------------------------
ID_RAIN, ID_RADAR_PRH, ID_H08IR = 19999, 4003, 8800
ID_TCLON, ID_TCLAT, ID_TCMIP = 99991, 99992, 99993
UNDEF = -9.99e33
# namelist defaults (scale/common/common_nml.f90:83-91,140-146,165-174)
SCALE_QC_DEFAULTS = dict(use_radar_ref=True, use_radar_vr=True, use_obserr_radar_ref=False,
                         use_obserr_radar_vr=False, min_radar_ref_member=1,
                         min_radar_ref_member_obsref=1, radar_ref_thres_dbz=15.0,
                         min_radar_ref_dbz=0.0, low_ref_shift=0.0, obserr_radar_ref=5.0,
                         obserr_radar_vr=3.0, gross_error=5.0, gross_error_rain=-1.0,
                         gross_error_radar_ref=-1.0, gross_error_radar_vr=-1.0,
                         gross_error_radar_prh=-1.0, gross_error_tcx=-1.0, gross_error_tcy=-1.0,
                         gross_error_tcp=-1.0)


def _scale_qc_cfg(**kw):
    c = dict(SCALE_QC_DEFAULTS)
    for k, v in kw.items():
        if k not in c:
            raise TypeError(f"unknown option {k}")
        c[k] = v
    return np.array([float(c[k]) for k in SCALE_QC_DEFAULTS])


def obs_departure_scale(elm, dat, err, hxf, qc_member=None, **cfg):
    """SCALE's departure and QC step (orc_obs_departure_scale) over nobs obs.

    hxf (nobs, ne): H(x) of each member; qc_member (nobs, ne) or None.  Returns
    (dat, err, dep, hdxf, qc): the preprocessed obs value and error, y - mean
    H(x) (0 unless qc is good), the perturbations (H(x) where qc is not good)
    and the qc."""
    lib = _lib_c()
    if not getattr(lib, "_dep_bound", False):
        lib.orc_obs_departure_scale.restype = c_int
        lib.orc_obs_departure_scale.argtypes = [c_int, c_double, c_double, _ip, c_int, _dp, _dp,
                                                _dp, _dp, _dp]
        lib._dep_bound = True
    c = _scale_qc_cfg(**cfg)
    hx = np.array(hxf, dtype=np.float64, order="F")
    n, ne = hx.shape
    qm = None if qc_member is None else np.asarray(qc_member, dtype=np.int32)
    elm = np.asarray(elm)
    out = [np.zeros(n) for _ in range(3)]
    qc = np.zeros(n, dtype=np.int32)
    y, e, d = np.zeros(1), np.zeros(1), np.zeros(1)
    for i in range(n):
        row = np.ascontiguousarray(hx[i])
        qrow = None if qm is None else np.ascontiguousarray(qm[i])
        q = lib.orc_obs_departure_scale(int(elm[i]), float(dat[i]), float(err[i]),
                                        _i(qrow) if qrow is not None else None, ne, _d(row), _d(c),
                                        _d(y), _d(e), _d(d))
        if q < 0:
            raise ValueError("H08 obs need the RTTOV operator's cloud flags")
        hx[i] = row
        out[0][i], out[1][i], out[2][i], qc[i] = y[0], e[0], d[0], q
    return out[0], out[1], out[2], hx, qc
