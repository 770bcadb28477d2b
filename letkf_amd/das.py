"""das_letkf-level host API: a handle that keeps the ensemble resident on the GPU.

Mirrors the reference drivers:

* :class:`GlobalLetkf` -- ``das_letkf`` of speedy/letkf/letkf_tools.f90:45-267 with
  ``set_letkf_obs`` (speedy/letkf/letkf_obs.f90:48-335): lat/lon grid, sigma
  levels, 3-D variables plus 2-D variables updated with the level-1 transform.
* :class:`L96Letkf` -- the Lorenz-96 analysis loop of lorenz96/letkf/letkf.f90:216-265.
* :class:`RegionalLetkf` -- SCALE-style cartesian grid with the per-type
  ``MAX_NOBS_PER_GRID`` cap (scale/letkf/letkf_tools.f90:853-1178).

Arrays cross the boundary in the reference's column-major layouts: numpy
arrays are passed as Fortran-ordered (e.g. ``gues3d`` of shape
(nij, nlev, ne, nv3d)).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr, iptr


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


class _Handle:
    h = None

    def __del__(self):
        self.close()

    def close(self):
        if self.h is not None:
            _lib.lib().lekf_destroy(self.h)
            self.h = None

    # -- common ---------------------------------------------------------------
    def set_ensemble(self, gues3d, gues2d=None):
        g3 = _f(gues3d)
        g2 = _f(gues2d) if gues2d is not None else None
        check(_lib.lib().lekf_set_ensemble(self.h, g3.ctypes.data_as(_lib._dp),
                                           g2.ctypes.data_as(_lib._dp) if g2 is not None else None),
              "set_ensemble")

    def set_infl(self, infl3d, infl2d=None):
        a = _f(infl3d)
        b = _f(infl2d) if infl2d is not None else None
        check(_lib.lib().lekf_set_infl(self.h, a.ctypes.data_as(_lib._dp),
                                       b.ctypes.data_as(_lib._dp) if b is not None else None),
              "set_infl")

    def run(self):
        check(_lib.lib().lekf_das_run(self.h), "das_run")

    def set_relax(self, mode="none", alpha=0.0, q_sprd_max=0.0, iq=-1, boundary_width=0.0):
        """SCALE/GFS driver options after the local solves: mode "rtpp" / "rtps" with
        weight alpha, the q spread limit on 3-D row iq, and (regional) the lateral
        boundary buffer width in metres."""
        check(_lib.lib().lekf_set_relax(self.h, _lib.RELAX[mode], float(alpha), float(q_sprd_max),
                                        int(iq), float(boundary_width)), "set_relax")

    def set_graph(self, enable=True):
        """Replay each cycle's launches as one HIP graph (captured on the first run
        after any change of inputs)."""
        check(_lib.lib().lekf_set_graph(self.h, int(bool(enable))), "set_graph")

    def select(self):
        check(_lib.lib().lekf_das_select(self.h), "das_select")

    def solve(self):
        check(_lib.lib().lekf_das_solve(self.h), "das_solve")

    def set_stream(self, stream):
        """Run on a caller stream (an int hipStream_t, e.g. torch.cuda.current_stream().cuda_stream)."""
        check(_lib.lib().lekf_set_stream(self.h, stream), "set_stream")

    def sync(self):
        check(_lib.lib().lekf_sync(self.h), "sync")

    def analysis(self):
        a3 = np.zeros((self.nij, self.nlev, self.ne, self.nv3d), order="F")
        a2 = np.zeros((self.nij, self.ne, max(self.nv2d, 1)), order="F")
        check(_lib.lib().lekf_get_analysis(self.h, a3.ctypes.data_as(_lib._dp),
                                           a2.ctypes.data_as(_lib._dp)), "get_analysis")
        return (a3, a2[:, :, :self.nv2d]) if self.nv2d else (a3, None)

    def infl(self):
        i3 = np.zeros((self.nij, self.nlev), order="F")
        i2 = np.zeros((self.nij, max(self.nv2d, 1)), order="F")
        check(_lib.lib().lekf_get_infl(self.h, i3.ctypes.data_as(_lib._dp),
                                       i2.ctypes.data_as(_lib._dp)), "get_infl")
        return i3, i2[:, :self.nv2d]

    def nobsl(self):
        n = np.zeros((self.nij, self.nlev), dtype=np.int32, order="F")
        check(_lib.lib().lekf_get_nobsl(self.h, n.ctypes.data_as(_lib._ip)), "get_nobsl")
        return n

    def local(self, b, cap=100000):
        idx = np.zeros(cap, dtype=np.int32)
        rho = np.zeros(cap)
        n = _lib.lib().lekf_get_local(self.h, int(b), cap, iptr(idx), dptr(rho))
        if n < 0:
            raise _lib.LetkfError(-n, "get_local")
        return idx[:min(n, cap)].copy(), rho[:min(n, cap)].copy()

    def flags(self):
        a = ctypes.c_long()
        b = ctypes.c_long()
        check(_lib.lib().lekf_get_flags(self.h, ctypes.byref(a), ctypes.byref(b)), "get_flags")
        return a.value, b.value

    def fp64_count(self):
        """Points of the last run that a mixed form solved again in FP64
        (their C was worse conditioned than the mixed_kappa tuning value)."""
        n = ctypes.c_long()
        check(_lib.lib().lekf_get_fp64_count(self.h, ctypes.byref(n)), "get_fp64_count")
        return n.value

    def cycle_swap(self):
        check(_lib.lib().lekf_cycle_swap(self.h), "cycle_swap")

    def device_bytes(self):
        return int(_lib.lib().lekf_device_bytes(self.h))

    def stream(self):
        return _lib.lib().lekf_stream(self.h)


class GlobalLetkf(_Handle):
    def __init__(self, *, lon, lat, sig, ne, nv3d=4, nv2d=1, ips=0, iq=3, q_top_lev=5,
                 sigma_obs=500.0e3, sigma_obsv=0.4, gross_error=10.0, var_local=None,
                 prec="f64", infl_mode="adaptive", infl_value=1.01, nobs_max=0):
        self._lon = np.ascontiguousarray(lon, dtype=np.float64)
        self._lat = np.ascontiguousarray(lat, dtype=np.float64)
        self._sig = np.ascontiguousarray(sig, dtype=np.float64)
        c = _lib.GlobalCfg()
        c.nlon, c.nlat, c.nlev = self._lon.size, self._lat.size, self._sig.size
        c.lon, c.lat, c.sig = dptr(self._lon), dptr(self._lat), dptr(self._sig)
        c.ne, c.nv3d, c.nv2d, c.ips, c.iq, c.q_top_lev = ne, nv3d, nv2d, ips, iq, q_top_lev
        c.sigma_obs, c.sigma_obsv, c.gross_error = sigma_obs, sigma_obsv, gross_error
        vl = np.ones(7) if var_local is None else np.asarray(var_local, dtype=np.float64)
        for i in range(7):
            c.var_local[i] = vl[i]
        c.prec = _lib.PREC[prec]
        c.infl_mode = _lib.LEKF_INFL_FIXED if infl_mode == "fixed" else _lib.LEKF_INFL_ADAPTIVE
        c.infl_value = infl_value
        c.nobs_max = nobs_max
        st = ctypes.c_int(0)
        self.h = _lib.lib().lekf_global_create(ctypes.byref(c), ctypes.byref(st))
        if not self.h:
            raise _lib.LetkfError(st.value, "lekf_global_create")
        self.nlon, self.nlat, self.nlev = c.nlon, c.nlat, c.nlev
        self.nij = c.nlon * c.nlat
        self.ne, self.nv3d, self.nv2d = ne, nv3d, nv2d

    def set_obs(self, elm, lon, lat, lev, dat, err, hxf, qc=None):
        """set_letkf_obs: raw obs records and H(x) of every member (nobs, ne)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (elm, lon, lat, lev, dat, err)]
        hx = _f(hxf)
        q = np.ascontiguousarray(qc, dtype=np.int32) if qc is not None else None
        nobs = arrs[0].size
        check(_lib.lib().lekf_set_obs_global(self.h, nobs, *[dptr(a) for a in arrs],
                                             iptr(q) if q is not None else None,
                                             hx.ctypes.data_as(_lib._dp)), "set_obs_global")

    def set_obs_hx(self, elm, lon, lat, lev, dat, err):
        """set_letkf_obs with H(x) computed on the GPU from the resident background
        (the reference obs operator; the obs go through the same QC and binning)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (elm, lon, lat, lev, dat, err)]
        check(_lib.lib().lekf_set_obs_global_hx(self.h, arrs[0].size, *[dptr(a) for a in arrs]),
              "set_obs_global_hx")

    def obs_operator(self, elm, lon, lat, lev, dat):
        """H(x) of every member of the resident background: (hxf (nobs, ne), qc (nobs),
        ri (nobs), rj (nobs), rk (nobs, ne))."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (elm, lon, lat, lev, dat)]
        n = arrs[0].size
        hxf = np.zeros((n, self.ne), order="F")
        rk = np.zeros((n, self.ne), order="F")
        qc = np.zeros(n, dtype=np.int32)
        ri = np.zeros(n)
        rj = np.zeros(n)
        check(_lib.lib().lekf_obsope_global(self.h, n, *[dptr(a) for a in arrs],
                                            hxf.ctypes.data_as(_lib._dp), iptr(qc), dptr(ri),
                                            dptr(rj), rk.ctypes.data_as(_lib._dp)),
              "obsope_global")
        return hxf, qc, ri, rj, rk

    def obs_index(self):
        nb = ctypes.c_int(0)
        grd = np.zeros((self.nlon, self.nlat), dtype=np.int32, order="F")
        check(_lib.lib().lekf_get_obs_index(self.h, None, grd.ctypes.data_as(_lib._ip),
                                            ctypes.byref(nb)), "get_obs_index")
        perm = np.zeros(max(nb.value, 1), dtype=np.int32)
        check(_lib.lib().lekf_get_obs_index(self.h, iptr(perm), None, None), "get_obs_index")
        return perm[:nb.value], grd


class L96Letkf(_Handle):
    def __init__(self, *, nx, ne, xlocal=3.0, prec="f64", infl_mode="fixed", infl_value=1.05):
        c = _lib.L96Cfg()
        c.nx, c.ne, c.xlocal = nx, ne, xlocal
        c.prec = _lib.PREC[prec]
        c.infl_mode = _lib.LEKF_INFL_FIXED if infl_mode == "fixed" else _lib.LEKF_INFL_ADAPTIVE
        c.infl_value = infl_value
        st = ctypes.c_int(0)
        self.h = _lib.lib().lekf_l96_create(ctypes.byref(c), ctypes.byref(st))
        if not self.h:
            raise _lib.LetkfError(st.value, "lekf_l96_create")
        self.nij, self.nlev, self.ne, self.nv3d, self.nv2d = nx, 1, ne, 1, 0

    def set_obs(self, gidx, y, err, hxf=None):
        """Obs at grid points gidx; hxf (ny, ne) or None for H(x) of the resident
        background computed on the GPU."""
        g = np.ascontiguousarray(gidx, dtype=np.int32)
        y = np.ascontiguousarray(y, dtype=np.float64)
        e = np.ascontiguousarray(err, dtype=np.float64)
        hx = _f(hxf) if hxf is not None else None
        check(_lib.lib().lekf_set_obs_l96(self.h, g.size, iptr(g), dptr(y), dptr(e),
                                          hx.ctypes.data_as(_lib._dp) if hx is not None else None),
              "set_obs_l96")

    def forecast(self, kt, dt=0.005, force=8.0):
        """kt RK4 steps of the Lorenz-96 model from the latest analysis; the
        forecast becomes the next cycle's background (lorenz96/model/lorenz96.f90:93-128)."""
        check(_lib.lib().lekf_l96_forecast(self.h, int(kt), float(dt), float(force)),
              "l96_forecast")


def _qc_cfg(cfg):
    """lekf_scale_qc_cfg from the namelist defaults updated by cfg."""
    c = _lib.ScaleQcCfg()
    opts = dict(_lib.SCALE_QC_DEFAULTS)
    for k, v in cfg.items():
        if k not in opts:
            raise TypeError(f"unknown option {k}")
        opts[k] = v
    for k, v in opts.items():
        setattr(c, k, int(v) if isinstance(getattr(c, k), int) else float(v))
    return c


class RegionalLetkf(_Handle):
    def __init__(self, *, nx, ny, zlev, dx, dy, ne, nv3d, sigma_h, sigma_v, nobs_max=1000,
                 prec="f64", infl_mode="adaptive", infl_value=1.0, minfl=0.0, list_budget=0):
        self._z = np.ascontiguousarray(zlev, dtype=np.float64)
        c = _lib.RegionalCfg()
        c.nx, c.ny, c.nlev = nx, ny, self._z.size
        c.dx, c.dy = dx, dy
        c.zlev = dptr(self._z)
        c.ne, c.nv3d = ne, nv3d
        c.sigma_h, c.sigma_v, c.nobs_max = sigma_h, sigma_v, nobs_max
        c.prec = _lib.PREC[prec]
        c.infl_mode = _lib.LEKF_INFL_FIXED if infl_mode == "fixed" else _lib.LEKF_INFL_ADAPTIVE
        c.infl_value = infl_value
        c.minfl = minfl
        c.list_budget = int(list_budget)
        st = ctypes.c_int(0)
        self.h = _lib.lib().lekf_regional_create(ctypes.byref(c), ctypes.byref(st))
        if not self.h:
            raise _lib.LetkfError(st.value, "lekf_regional_create")
        self.nij, self.nlev, self.ne, self.nv3d, self.nv2d = nx * ny, c.nlev, ne, nv3d, 0
        self.nx, self.ny = nx, ny

    def set_obs(self, ri, rj, z, err, dep, bucket, hdxf, qc=None):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (ri, rj, z, err, dep)]
        b = np.ascontiguousarray(bucket, dtype=np.int32)
        q = np.ascontiguousarray(qc, dtype=np.int32) if qc is not None else None
        hx = _f(hdxf)
        check(_lib.lib().lekf_set_obs_regional(self.h, a[0].size, *[dptr(x) for x in a], iptr(b),
                                               iptr(q) if q is not None else None,
                                               hx.ctypes.data_as(_lib._dp)), "set_obs_regional")

    def set_heights(self, hgt):
        """Heights (m) of the model levels, hgt (nij, nlev), for the radar operator."""
        self._hgt = _f(hgt)
        check(_lib.lib().lekf_set_heights(self.h, self._hgt.ctypes.data_as(_lib._dp)), "set_heights")

    def radar_operator(self, elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z,
                       var=tuple(range(11)), method=3, use_terminal_velocity=False,
                       min_ref_dbz=0.0, low_ref_shift=0.0, zmax=99.0e3):
        """SCALE's radar operator on the resident ensemble: (hxf (nobs, ne) in dBZ or
        m/s, qc (nobs, ne), rk (nobs)).  var: the 3-D variables holding u, v, w, T,
        p, q, qc, qr, qi, qs, qg."""
        arrs = [np.ascontiguousarray(np.broadcast_to(a, np.shape(elm)), dtype=np.float64)
                for a in (elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z)]
        n = arrs[0].size
        c = _lib.RadarCfg()
        for q in range(11):
            c.var[q] = int(var[q])
        c.method, c.use_terminal_velocity = int(method), int(bool(use_terminal_velocity))
        c.min_ref_dbz, c.low_ref_shift, c.zmax = min_ref_dbz, low_ref_shift, zmax
        hxf = np.zeros((n, self.ne), order="F")
        qc = np.zeros((n, self.ne), dtype=np.int32, order="F")
        rk = np.zeros(n)
        check(_lib.lib().lekf_obsope_radar(self.h, ctypes.byref(c), n, *[dptr(a) for a in arrs],
                                           hxf.ctypes.data_as(_lib._dp),
                                           qc.ctypes.data_as(_lib._ip), dptr(rk)),
              "obsope_radar")
        return hxf, qc, rk

    def obs_departure(self, elm, dat, err, hxf, qc_member=None, **cfg):
        """SCALE's departure and QC step of set_letkf_obs on the GPU.

        elm, dat, err (nobs): obs as observed (reflectivity in mm^6 m^-3);
        hxf (nobs, ne): H(x) of each member (the radar operator's output);
        qc_member (nobs, ne): its qc.  cfg: the namelist options of
        _lib.SCALE_QC_DEFAULTS.  Returns (dat, err, dep, hdxf, qc): the obs in dBZ
        where it is a reflectivity, its error, y - mean H(x), the perturbations
        (H(x) where qc is not good) and the qc."""
        c = _qc_cfg(cfg)
        a = [np.ascontiguousarray(np.broadcast_to(v, np.shape(elm)), dtype=np.float64)
             for v in (elm, dat, err)]
        n = a[0].size
        hx = np.array(hxf, dtype=np.float64, order="F")
        if hx.shape != (n, self.ne):
            raise ValueError(f"hxf must be ({n}, {self.ne})")
        q = None
        if qc_member is not None:
            q = np.asfortranarray(qc_member, dtype=np.int32)
            if q.shape != (n, self.ne):
                raise ValueError(f"qc_member must be ({n}, {self.ne})")
        out = [np.zeros(n) for _ in range(3)]
        qc = np.zeros(n, dtype=np.int32)
        check(_lib.lib().lekf_obs_departure_scale(
            self.h, ctypes.byref(c), n, *[dptr(x) for x in a],
            q.ctypes.data_as(_lib._ip) if q is not None else None, hx.ctypes.data_as(_lib._dp),
            *[dptr(x) for x in out], iptr(qc)), "obs_departure_scale")
        return out[0], out[1], out[2], hx, qc

    def set_obs_radar(self, elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z, dat, err,
                      bucket, var=tuple(range(11)), method=3, use_terminal_velocity=False,
                      zmax=99.0e3, **qc):
        """Radar obs set up on the device: the radar operator, SCALE's departure and
        QC step and the obs index, with nothing copied back but the qc (returned).
        dat: the obs as observed (reflectivity in mm^6 m^-3); bucket: the (element,
        obs type) index of the cap; qc: the namelist options of
        _lib.SCALE_QC_DEFAULTS (min_radar_ref_dbz and low_ref_shift also go to the
        operator)."""
        arrs = [np.ascontiguousarray(np.broadcast_to(a, np.shape(elm)), dtype=np.float64)
                for a in (elm, lon, lat, z, ri, rj, radar_lon, radar_lat, radar_z, dat, err)]
        n = arrs[0].size
        b = np.ascontiguousarray(np.broadcast_to(bucket, (n,)), dtype=np.int32)
        qcfg = _qc_cfg(qc)
        c = _lib.RadarCfg()
        for q in range(11):
            c.var[q] = int(var[q])
        c.method, c.use_terminal_velocity = int(method), int(bool(use_terminal_velocity))
        c.min_ref_dbz, c.low_ref_shift, c.zmax = qcfg.min_radar_ref_dbz, qcfg.low_ref_shift, zmax
        out = np.zeros(n, dtype=np.int32)
        check(_lib.lib().lekf_set_obs_radar(self.h, ctypes.byref(c), ctypes.byref(qcfg), n,
                                            *[dptr(a) for a in arrs], iptr(b), iptr(out)),
              "set_obs_radar")
        return out

    def obs_index(self):
        """(perm, cum): input index of each binned obs and the cumulative obs count
        through each cell, shape (ny, nx) (cell (i, j) at [j-1, i-1])."""
        nb = ctypes.c_int(0)
        cum = np.zeros(self.nij, dtype=np.int32)
        check(_lib.lib().lekf_get_obs_index(self.h, None, iptr(cum), ctypes.byref(nb)),
              "get_obs_index")
        perm = np.zeros(max(nb.value, 1), dtype=np.int32)
        check(_lib.lib().lekf_get_obs_index(self.h, iptr(perm), None, None), "get_obs_index")
        return perm[:nb.value], cum.reshape(self.ny, self.nx)


def run_many(handles):
    """Batch the local solves of several handles into shared launches (waits)."""
    arr = (ctypes.c_void_p * len(handles))(*[h.h for h in handles])
    check(_lib.lib().lekf_das_run_many(arr, len(handles)), "das_run_many")


def run_batch(handles, stream):
    """Enqueue the batched cycle of several handles on one stream without waiting;
    the handles run on that stream afterwards."""
    arr = (ctypes.c_void_p * len(handles))(*[h.h for h in handles])
    check(_lib.lib().lekf_das_run_batch(arr, len(handles), stream), "das_run_batch")
