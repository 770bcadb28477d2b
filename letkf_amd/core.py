"""Host-side mirror of the reference's ``letkf_core`` (common/common_letkf.f90:50-243).

``letkf_core`` keeps the reference's argument order and meaning:

    trans, parm_infl[, transm][, pao] = letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep,
                                                    parm_infl, transm=..., pao=...,
                                                    rdiag_wloc=..., minfl=...)

``hdxb`` is the (nobs, ne) obs-space perturbation array of which the first
``nobsl`` rows are used.  The solve runs on the GPU through ``lekf_core_f64``;
a nonzero status raises :class:`LetkfError` where the reference would
``STOP`` (2: eigensolver failure, 99: sizes).  The rank-deficient case the
reference turns into NaN (common/common_mtx.f90:66-74) returns finite values
and raises only when ``strict=True``.

``letkf_core_batch`` solves many local problems given in device memory
(torch tensors) in one launch through ``lekf_core_batch``.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import LekfJob, check, dptr


def letkf_core(ne, nobs, nobsl, hdxb, rdiag, rloc, dep, parm_infl, *, transm=False, pao=False,
               rdiag_wloc=False, minfl=None, strict=False):
    L = _lib.lib()
    h = np.asfortranarray(np.asarray(hdxb, dtype=np.float64).reshape(nobs, ne))
    rd = np.ascontiguousarray(rdiag, dtype=np.float64)
    rl = np.ascontiguousarray(rloc, dtype=np.float64)
    dp = np.ascontiguousarray(dep, dtype=np.float64)
    if nobsl > nobs or rd.size < nobsl or rl.size < nobsl or dp.size < nobsl:
        raise _lib.LetkfError(_lib.LEKF_ERR_SIZE, "letkf_core")
    parm = np.array([parm_infl], dtype=np.float64)
    tr = np.zeros((ne, ne), dtype=np.float64, order="F")
    tm = np.zeros(ne) if transm else None
    pa = np.zeros((ne, ne), order="F") if pao else None
    mf = np.array([minfl], dtype=np.float64) if minfl is not None else None
    st = L.lekf_core_f64(ne, nobs, nobsl, h.ctypes.data_as(_lib._dp), dptr(rd), dptr(rl),
                         dptr(dp), dptr(parm), tr.ctypes.data_as(_lib._dp),
                         dptr(tm) if tm is not None else None,
                         pa.ctypes.data_as(_lib._dp) if pa is not None else None,
                         int(bool(rdiag_wloc)), dptr(mf) if mf is not None else None)
    if st == _lib.LEKF_WARN_RANKDEF and not strict:
        st = _lib.LEKF_OK
    check(st, "letkf_core")
    out = [np.ascontiguousarray(tr), float(parm[0])]
    if transm:
        out.append(tm)
    if pao:
        out.append(np.ascontiguousarray(pa))
    return tuple(out)


def _ptr(t):
    return None if t is None else t.data_ptr()


def make_job(*, k, npt, hdxf, rdiag, dep, loc_off, loc_idx, loc_rho, loc_cnt=None, loc_rdiag=None,
             infl_in=None, infl_out=None, infl_fixed=1.0, trans=None, transm=None, pao=None,
             status=None, hdxf32=None, y_so=None, y_sm=1, wloc=False, minfl=None,
             fixed_infl=False, apply=None):
    """Build an ``lekf_job`` from torch device tensors (row-major obs table by default)."""
    j = LekfJob()
    j.npt = npt
    j.k = k
    flags = 0
    if wloc:
        flags |= _lib.LEKF_F_WLOC
    if minfl is not None:
        flags |= _lib.LEKF_F_MINFL
        j.minfl = float(minfl)
    if fixed_infl:
        flags |= _lib.LEKF_F_FIXED_INFL
    if transm is not None:
        flags |= _lib.LEKF_F_TRANSM
    j.infl_fixed = float(infl_fixed)
    j.hdxf = _ptr(hdxf)
    j.y_so = k if y_so is None else y_so
    j.y_sm = y_sm
    j.hdxf32 = _ptr(hdxf32)
    j.rdiag = _ptr(rdiag)
    j.dep = _ptr(dep)
    j.loc_off = _ptr(loc_off)
    j.loc_cnt = _ptr(loc_cnt)
    j.loc_idx = _ptr(loc_idx)
    j.loc_rho = _ptr(loc_rho)
    j.loc_rdiag = _ptr(loc_rdiag)
    j.infl_in = _ptr(infl_in)
    j.infl_out = _ptr(infl_out)
    j.trans = _ptr(trans)
    j.transm = _ptr(transm)
    j.pao = _ptr(pao)
    j.status = _ptr(status)
    if apply is not None:
        flags |= _lib.LEKF_F_APPLY
        j.nij = apply["nij"]
        j.nv3d = apply["nv3d"]
        j.nv2d = apply.get("nv2d", 0)
        j.x3b = _ptr(apply["x3b"])
        j.x3a = _ptr(apply["x3a"])
        j.x2b = _ptr(apply.get("x2b"))
        j.x2a = _ptr(apply.get("x2a"))
        j.upd_mask = _ptr(apply.get("upd_mask"))
    j.flags = flags
    # the descriptor holds raw device pointers: keep the buffers alive with it
    j._keep = [hdxf, rdiag, dep, loc_off, loc_idx, loc_rho, loc_cnt, loc_rdiag, infl_in,
               infl_out, trans, transm, pao, status, hdxf32, apply]
    return j


def letkf_core_batch(jobs, prec="f64", stream=None):
    """Launch the batched solve over a list of ``lekf_job`` (asynchronous).

    The kernels read the jobs' buffers after this returns: keep the job
    objects (which hold their buffers) alive until the stream is synchronised.
    """
    L = _lib.lib()
    arr = (LekfJob * len(jobs))(*jobs)
    st = L.lekf_core_batch(arr, len(jobs), _lib.PREC[prec] if isinstance(prec, str) else prec,
                           stream)
    check(st, "lekf_core_batch")
