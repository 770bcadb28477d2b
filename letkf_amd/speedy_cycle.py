"""The SPEEDY LETKF programs on the GPU, reading and writing the reference's
files, so that the cycle script (speedy/letkf/run/run_cycle.sh:90-140) can
call them in place of its ``obsope`` and ``letkf020.m01`` executables.

``letkf``  -- the letkf program (speedy/letkf/letkf.f90:63-135):
    obs01MMM.dat (8-value records of each member, set_letkf_obs,
    speedy/letkf/letkf_obs.f90:76-160) and gs01MMM.grd in; analMMM.grd,
    gues_me/gues_sp/anal_me/anal_sp.grd out; the multiplicative inflation
    field infl_mul.grd read and written as das_letkf does
    (speedy/letkf/letkf_tools.f90:113-210).
``obsope`` -- the obsope program (speedy/letkf/obsope.f90) for every member
    at once: obsin.dat (6-value records) and gs01MMM.grd in, obs01MMM.dat out.
``cycle``  -- both in one pass: obsin.dat and the background go to the GPU
    once, H(x) is taken there, and the obs values pass through the same
    single-precision record conversions the obs01MMM.dat files apply, so the
    analysis equals that of ``obsope`` followed by ``letkf``.

Command line: ``python -m letkf_amd.speedy_cycle {letkf,obsope,cycle} --ne N``
in the working directory the script prepares.

The LETKF itself runs in :class:`letkf_amd.das.GlobalLetkf`; this module
only moves files.  The ensemble mean and spread files are formed on the host
in the reference's summation order (ensmean_grd, speedy/common/
common_speedy.f90:335-370; write_ensmspr_mpi, common_mpi_speedy.f90:460-515).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import das
from . import speedy_io as sio

# speedy/common/common_speedy.f90:61-120: T30 grid
LON = 360.0 / sio.NLON * np.arange(sio.NLON)
LAT = np.array([
    -87.159, -83.479, -79.777, -76.070, -72.362, -68.652, -64.942, -61.232, -57.521, -53.810,
    -50.099, -46.389, -42.678, -38.967, -35.256, -31.545, -27.833, -24.122, -20.411, -16.700,
    -12.989, -9.278, -5.567, -1.856, 1.856, 5.567, 9.278, 12.989, 16.700, 20.411, 24.122, 27.833,
    31.545, 35.256, 38.967, 42.678, 46.389, 50.099, 53.810, 57.521, 61.232, 64.942, 68.652,
    72.362, 76.070, 79.777, 83.479, 87.159])
SIG = np.array([0.95, 0.835, 0.685, 0.51, 0.34, 0.2, 0.08])

# speedy/letkf/letkf_obs.f90:23-50 and letkf_tools.f90:24
NSLOTS, NBSLOT = 1, 1
SIGMA_OBS, SIGMA_OBSV, GROSS_ERROR = 1000.0e3, 0.1, 10.0
COV_INFL_MUL = -1.01
IPS, IQ, Q_TOP_LEV = 0, 3, 5
NIJ = sio.NLON * sio.NLAT


class Files:
    """File names of one working directory (letkf.f90:27,91; letkf_obs.f90:78,93;
    common_mpi_speedy.f90:362,394,472; letkf_tools.f90:47)."""

    def __init__(self, workdir="."):
        self.dir = str(workdir)

    def path(self, name):
        return os.path.join(self.dir, name)

    def gues(self, m, slot=NBSLOT):
        return self.path(f"gs{slot:02d}{m + 1:03d}.grd")

    def anal(self, m):
        return self.path(f"anal{m + 1:03d}.grd")

    def obs(self, m, slot=1):
        return self.path(f"obs{slot:02d}{m + 1:03d}.dat")

    def mean(self, kind):
        return self.path(f"{kind}_me.grd")

    def sprd(self, kind):
        return self.path(f"{kind}_sp.grd")

    @property
    def infl(self):
        return self.path("infl_mul.grd")

    @property
    def obsin(self):
        return self.path("obsin.dat")


# ---------------------------------------------------------------------------
# grid <-> handle layout
# ---------------------------------------------------------------------------
def read_ensemble(files, ne, big_endian=True, slot=NBSLOT):
    """gsTTMMM.grd of every member as (gues3d (nij, nlev, ne, nv3d), gues2d (nij, ne, nv2d))."""
    g3 = np.zeros((NIJ, sio.NLEV, ne, sio.NV3D), order="F")
    g2 = np.zeros((NIJ, ne, sio.NV2D), order="F")
    for m in range(ne):
        v3, v2 = sio.read_grd(files.gues(m, slot), big_endian=big_endian)
        g3[:, :, m, :] = v3.reshape(NIJ, sio.NLEV, sio.NV3D, order="F")
        g2[:, m, :] = v2.reshape(NIJ, sio.NV2D, order="F")
    return g3, g2


def _grid3(a):
    """(nij, nlev, nv) -> (nlon, nlat, nlev, nv)."""
    return a.reshape(sio.NLON, sio.NLAT, a.shape[1], a.shape[2], order="F")


def _grid2(a):
    return a.reshape(sio.NLON, sio.NLAT, a.shape[1], order="F")


def write_ensemble(files, a3, a2, big_endian=True):
    for m in range(a3.shape[2]):
        sio.write_grd(files.anal(m), _grid3(a3[:, :, m, :]), _grid2(a2[:, m, :]), big_endian)


def mean_sprd(x3, x2):
    """Ensemble mean and spread sqrt(sum (x - mean)^2 / (ne - 1)), summed over
    members in order."""
    ne = x3.shape[2]
    m3 = x3[:, :, 0, :].copy()
    m2 = x2[:, 0, :].copy()
    for m in range(1, ne):
        m3 += x3[:, :, m, :]
        m2 += x2[:, m, :]
    m3 /= float(ne)
    m2 /= float(ne)
    s3 = (x3[:, :, 0, :] - m3) ** 2
    s2 = (x2[:, 0, :] - m2) ** 2
    for m in range(1, ne):
        s3 += (x3[:, :, m, :] - m3) ** 2
        s2 += (x2[:, m, :] - m2) ** 2
    return m3, m2, np.sqrt(s3 / float(ne - 1)), np.sqrt(s2 / float(ne - 1))


def write_mean_sprd(files, kind, x3, x2, big_endian=True):
    m3, m2, s3, s2 = mean_sprd(x3, x2)
    sio.write_grd(files.mean(kind), _grid3(m3), _grid2(m2), big_endian)
    sio.write_grd(files.sprd(kind), _grid3(s3), _grid2(s2), big_endian)


# ---------------------------------------------------------------------------
# observations
# ---------------------------------------------------------------------------
def read_member_obs(files, ne, big_endian=True, nslots=NSLOTS):
    """set_letkf_obs's reading (letkf_obs.f90:89-160): the records of every
    slot, H(x) of each member (nobs, ne), qc = the minimum over members."""
    recs, hx, qc = [], [], []
    for slot in range(1, nslots + 1):
        n = sio.count_obs(files.obs(0, slot), 8)
        if n == 0:
            continue
        h = np.zeros((n, ne), order="F")
        q = np.zeros((n, ne), dtype=np.int32)
        for m in range(ne):
            o = sio.read_obs(files.obs(m, slot), nval=8, big_endian=big_endian)
            if o["elem"].size != n:
                raise ValueError(f"{files.obs(m, slot)}: {o['elem'].size} records, member 1 has {n}")
            h[:, m] = o["ohx"]
            q[:, m] = o["oqc"]
        recs.append(o)
        hx.append(h)
        qc.append(q.min(axis=1))
    if not recs:
        return None
    obs = {k: np.concatenate([r[k] for r in recs]) for k in sio.OBS_KEYS}
    obs["hxf"] = np.asfortranarray(np.concatenate(hx, axis=0))
    obs["qc"] = np.concatenate(qc).astype(np.int32)
    return obs


def _obs_through_files(obs_in, hxf, qc):
    """The values letkf reads back from obs01MMM.dat written by obsope: each
    record encoded by write_obs2 and decoded by read_obs2, member by member."""
    out = None
    ne = hxf.shape[1]
    h = np.zeros_like(hxf, order="F")
    for m in range(ne):
        rec = dict(obs_in, ohx=hxf[:, m], oqc=qc)
        o = sio.decode_obs(sio.encode_obs2(rec))
        h[:, m] = o["ohx"]
        out = o
    out = {k: out[k] for k in sio.OBS_KEYS}
    out["hxf"] = h
    out["qc"] = qc.astype(np.int32)
    return out


# ---------------------------------------------------------------------------
# inflation field (letkf_tools.f90:113-140, 203-210)
# ---------------------------------------------------------------------------
def initial_infl(files, cov_infl_mul, big_endian=True):
    """(infl3d (nij, nlev), infl2d (nij, nv2d)) entering das_letkf.  All 3-D
    variables take the field of the first one, which the reference uses for the
    solve (variable localization factors all equal, letkf_tools.f90:28-36,82-88)."""
    if cov_infl_mul > 0.0:
        i3 = np.full((NIJ, sio.NLEV), cov_infl_mul, order="F")
        i3[:, sio.NLEV - 1] = 1.01
        return i3, np.full((NIJ, sio.NV2D), cov_infl_mul, order="F")
    if os.path.exists(files.infl):
        v3, v2 = sio.read_grd(files.infl, big_endian=big_endian)
        return (np.asfortranarray(v3[:, :, :, 0].reshape(NIJ, sio.NLEV, order="F"), dtype=np.float64),
                np.asfortranarray(v2.reshape(NIJ, sio.NV2D, order="F"), dtype=np.float64))
    return (np.full((NIJ, sio.NLEV), -cov_infl_mul, order="F"),
            np.full((NIJ, sio.NV2D), -cov_infl_mul, order="F"))


def write_infl(files, i3, i2_in, big_endian=True):
    """infl_mul.grd after the analysis: every 3-D variable carries the updated
    field; every 2-D variable carries the first 2-D variable's input field (the
    reference copies work2d(ij, 1) into each 2-D slot, letkf_tools.f90:178-181)."""
    v3 = np.repeat(i3.reshape(NIJ, sio.NLEV, 1, order="F"), sio.NV3D, axis=2)
    v2 = np.repeat(i2_in[:, :1], sio.NV2D, axis=1)
    sio.write_grd(files.infl, _grid3(v3), _grid2(v2), big_endian)


# ---------------------------------------------------------------------------
# programs
# ---------------------------------------------------------------------------
def _handle(ne, prec, sigma_obs, sigma_obsv, gross_error):
    return das.GlobalLetkf(lon=LON, lat=LAT, sig=SIG, ne=ne, nv3d=sio.NV3D, nv2d=sio.NV2D,
                           ips=IPS, iq=IQ, q_top_lev=Q_TOP_LEV, sigma_obs=sigma_obs,
                           sigma_obsv=sigma_obsv, gross_error=gross_error, prec=prec,
                           infl_mode="adaptive", infl_value=1.0)


class _Timer:
    def __init__(self, log):
        self.log = log
        self.t = time.perf_counter()
        self.times = {}

    def __call__(self, name):
        t = time.perf_counter()
        self.times[name] = t - self.t
        if self.log:
            self.log(f"### TIMER({name}): {t - self.t:10.3f}")
        self.t = t


def _analyse(files, h, g3, g2, obs, cov_infl_mul, big_endian, tm):
    """das_letkf and the output files; obs None or with no accepted record: no
    analysis (anal = gues) and no inflation file (letkf_tools.f90:73-79)."""
    nacc = 0
    if obs is not None and obs["elem"].size:
        h.set_obs(obs["elem"], obs["rlon"], obs["rlat"], obs["rlev"], obs["odat"], obs["oerr"],
                  obs["hxf"], obs["qc"])
        nacc = h.obs_index()[0].size
    if nacc == 0:
        a3, a2 = g3, g2
    else:
        i3, i2 = initial_infl(files, cov_infl_mul, big_endian)
        h.set_infl(i3, i2)
        h.run()
        a3, a2 = h.analysis()
        if cov_infl_mul < 0.0:
            write_infl(files, h.infl()[0], i2, big_endian)
    tm("DAS_LETKF")
    write_ensemble(files, a3, a2, big_endian)
    write_mean_sprd(files, "anal", a3, a2, big_endian)
    tm("WRITE_ANAL")
    return nacc


def letkf(workdir=".", *, ne, prec="f64", cov_infl_mul=COV_INFL_MUL, sigma_obs=SIGMA_OBS,
          sigma_obsv=SIGMA_OBSV, gross_error=GROSS_ERROR, big_endian=True, log=None):
    """The letkf program.  Returns {"nobs": records read, "nobs_assim": records
    accepted, "timers": {...}}."""
    files = Files(workdir)
    tm = _Timer(log)
    h = _handle(ne, prec, sigma_obs, sigma_obsv, gross_error)
    try:
        tm("INITIALIZE")
        obs = read_member_obs(files, ne, big_endian)
        tm("READ_OBS")
        g3, g2 = read_ensemble(files, ne, big_endian)
        write_mean_sprd(files, "gues", g3, g2, big_endian)
        h.set_ensemble(g3, g2)
        tm("READ_GUES")
        nacc = _analyse(files, h, g3, g2, obs, cov_infl_mul, big_endian, tm)
    finally:
        h.close()
    return dict(nobs=0 if obs is None else int(obs["elem"].size), nobs_assim=int(nacc),
                timers=tm.times)


def _obsope_on(h, obs_in):
    hxf, qc, _, _, _ = h.obs_operator(obs_in["elem"], obs_in["rlon"], obs_in["rlat"],
                                      obs_in["rlev"], obs_in["odat"])
    return hxf, qc


def obsope(workdir=".", *, ne, big_endian=True, log=None):
    """The obsope program for members 1..ne: obs01MMM.dat from obsin.dat and
    gs01MMM.grd.  H(x) of rejected records is 0; the qc written to every
    member's file is 1 where all members accept the record, which is what
    set_letkf_obs reduces the per-member flags to."""
    files = Files(workdir)
    obs_in = sio.read_obs(files.obsin, nval=6, big_endian=big_endian)
    g3, g2 = read_ensemble(files, ne, big_endian)
    h = _handle(ne, "f64", SIGMA_OBS, SIGMA_OBSV, GROSS_ERROR)
    try:
        h.set_ensemble(g3, g2)
        hxf, qc = _obsope_on(h, obs_in)
    finally:
        h.close()
    for m in range(ne):
        sio.write_obs2(files.obs(m), dict(obs_in, ohx=hxf[:, m], oqc=qc), big_endian)
    if log:
        log(f"{obs_in['elem'].size} records, {int(qc.sum())} accepted by every member")
    return dict(nobs=int(obs_in["elem"].size), accepted=int(qc.sum()))


def cycle(workdir=".", *, ne, prec="f64", cov_infl_mul=COV_INFL_MUL, sigma_obs=SIGMA_OBS,
          sigma_obsv=SIGMA_OBSV, gross_error=GROSS_ERROR, big_endian=True, log=None):
    """obsope + letkf with one upload of the background and no obs01MMM.dat files."""
    files = Files(workdir)
    tm = _Timer(log)
    h = _handle(ne, prec, sigma_obs, sigma_obsv, gross_error)
    try:
        tm("INITIALIZE")
        g3, g2 = read_ensemble(files, ne, big_endian)
        write_mean_sprd(files, "gues", g3, g2, big_endian)
        h.set_ensemble(g3, g2)
        tm("READ_GUES")
        obs = None
        if sio.count_obs(files.obsin, 6):
            obs_in = sio.read_obs(files.obsin, nval=6, big_endian=big_endian)
            hxf, qc = _obsope_on(h, obs_in)
            obs = _obs_through_files(obs_in, hxf, qc)
        tm("OBSOPE")
        nacc = _analyse(files, h, g3, g2, obs, cov_infl_mul, big_endian, tm)
    finally:
        h.close()
    return dict(nobs=0 if obs is None else int(obs["elem"].size), nobs_assim=int(nacc),
                timers=tm.times)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m letkf_amd.speedy_cycle",
                                 description="SPEEDY obsope / letkf on the GPU with the reference's files")
    ap.add_argument("program", choices=("letkf", "obsope", "cycle"))
    ap.add_argument("--ne", type=int, required=True, help="ensemble size (nbv)")
    ap.add_argument("--dir", default=".", help="working directory")
    ap.add_argument("--prec", default="f64", choices=("f64", "f32", "bf16x3"))
    ap.add_argument("--cov-infl-mul", type=float, default=COV_INFL_MUL)
    ap.add_argument("--sigma-obs", type=float, default=SIGMA_OBS)
    ap.add_argument("--sigma-obsv", type=float, default=SIGMA_OBSV)
    ap.add_argument("--gross-error", type=float, default=GROSS_ERROR)
    ap.add_argument("--little-endian", action="store_true",
                    help="files in little-endian byte order (the reference writes big-endian)")
    a = ap.parse_args(argv)
    be = not a.little_endian
    if a.program == "obsope":
        r = obsope(a.dir, ne=a.ne, big_endian=be, log=print)
    else:
        run = letkf if a.program == "letkf" else cycle
        r = run(a.dir, ne=a.ne, prec=a.prec, cov_infl_mul=a.cov_infl_mul, sigma_obs=a.sigma_obs,
                sigma_obsv=a.sigma_obsv, gross_error=a.gross_error, big_endian=be, log=print)
        print(f"{r['nobs']} OBSERVATIONS INPUT, {r['nobs_assim']} TO BE ASSIMILATED")
    return 0


if __name__ == "__main__":
    sys.exit(main())
