"""Write the golden fixtures under tests/golden/ from the reference Fortran.

Run in a container where /root/reference exists (so oracle/_ref is built):

    make -C oracle && python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  core_cases.npz   letkf_core inputs/outputs of the reference Fortran
                   (common/common_letkf.f90:50-243) for k in {3..96}, p in {0..1000}
                   and every OPTIONAL-argument combination;
  eigen_cases.npz  mtx_eigen (common/common_mtx.f90:41-99) inputs/outputs;
  l96_step.npz     one Lorenz-96 analysis step (config 1): the L96 selection rule
                   (lorenz96/letkf/letkf.f90:216-254) and the reference letkf_core +
                   apply (:255-262) driven over the 40 points;
  global_patch.npz a reduced global sigma-level problem: selection by the oracle's
                   obs_local restatement, local solves + apply by the reference
                   Fortran (patch_by_reference: ref_solve_apply over the lists), for
                   the das-level parity test.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from letkf_amd import synth  # noqa: E402

CORE = [(0, 20), (1, 20), (21, 20), (200, 40), (1000, 50), (200, 64), (150, 96), (7, 3), (60, 31)]
OPTS = [(0, 0, 0, None), (1, 0, 0, None), (0, 1, 0, None), (1, 1, 1, 1.1), (0, 0, 0, 1.3)]


def core_cases():
    """Inputs stored once per (p, k); outputs per OPTIONAL-argument combination.
    With rdiag_wloc the input rdiag is divided by rloc first."""
    out = {}
    for c, (p, k) in enumerate(CORE):
        h, rd, rl, dp = synth.local_problem(p, k, seed=31 * p + k)
        out[f"c{c}_h"] = h
        out[f"c{c}_rd"] = rd
        out[f"c{c}_rl"] = rl
        out[f"c{c}_dp"] = dp
        for o, (tm, pa, wl, mf) in enumerate(OPTS):
            r = oracle.letkf_core(h, rd / rl if wl else rd, rl, dp, 1.03, transm=bool(tm),
                                  pao=bool(pa), rdiag_wloc=bool(wl), minfl=mf, impl="ref")
            key = f"c{c}_o{o}"
            out[key + "_trans"] = r["trans"]
            out[key + "_parm"] = np.array([r["parm_infl"]])
            if tm:
                out[key + "_transm"] = r["transm"]
            if pa:
                out[key + "_pao"] = r["pao"]
    out["cases"] = np.array(CORE)
    out["opts"] = np.array([[tm, pa, wl, -1.0 if mf is None else mf] for tm, pa, wl, mf in OPTS])
    out["parm0"] = np.array([1.03])
    np.savez_compressed(os.path.join(HERE, "core_cases.npz"), **out)


def eigen_cases():
    out = {}
    for i, n in enumerate([2, 5, 20, 40]):
        rng = np.random.default_rng(400 + n)
        a = rng.standard_normal((n, n))
        a = a @ a.T + n * np.eye(n)
        ev, vec, nr, _ = oracle.mtx_eigen(a, impl="ref")
        out[f"a{i}"] = a
        out[f"ev{i}"] = ev
        out[f"vec{i}"] = vec
    np.savez_compressed(os.path.join(HERE, "eigen_cases.npz"), **out)


def l96_step():
    c = synth.l96()
    nx, ne = c["nx"], c["ne"]
    idx, rho, off = [], [], [0]
    for ix in range(nx):
        i, r = oracle.obs_local_l96(nx, c["gidx"], ix, c["xlocal"])
        idx.append(i)
        rho.append(r)
        off.append(off[-1] + i.size)
    hx = c["hxf"]
    hm = hx.mean(axis=1)
    hdxf = hx - hm[:, None]
    dep = c["y"] - hm
    xa, infl = oracle.ref_solve_apply(np.array(off), np.concatenate(idx), np.concatenate(rho), hdxf,
                                      c["err"] ** 2, dep, c["xb"][:, None, :],
                                      np.full(nx, c["infl"]), fixed_infl=c["infl"])
    np.savez_compressed(os.path.join(HERE, "l96_step.npz"), xb=c["xb"], gidx=c["gidx"], y=c["y"],
                        err=c["err"], hxf=c["hxf"], xa=xa[:, 0, :], truth=c["truth"],
                        loc_off=np.array(off), loc_idx=np.concatenate(idx),
                        loc_rho=np.concatenate(rho))


def patch_by_reference(p, r):
    """The patch's analysis with every local solve and apply done by the reference
    Fortran letkf_core (oracle/_ref, ref_solve_apply) over the oracle's obs_local
    lists: 3-D rows at every level, the 2-D rows with the level-1 transform
    (speedy/letkf/letkf_tools.f90:178-197), adaptive inflation.  Returns (anal3d,
    anal2d, infl3d) in the layouts of das_global_problem."""
    s = r["sobs"]
    nij, nlev, ne = p["nij"], p["nlev"], p["ne"]
    lon1 = np.tile(p["lon"], p["nlat"])
    lat1 = np.repeat(p["lat"], p["nlon"])
    dz = p["sigma_obs"] * np.sqrt(10.0 / 3.0) * 2.0
    dzv = p["sigma_obsv"] * np.sqrt(10.0 / 3.0) * 2.0
    dlat = dz / oracle.PI_REF / oracle.RE_REF * 180.0
    g2 = p["gues2d"][:, :, 0]
    psm = g2[:, 0].copy()
    for m in range(1, ne):
        psm = psm + g2[:, m]
    psm = psm / ne
    a3 = np.zeros_like(p["gues3d"])
    a2 = np.zeros_like(p["gues2d"])
    i3 = np.zeros((nij, nlev))
    for lev in range(nlev):
        idx_all, rho_all, off, xb = [], [], [0], []
        for ij in range(nij):
            dlon = dlat / np.cos(oracle.PI_REF * lat1[ij] / 180.0)
            idx, rho, _ = oracle.obs_local_global(lon1[ij], lat1[ij], dlon, dlat,
                                                  np.log(psm[ij] * p["sig"][lev]), lev + 1,
                                                  p["lon"], p["lat"], r["nobsgrd"], s, dz, dzv,
                                                  p["sigma_obs"], p["sigma_obsv"])
            idx_all.append(idx)
            rho_all.append(rho)
            off.append(off[-1] + idx.size)
            rows = [p["gues3d"][ij, lev, :, v] for v in range(4)]
            if lev == 0:
                rows.append(p["gues2d"][ij, :, 0])
            xb.append(np.stack(rows + [np.zeros(ne)] * (5 - len(rows))))
        xa, io = oracle.ref_solve_apply(np.array(off), np.concatenate(idx_all),
                                        np.concatenate(rho_all), np.ascontiguousarray(s["hdxf"].T),
                                        s["err"] ** 2, s["dep"], np.stack(xb),
                                        np.full(nij, p["infl"]))
        a3[:, lev] = xa[:, :4].transpose(0, 2, 1)
        if lev == 0:
            a2[:, :, 0] = xa[:, 4]
        i3[:, lev] = io
    return a3, a2, i3


def global_patch():
    p = synth.global_problem(nlon=24, lat=synth.SPEEDY_LAT[2::4], sig=synth.SPEEDY_SIG[::2], ne=10,
                             nstation=60, seed=synth.SEED_BASE + 90, nmode=16)
    p.update(sigma_obs=1500.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.01, ips=0)
    o = p["obs"]
    r = oracle.das_global_problem(p)
    a3, a2, i3 = patch_by_reference(p, r)
    # the oracle's das restatement gives the reference's analysis to rounding
    sc = np.abs(a3 - a3.mean(axis=2, keepdims=True)).max()
    assert np.abs(r["anal3d"] - a3).max() <= 1e-12 * sc
    r.update(anal3d=a3, anal2d=a2, infl3d=i3)
    np.savez_compressed(os.path.join(HERE, "global_patch.npz"), lon=p["lon"], lat=p["lat"],
                        sig=p["sig"], gues3d=p["gues3d"], gues2d=p["gues2d"], elm=o["elm"],
                        olon=o["lon"], olat=o["lat"], olev=o["lev"], dat=o["dat"], err=o["err"],
                        hxf=o["hxf"], anal3d=r["anal3d"], anal2d=r["anal2d"], infl3d=r["infl3d"],
                        nobsl=r["nobsl"], perm=r["perm"], nobsgrd=r["nobsgrd"],
                        sigma=np.array([p["sigma_obs"], p["sigma_obsv"], p["gross_error"], 1.01]))


if __name__ == "__main__":
    if not oracle.ref_available():
        raise SystemExit("build oracle/_ref first (make -C oracle)")
    core_cases()
    eigen_cases()
    l96_step()
    global_patch()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
