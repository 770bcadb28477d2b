#!/usr/bin/env python3
"""Benchmark: LETKF local analyses/s on one MI355X (SPEEDY-sized global config).

Workload (one "step" = one analysis cycle of das_letkf over resident inputs):
T30 96x48 lon-lat grid, 7 sigma levels, u,v,T,q + ps, k=20 members, ~12k
obs (415 stations x 7 levels x 4 + ps), sigma_h 500 km, sigma_v 0.4 ln p,
adaptive inflation.  32,256 local analyses per cycle: for each one the
selection (obs_local), the local solve (letkf_core) and the apply.

Prints ONE JSON line.  value = plain-FP64 local analyses/s (inputs resident in
HBM); mixed-precision forms, p50/p99 cycle latency, the roofline of the
dominant kernel (the local-solve kernel, FP64 MFMA bound) and the reference
Fortran letkf_core timed on this host's cores are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "local analyses/s + p50/p99 cycle latency, mixed vs FP64 GPU at stated err"
FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (matrix and vector)
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0


def flops_per_point(p, k, nv):
    """Algorithmic FLOPs of one local analysis (minimal association):
    2pk^2 (Y^T R^-1 Y) + 3pk (R^-1 Y, Y^T R^-1 d) + 13k^3 (eigensolve 9k^3 + Pa, Wa 4k^3)
    + 3k^2 (w-bar, T) + 2 nv k^2 (apply); a point without obs costs 2 nv k^2 + k^2."""
    p = np.asarray(p, dtype=np.float64)
    full = 2 * p * k * k + 3 * p * k + 13 * k ** 3 + 3 * k * k + 2 * nv * k * k
    empty = 2 * nv * k * k + k * k
    return np.where(p > 0, full, empty)


def measured_traffic(name):
    """HBM bytes per launch of a kernel/workload from its rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE scaled by a plain copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q))


def make_handle(p, prec):
    from letkf_amd import das
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                        ips=0, iq=p["iq"], q_top_lev=p["q_top_lev"], sigma_obs=p["sigma_obs"],
                        sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"], prec=prec,
                        infl_value=p["infl"])
    h.set_ensemble(p["gues3d"], p["gues2d"])
    o = p["obs"]
    t0 = time.perf_counter()
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    obs_ms = (time.perf_counter() - t0) * 1e3
    return h, obs_ms


def time_cycles(h, steps, warmup):
    """Warm up, then time `steps` cycles; HIP events on the handle's stream."""
    from letkf_amd.device import Event
    st = h.stream()
    for _ in range(warmup):
        h.run()
    h.sync()
    ev = [(Event(), Event(), Event()) for _ in range(steps)]
    lat = []
    t_start = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        ev[i][0].record(st)
        h.select()
        ev[i][1].record(st)
        h.solve()
        ev[i][2].record(st)
        h.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
    wall_ms = (time.perf_counter() - t_start) * 1e3
    sel = [Event.elapsed_ms(a, b) for a, b, _ in ev]
    sol = [Event.elapsed_ms(b, c) for _, b, c in ev]
    return dict(wall_ms=wall_ms, lat=lat, select_ms=sel, solve_ms=sol)


def cpu_baseline(p, ref, max_threads=16):
    """Reference Fortran letkf_core + apply over the full cycle's local problems.

    Local lists come from the oracle's obs_local restatement (identical to the
    GPU selection).  OpenMP over points (scale/letkf/letkf_tools.f90:175)."""
    import oracle
    if not oracle.ref_available():
        return None
    threads = max(1, min(max_threads, os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    s = ref["sobs"]
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
    idx_all, rho_all, off = [], [], [0]
    for lev in range(nlev):
        for ij in range(nij):
            dlon = dlat / np.cos(oracle.PI_REF * lat1[ij] / 180.0)
            idx, rho, _ = oracle.obs_local_global(lon1[ij], lat1[ij], dlon, dlat,
                                                  np.log(psm[ij] * p["sig"][lev]), lev + 1,
                                                  p["lon"], p["lat"], ref["nobsgrd"], s, dz, dzv,
                                                  p["sigma_obs"], p["sigma_obsv"])
            idx_all.append(idx)
            rho_all.append(rho)
            off.append(off[-1] + idx.size)
    off = np.array(off)
    idx_all = np.concatenate(idx_all)
    rho_all = np.concatenate(rho_all)
    hdxf = np.ascontiguousarray(s["hdxf"].T)
    g3 = p["gues3d"]
    xb = np.ascontiguousarray(g3.transpose(1, 0, 3, 2)).reshape(nlev * nij, 4, ne)
    infl = np.full(nlev * nij, p["infl"])
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.ref_solve_apply(off, idx_all, rho_all, hdxf, s["err"] ** 2, s["dep"], xb, infl)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    npt = nlev * nij
    return {"value": npt / t, "unit": "local analyses/s", "cores": oracle.ref_threads(),
            "kind": "reference",
            "sample": (f"full SPEEDY-sized cycle: {npt} local analyses (reference Fortran "
                       f"letkf_core + EISPACK + reference DGEMM, amdflang -O3, OpenMP over points) "
                       f"+ apply to 4 variables; median of 3 runs, {t:.3f} s each; selection "
                       f"(oracle restatement) not timed")}


def radar_bench(sub, steps, precs=("f64", "f32", "bf16x3"), cpu_sample=0):
    """Regional radar config on a grid of sub = (nx, ny, nlev) columns:
    k=50, u,v,w,T,qv,qr, ~60% of columns observed at every level, sigma_h = sigma_v = 2 km,
    nobs_max = 1000 per bucket.  Each form's analysis is compared with the FP64 GPU
    analysis (itself held to 1e-10 of the reference by the tests)."""
    from letkf_amd import das, synth
    from letkf_amd.device import Event
    t0 = time.perf_counter()
    p = synth.radar(sub=sub)
    gen_s = time.perf_counter() - t0
    npt = p["nx"] * p["ny"] * p["nlev"]
    spread = p["gues3d"].std(axis=2, ddof=1)
    res, ref = {}, None
    for prec in precs:
        h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                              ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"],
                              sigma_v=p["sigma_v"], nobs_max=p["nobs_max"], infl_value=1.0,
                              prec=prec)
        h.set_ensemble(p["gues3d"])
        t1 = time.perf_counter()
        h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
        h.sync()
        obs_ms = (time.perf_counter() - t1) * 1e3
        st = h.stream()
        h.run()
        h.sync()
        lat = []
        e0, e1 = Event(), Event()
        for _ in range(steps):
            t1 = time.perf_counter()
            e0.record(st)
            h.run()
            e1.record(st)
            h.sync()
            lat.append((time.perf_counter() - t1) * 1e3)
        a3, _ = h.analysis()
        nl = h.nobsl()
        h.close()
        r = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3), "cycle_ms": float(np.mean(lat)),
             "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
             "obs_setup_ms": obs_ms, "mean_nobsl": float(nl.mean())}
        if ref is None:
            ref = a3
        else:
            r["err_vs_f64_over_bg_spread"] = float((np.abs(a3 - ref).max(axis=2) / spread).max())
        res[prec] = r
        print(f"[bench] radar {sub} {prec}: {r['cycle_ms']:.0f} ms/cycle", file=sys.stderr,
              flush=True)
    if "f64" in res:
        for prec in res:
            res[prec]["speedup_vs_f64"] = res[prec]["local_analyses_per_s"] / res["f64"]["local_analyses_per_s"]
    out = {"grid": list(sub), "k": p["ne"], "local_analyses_per_cycle": npt,
           "nobs": int(p["ri"].size), "nobs_max_per_bucket": p["nobs_max"], "gen_s": gen_s,
           "forms": res}
    if cpu_sample:
        out["cpu_baseline"] = radar_cpu_sample(p, cpu_sample)
    return out


def radar_cpu_sample(p, npts):
    """Reference letkf_core + apply on a seeded sample of radar points (their capped
    local lists from the oracle's SCALE selection restatement), OpenMP over points."""
    import oracle
    if not oracle.ref_available():
        return None
    nx, ny, nlev, nij = p["nx"], p["ny"], p["nlev"], p["nx"] * p["ny"]
    perm, grd = oracle.obs_bins_regional(p["ri"], p["rj"], nx, ny)
    so = {k: np.asarray(p[k])[perm] for k in ("ri", "rj", "z", "err", "bucket")}
    rng = np.random.default_rng(7)
    pts = np.sort(rng.choice(nij * nlev, npts, replace=False))
    idx_all, rho_all, off = [], [], [0]
    for b in pts:
        lev, ij = divmod(int(b), nij)
        idx, rl, _ = oracle.obs_local_regional(float(ij % nx + 1), float(ij // nx + 1),
                                               p["zlev"][lev], nx, ny, grd, so, p["dx"], p["dy"],
                                               p["sigma_h"], p["sigma_v"], p["nobs_max"], 2)
        idx_all.append(idx)
        rho_all.append(rl)
        off.append(off[-1] + idx.size)
    g3 = p["gues3d"]
    xb = np.ascontiguousarray(g3.transpose(1, 0, 3, 2)).reshape(nlev * nij, p["nv3d"], p["ne"])[pts]
    hd = np.asarray(p["hdxf"])[perm]
    t0 = time.perf_counter()
    oracle.ref_solve_apply(np.array(off), np.concatenate(idx_all), np.concatenate(rho_all), hd,
                           so["err"] ** 2, np.asarray(p["dep"])[perm], xb, np.ones(npts))
    t = time.perf_counter() - t0
    return {"value": npts / t, "unit": "local analyses/s", "cores": oracle.ref_threads(),
            "kind": "reference",
            "sample": (f"{npts} seeded radar points of the {nx}x{ny}x{nlev} grid, their capped "
                       f"local lists (mean {np.mean(np.diff(off)):.0f} obs): reference Fortran "
                       f"letkf_core + EISPACK + reference DGEMM + apply, {t:.2f} s; selection "
                       f"not timed")}


def gfs_bench(steps=2, precs=("f64", "f32", "bf16x3")):
    """GFS-sized global config: 192x94x64, u,v,T,q + ps, k=40, ~600k
    clustered obs, sigma_h 700 km, sigma_v 0.4 ln p, nobs_max = 1000 nearest-first.
    Each form's analysis is compared with the FP64 GPU analysis."""
    from letkf_amd import das, synth
    from letkf_amd.device import Event
    t0 = time.perf_counter()
    p = synth.gfs()
    gen_s = time.perf_counter() - t0
    npt = p["nij"] * p["nlev"]
    spread = p["gues3d"].std(axis=2, ddof=1)
    res, ref = {}, None
    for prec in precs:
        h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                            iq=p["iq"], q_top_lev=p["q_top_lev"], sigma_obs=p["sigma_obs"],
                            sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"], prec=prec,
                            infl_value=p["infl"], nobs_max=1000)
        h.set_ensemble(p["gues3d"], p["gues2d"])
        o = p["obs"]
        h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
        h.run()
        h.sync()
        st = h.stream()
        lat, sel, sol = [], [], []
        e = [Event() for _ in range(3)]
        for _ in range(steps):
            t1 = time.perf_counter()
            e[0].record(st)
            h.select()
            e[1].record(st)
            h.solve()
            e[2].record(st)
            h.sync()
            lat.append((time.perf_counter() - t1) * 1e3)
            sel.append(Event.elapsed_ms(e[0], e[1]))
            sol.append(Event.elapsed_ms(e[1], e[2]))
        a3, _ = h.analysis()
        nl = h.nobsl()
        h.close()
        r = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3), "cycle_ms": float(np.mean(lat)),
             "select_ms": float(np.mean(sel)), "solve_apply_ms": float(np.mean(sol)),
             "mean_nobsl": float(nl.mean())}
        if ref is None:
            ref = a3
        else:
            r["err_vs_f64_over_bg_spread"] = float((np.abs(a3 - ref).max(axis=2) /
                                                   np.maximum(spread, 1e-300)).max())
        res[prec] = r
        print(f"[bench] gfs {prec}: {r['cycle_ms']:.0f} ms/cycle", file=sys.stderr, flush=True)
    for prec in res:
        if "f64" in res:
            res[prec]["speedup_vs_f64"] = (res[prec]["local_analyses_per_s"] /
                                           res["f64"]["local_analyses_per_s"])
    return {"grid": [p["lon"].size, p["lat"].size, p["sig"].size], "k": p["ne"],
            "local_analyses_per_cycle": npt, "nobs": int(p["obs"]["elm"].size),
            "nobs_max": 1000, "gen_s": gen_s, "forms": res}


def speedy_k256_bench(steps=2, precs=("f64", "f32", "bf16x3"), cpu_sample=0):
    """Large-k variant of the SPEEDY-sized config (configs[3], last sentence): the same
    96x48x7 grid, fields and obs with k=256 members.  Points with fewer local obs than
    members take the dual (p x p) form, the rest the k x k form; each form's analysis is
    compared with the FP64 GPU analysis."""
    from letkf_amd import das, synth
    from letkf_amd.device import Event
    t0 = time.perf_counter()
    p = synth.speedy(ne=256)
    gen_s = time.perf_counter() - t0
    npt = p["nij"] * p["nlev"]
    spread = p["gues3d"].std(axis=2, ddof=1)
    res, ref, keep = {}, None, None
    for prec in precs:
        h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                            sigma_obs=p["sigma_obs"], sigma_obsv=p["sigma_obsv"],
                            gross_error=p["gross_error"], prec=prec, infl_value=p["infl"])
        h.set_ensemble(p["gues3d"], p["gues2d"])
        o = p["obs"]
        h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
        h.run()
        h.sync()
        st = h.stream()
        lat = []
        for _ in range(steps):
            e0, e1 = Event(), Event()
            t1 = time.perf_counter()
            e0.record(st)
            h.run()
            e1.record(st)
            h.sync()
            lat.append((time.perf_counter() - t1) * 1e3)
        a3, _ = h.analysis()
        nl = h.nobsl()
        r = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3), "cycle_ms": float(np.mean(lat)),
             "mean_nobsl": float(nl.mean()), "frac_points_dual": float(np.mean((nl > 0) & (nl < 256)))}
        if ref is None:
            ref = a3
            keep = h
        else:
            r["err_vs_f64_over_bg_spread"] = float((np.abs(a3 - ref).max(axis=2) /
                                                   np.maximum(spread, 1e-300)).max())
            h.close()
        res[prec] = r
        print(f"[bench] speedy k=256 {prec}: {r['cycle_ms']:.0f} ms/cycle", file=sys.stderr,
              flush=True)
    for prec in res:
        if "f64" in res:
            res[prec]["speedup_vs_f64"] = (res[prec]["local_analyses_per_s"] /
                                           res["f64"]["local_analyses_per_s"])
    out = {"grid": [96, 48, 7], "k": 256, "local_analyses_per_cycle": npt, "gen_s": gen_s,
           "forms": res}
    if cpu_sample and keep is not None:
        out["cpu_baseline"] = k256_cpu_sample(p, keep, cpu_sample)
    if keep is not None:
        keep.close()
    return out


def k256_cpu_sample(p, h, npts):
    """Reference letkf_core + apply (u, v, T, q) on a seeded sample of the k=256 points,
    on the local lists the GPU selected (selection is pinned bit-exact by the tests)."""
    import oracle
    if not oracle.ref_available():
        return None
    sobs, _, _ = oracle.set_letkf_obs_global(p["obs"], p["lon"], p["lat"], p["gross_error"])
    nij = p["nij"]
    rng = np.random.default_rng(9)
    pts = np.sort(rng.choice(nij * p["nlev"], npts, replace=False))
    lists = [h.local(int(b)) for b in pts]
    off = np.cumsum([0] + [l[0].size for l in lists])
    xb = np.stack([p["gues3d"][int(b) % nij, int(b) // nij, :, :].T for b in pts])
    t0 = time.perf_counter()
    oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                           np.concatenate([l[1] for l in lists]),
                           np.ascontiguousarray(sobs["hdxf"].T), sobs["err"] ** 2, sobs["dep"],
                           np.ascontiguousarray(xb), np.full(npts, p["infl"]))
    t = time.perf_counter() - t0
    return {"value": npts / t, "unit": "local analyses/s", "cores": oracle.ref_threads(),
            "kind": "reference",
            "sample": (f"{npts} seeded points of the k=256 SPEEDY-sized grid on their local "
                       f"lists (mean {np.mean(np.diff(off)):.0f} obs): reference Fortran "
                       f"letkf_core + EISPACK + reference DGEMM + apply to 4 variables, "
                       f"{t:.2f} s; selection not timed")}


def serve_bench(duration_s=20.0, rate=2.0, precs=("f64", "bf16x3"), seed=5):
    """Served analysis stream: Poisson arrivals of LETKF cycles of mixed sizes,
    each with a 30-s deadline.  Mix (by arrival): 45% Lorenz-96 (N in 40..960,
    k in 20..200), 40% SPEEDY-sized (k=20 or 64), 15% regional radar on 64x64x60
    columns (k=50).  Every job is a fresh instance with its own background and obs."""
    from letkf_amd import serve, synth
    rng = np.random.default_rng(seed)
    tmpl = {"l96_40_20": serve.l96_job, "l96_240_40": serve.l96_job, "l96_960_80": serve.l96_job,
            "l96_960_200": serve.l96_job, "speedy_20": serve.global_job,
            "speedy_64": serve.global_job, "radar_64_50": serve.regional_job}
    probs = {"l96_40_20": synth.l96(nx=40, ne=20), "l96_240_40": synth.l96(nx=240, ne=40, seed=2),
             "l96_960_80": synth.l96(nx=960, ne=80, seed=3),
             "l96_960_200": synth.l96(nx=960, ne=200, seed=6), "speedy_20": synth.speedy(),
             "speedy_64": synth.speedy(ne=64, seed=8),
             "radar_64_50": synth.radar(sub=(64, 64, 60), seed=4)}
    names = list(tmpl)
    weights = np.array([0.15, 0.1, 0.1, 0.1, 0.25, 0.15, 0.15])
    arr = serve.poisson_arrivals(rate, duration_s, seed)
    kinds = rng.choice(len(names), size=arr.size, p=weights / weights.sum())
    out = {}
    for prec in precs:
        srv = serve.Server(streams=2)
        # warm-up: one job of each kind (kernel loading, first allocations)
        for n in names:
            srv.submit(tmpl[n](f"warm-{n}", probs[n], prec)).result(timeout=300)
        srv.completed.clear()
        t0 = time.perf_counter()
        futs = []
        for i, (ta, kk) in enumerate(zip(arr, kinds)):
            now = time.perf_counter() - t0
            if ta > now:
                time.sleep(ta - now)
            n = names[kk]
            futs.append(srv.submit(tmpl[n](f"{n}-{i}", probs[n], prec)))
        srv.drain(timeout=600)
        wall = time.perf_counter() - t0
        res = [f.result() for f in futs]
        stt = srv.stats()
        srv.close()
        lat = np.array([r.latency_s for r in res])
        out[prec] = {"jobs": len(res), "wall_s": wall, "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max()),
                     "deadline_met": float(np.mean(lat <= 30.0)),
                     "local_analyses_per_s": stt["local_analyses"] / wall,
                     "max_jobs_per_batch": int(max(r.batch_jobs for r in res))}
        print(f"[bench] serve {prec}: {len(res)} jobs p50 {out[prec]['p50_s']:.3f}s "
              f"p99 {out[prec]['p99_s']:.3f}s", file=sys.stderr, flush=True)
    return {"arrivals_per_s": rate, "duration_s": duration_s, "mix": dict(zip(names, weights.tolist())),
            "forms": out}


def copy_calibration(nbytes=1 << 30, reps=5):
    """A plain device copy of known size: FETCH_SIZE / WRITE_SIZE of its dispatches
    give the scale factor between the counters and HBM bytes."""
    from letkf_amd.device import DeviceArray, sync
    a = DeviceArray((nbytes,), np.uint8)
    b = DeviceArray((nbytes,), np.uint8)
    a.memset(1)
    for _ in range(reps):
        b.copy_from_device(a)
    sync()
    print(json.dumps({"copy_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--profile", action="store_true",
                    help="only the plain FP64 timed loop (for rocprofv3)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only a 1 GiB device-to-device copy, 5 times (HBM counter calibration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the radar and served-stream measurements")
    ap.add_argument("--radar-full", action="store_true",
                    help="radar on the full 256x256x60 grid instead of 128x128x60")
    ap.add_argument("--serve-s", type=float, default=20.0, help="served-stream duration")
    ap.add_argument("--only", choices=["radar", "radar-full", "gfs", "k256", "serve"],
                    help="run one of the other configurations alone and print its JSON "
                         "(for profiling); the default run prints the SPEEDY line with all "
                         "of them attached")
    ap.add_argument("--precs", default="f64,f32,bf16x3",
                    help="arithmetic forms for --only (comma separated)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the LETKF local analyses of a job run on a single "
              "MI355X (the multi-GPU design shards spatial blocks; it is not built). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from letkf_amd import synth
    from letkf_amd._lib import lib
    if lib().lekf_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calib:
        copy_calibration()
        return
    if args.only:
        precs = tuple(args.precs.split(","))
        if args.only in ("radar", "radar-full"):
            sub = (256, 256, 60) if args.only == "radar-full" else (128, 128, 60)
            r = radar_bench(sub, steps=args.steps, precs=precs)
        elif args.only == "gfs":
            r = gfs_bench(steps=args.steps, precs=precs)
        elif args.only == "k256":
            r = speedy_k256_bench(steps=args.steps, precs=precs)
        else:
            r = serve_bench(duration_s=args.serve_s, precs=precs)
        print(json.dumps(r), flush=True)
        return
    t0 = time.perf_counter()
    p = synth.speedy()
    gen_s = time.perf_counter() - t0
    print(f"[bench] generated SPEEDY-sized inputs in {gen_s:.1f}s", file=sys.stderr, flush=True)

    h, obs_ms = make_handle(p, "f64")
    npt = p["nij"] * p["nlev"]
    r = time_cycles(h, args.steps, args.warmup)
    nobsl = h.nobsl()
    ms_per_step = r["wall_ms"] / args.steps
    value = npt / (ms_per_step / 1e3)
    nv = np.full((p["nij"], p["nlev"]), 4)
    nv[:, 0] = 5
    flops = float(flops_per_point(nobsl, p["ne"], nv).sum())
    t_solve = float(np.mean(r["solve_ms"])) / 1e3
    achieved = flops / t_solve / 1e12
    out = {
        "metric": METRIC, "value": value, "unit": "local analyses/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SPEEDY-sized fields and obs, letkf_amd.synth.speedy)",
        "config": {"workload": "SPEEDY-sized global LETKF, plain FP64: T30 96x48x7, u,v,T,q+ps, "
                               "k=20, 415 stations x (7x4+1) obs, sigma_h 500 km, sigma_v 0.4",
                   "local_analyses_per_cycle": npt, "k": p["ne"], "nobs": int(p["obs"]["elm"].size),
                   "mean_nobsl": float(nobsl.mean()), "max_nobsl": int(nobsl.max())},
        "latency_ms": {"p50": percentile(r["lat"], 50), "p99": percentile(r["lat"], 99)},
        "phases_ms": {"select": float(np.mean(r["select_ms"])),
                      "solve_apply": float(np.mean(r["solve_ms"])), "obs_setup_once": obs_ms},
        "roofline": {"bound": "mfma", "kernel": "lekf_solve_kernel (FP64)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic("lekf_solve_speedy_f64.md"),
                     "flops_per_launch": flops},
    }
    if args.profile:
        print(json.dumps(out), flush=True)
        return

    # FP64 oracle (CPU restatement of das_letkf) for the error of every form
    import oracle
    t0 = time.perf_counter()
    ref = oracle.das_global_problem(p)
    print(f"[bench] FP64 oracle cycle on CPU in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
          flush=True)
    spread = p["gues3d"].std(axis=2, ddof=1)
    spread = np.maximum(spread, 1e-300)

    def err_vs_oracle(hh):
        a3, _ = hh.analysis()
        e = np.abs(a3 - ref["anal3d"]).max(axis=2)
        dev = np.abs(ref["anal3d"] - ref["anal3d"].mean(axis=2, keepdims=True)).max(axis=(0, 1, 2))
        return {"max_rel_to_bg_spread": float((e / spread).max()),
                "max_abs_over_analysis_perturbation_scale": float(
                    (np.abs(a3 - ref["anal3d"]).max(axis=(0, 1, 2)) / dev).max())}

    out["err_vs_oracle"] = err_vs_oracle(h)
    if not args.no_mixed:
        mixed = {}
        for prec in ("f32", "bf16x3"):
            hm, _ = make_handle(p, prec)
            rm = time_cycles(hm, args.steps, args.warmup)
            msm = rm["wall_ms"] / args.steps
            mixed[prec] = {"value": npt / (msm / 1e3), "ms_per_step": msm,
                           "latency_ms": {"p50": percentile(rm["lat"], 50),
                                          "p99": percentile(rm["lat"], 99)},
                           "solve_ms": float(np.mean(rm["solve_ms"])),
                           "speedup_vs_f64": (npt / (msm / 1e3)) / value,
                           "err_vs_oracle": err_vs_oracle(hm)}
            hm.close()
        out["mixed"] = mixed
    if not args.no_cpu:
        t0 = time.perf_counter()
        cb = cpu_baseline(p, ref)
        print(f"[bench] CPU baseline in {time.perf_counter() - t0:.1f}s", file=sys.stderr,
              flush=True)
        out["cpu_baseline"] = cb
    h.close()
    if not args.no_extra:
        sub = (256, 256, 60) if args.radar_full else (128, 128, 60)
        out["radar"] = radar_bench(sub, steps=3, cpu_sample=0 if args.no_cpu else 1500)
        out["gfs"] = gfs_bench()
        out["speedy_k256"] = speedy_k256_bench(cpu_sample=0 if args.no_cpu else 480)
        out["serve"] = serve_bench(duration_s=args.serve_s)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
