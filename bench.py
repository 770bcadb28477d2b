#!/usr/bin/env python3
"""Benchmark: LETKF local analyses/s on one MI355X, regional radar 256x256x60, k=50.

Workload (one "step" = one analysis cycle of the regional das_letkf over
resident inputs, scale/letkf/letkf_tools.f90:38-476): 256x256 columns x 60
levels of u,v,w,T,qv,qr, k=50 members, pseudo-radar obs on 60% of the
columns at every level (Doppler and reflectivity buckets), sigma_h = sigma_v
= 2 km, at most 1000 obs per bucket per point (the reference's
MAX_NOBS_PER_GRID).  3,932,160 local analyses per cycle: for each one the
selection (obs_local), the local solve (letkf_core) and the apply.

Prints ONE JSON line.  value = plain-FP64 local analyses/s (inputs resident in
HBM).  Beside it: the FP32 and split-BF16x3 forms, p50/p99 cycle latency of
every form (with and without the host-to-device obs upload), every form's
error against the reference Fortran letkf_core on a seeded sample of points,
the roofline of the dominant kernel (HIP events around its launches on the
handle's stream), the reference Fortran timed on this host's cores, and the
other configurations (SPEEDY-sized, GFS-sized, k=256, a served stream).
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
HBM_PEAK_GBS = 8000.0
MIXED_BUDGET = 1e-4       # analysis error / background spread of the mixed forms


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def percentile(a, q):
    return float(np.percentile(np.asarray(a), q))


# ---------------------------------------------------------------------------
# algorithmic work per local analysis
# ---------------------------------------------------------------------------
def gram_flops(p, k):
    """Y^T R^-1 Y (2pk^2), R^-1 Y (pk) and Y^T R^-1 d (2pk) of a point with p local obs."""
    p = np.asarray(p, dtype=np.float64)
    return 2.0 * p * k * k + 3.0 * p * k


def eig_flops(p, k, nv):
    """Eigensolve (9k^3), Pa and W (4k^3), w-bar and T (3k^2) and the apply (2 nv k^2)
    of a point with obs; the apply of the background alone (k^2 + 2 nv k^2) without."""
    p = np.asarray(p, dtype=np.float64)
    full = 13.0 * k ** 3 + 3.0 * k * k + 2.0 * nv * k * k
    return np.where(p > 0, full, 2.0 * nv * k * k + k * k)


def flops_per_point(p, k, nv):
    """Whole local analysis: the Gram and everything after it (a point without obs
    costs the apply of the background only)."""
    p = np.asarray(p, dtype=np.float64)
    return np.where(p > 0, gram_flops(p, k), 0.0) + eig_flops(p, k, nv)


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


def radar_scan_records(p):
    """Obs records the regional selection scans per cycle: at every point, the obs
    of the cells of its search rectangle's rows that can hold an obs within the
    cutoff (rows trimmed as the selection trims them; the same horizontal set at
    each of the nlev levels of a column).  Each record is 36 B (ri, rj, z, err and
    the obs index)."""
    nx, ny, nlev = p["nx"], p["ny"], p["nlev"]
    dx, dy, sh = p["dx"], p["dy"], p["sigma_h"]
    ci = np.clip(np.ceil(np.asarray(p["ri"]) - 0.5).astype(np.int64), 1, nx)
    cj = np.clip(np.ceil(np.asarray(p["rj"]) - 0.5).astype(np.int64), 1, ny)
    cnt = np.bincount((cj - 1) * nx + (ci - 1), minlength=nx * ny).reshape(ny, nx)
    cum = np.zeros((ny, nx + 1), np.int64)
    cum[:, 1:] = np.cumsum(cnt, axis=1)
    dlon, dlat = sh * 3.651483717 / dx, sh * 3.651483717 / dy
    r2cut = 13.33333333 * sh * sh * (1.0 + 1e-9)
    iy, ix = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    ri, rj = ix.astype(float), iy.astype(float)
    imin = np.maximum(1, np.floor(ri - dlon).astype(np.int64))
    imax = np.minimum(nx, np.ceil(ri + dlon).astype(np.int64))
    jmin = np.maximum(1, np.floor(rj - dlat).astype(np.int64))
    jmax = np.minimum(ny, np.ceil(rj + dlat).astype(np.int64))
    tot = np.zeros((ny, nx), np.int64)
    R = int(np.ceil(dlat)) + 1
    for dj in range(-R, R + 1):
        j = iy + dj
        ok = (j >= jmin) & (j <= jmax)
        dyg = np.maximum(0.0, np.abs(j - rj) - 0.5) * dy
        rem = r2cut - dyg * dyg
        ok &= rem >= 0.0
        rx = np.sqrt(np.maximum(rem, 0.0)) / dx + 0.5
        ilo = np.maximum(np.ceil(ri - rx).astype(np.int64), imin)
        ihi = np.minimum(np.floor(ri + rx).astype(np.int64), imax)
        ok &= ilo <= ihi
        jj = np.clip(j, 1, ny) - 1
        tot += np.where(ok, cum[jj, np.clip(ihi, 0, nx)] - cum[jj, np.clip(ilo - 1, 0, nx)], 0)
    return int(tot.sum()) * nlev


def roofline(kt, nobsl, k, nv, steps, profile_name, scan_records=None):
    """Roofline of the dominant kernel of the timed cycles, from the per-kernel
    HIP-event times kt = {name: (total ms, launches)}."""
    name, (ms, nl) = max(kt.items(), key=lambda kv: kv[1][0])
    per_cycle_s = ms / steps / 1e3
    launches = nl / steps
    nv_arr = np.full(nobsl.shape, nv)
    if name in ("gram", "solve", "big"):
        fl = gram_flops(nobsl[nobsl > 0], k).sum() if name == "gram" else \
            flops_per_point(nobsl, k, nv_arr).sum()
        bound, unit, peak = "mfma", "TFLOP/s", FP64_PEAK_TFLOPS
        achieved = fl / per_cycle_s / 1e12
        work = {"flops_per_cycle": float(fl), "flops_per_launch": float(fl / launches)}
    elif name.startswith("eig"):
        names = [n for n in kt if n.startswith("eig")]
        tot_s = sum(kt[n][0] for n in names) / steps / 1e3
        fl = eig_flops(nobsl, k, nv_arr).sum()
        bound, unit, peak = "mfma", "TFLOP/s", FP64_PEAK_TFLOPS
        achieved = fl / tot_s / 1e12
        work = {"flops_per_cycle": float(fl), "kernels": names}
    else:
        # selection, a scan: the obs records read (36 B per candidate and point)
        # and the lists written (index + rho, 12 B per kept entry)
        lb = 12.0 * float(nobsl.sum())
        sb = 36.0 * float(scan_records) if scan_records else 0.0
        by = sb + lb
        bound, unit, peak = "hbm", "GB/s", HBM_PEAK_GBS
        achieved = by / per_cycle_s / 1e9
        work = {"bytes_per_cycle": by, "bytes_per_launch": by / launches,
                "scan_bytes_per_cycle": sb, "list_bytes_per_cycle": lb}
    return {"bound": bound, "kernel": name, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": measured_traffic(profile_name),
            "ms_per_cycle": ms / steps, "launches_per_cycle": launches, **work}


def kernel_table(kt, steps):
    return {n: {"ms_per_cycle": t / steps, "launches_per_cycle": c / steps}
            for n, (t, c) in sorted(kt.items(), key=lambda kv: -kv[1][0])}


def host_cores():
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    return {"nproc": os.cpu_count() or 1, "affinity": aff,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


# ---------------------------------------------------------------------------
# regional radar, 256 x 256 x 60, k = 50 (the headline)
# ---------------------------------------------------------------------------
def radar_handle(p, prec):
    from letkf_amd import das
    return das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                             ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"],
                             sigma_v=p["sigma_v"], nobs_max=p["nobs_max"], infl_mode="fixed",
                             infl_value=1.0, prec=prec)


def time_radar_form(p, prec, steps, warmup, sample_pts=None, ktime=True):
    """One arithmetic form: set_ensemble, set_obs (timed, host buffers), `warmup`
    untimed cycles, `steps` timed cycles (HIP events on the handle's stream and
    the host clock); per-kernel times from the library's launch events."""
    from letkf_amd.device import Event, ktime_all, ktime_enable
    h = radar_handle(p, prec)
    h.set_ensemble(p["gues3d"])
    obs_ms = []
    for _ in range(2):
        t0 = time.perf_counter()
        h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
        h.sync()
        obs_ms.append((time.perf_counter() - t0) * 1e3)
    st = h.stream()
    for _ in range(warmup):
        h.run()
    h.sync()
    ev = [(Event(), Event()) for _ in range(steps)]
    lat = []
    if ktime:
        ktime_enable(True)
    t_start = time.perf_counter()
    for i in range(steps):
        t0 = time.perf_counter()
        ev[i][0].record(st)
        h.run()
        ev[i][1].record(st)
        h.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
    wall_ms = (time.perf_counter() - t_start) * 1e3
    kt = ktime_all() if ktime else {}
    if ktime:
        ktime_enable(False)
    gpu_ms = [Event.elapsed_ms(a, b) for a, b in ev]
    out = {"h": h, "wall_ms": wall_ms, "lat": lat, "gpu_ms": gpu_ms, "kt": kt,
           "obs_ms": float(np.median(obs_ms))}
    return out


def radar_headline(args):
    from letkf_amd import synth
    t0 = time.perf_counter()
    p = synth.radar()
    gen_s = time.perf_counter() - t0
    npt = p["nx"] * p["ny"] * p["nlev"]
    k, nv = p["ne"], p["nv3d"]
    log(f"generated radar 256x256x60 inputs ({p['ri'].size} obs) in {gen_s:.1f}s")

    # seeded sample of points for the reference Fortran (the FP64 oracle at this size)
    rng = np.random.default_rng(11)
    nij = p["nx"] * p["ny"]
    sample = np.sort(rng.choice(npt, args.cpu_points, replace=False))  # b = lev * nij + ij
    s_ij, s_lev = sample % nij, sample // nij
    xb_s = np.ascontiguousarray(p["gues3d"][s_ij, s_lev].transpose(0, 2, 1))  # (pt, var, member)
    spread_s = np.maximum(xb_s.std(axis=2, ddof=1), 1e-300)

    forms = {}
    lists = None
    xa_s = {}
    nobsl = None
    res64 = None
    for prec in args.precs:
        steps = args.steps if prec == "f64" else min(args.steps, args.mixed_steps)
        warm = args.warmup if prec == "f64" else min(args.warmup, 1)
        r = time_radar_form(p, prec, steps, warm)
        h = r["h"]
        if nobsl is None:
            nobsl = h.nobsl()
        if lists is None:
            lists = [h.local(int(b), cap=2 * p["nobs_max"] + 64) for b in sample]
        a3, _ = h.analysis()
        xa_s[prec] = np.ascontiguousarray(a3[s_ij, s_lev].transpose(0, 2, 1))
        del a3
        nfp64 = h.fp64_count()
        h.close()
        ms = r["wall_ms"] / steps
        f = {"local_analyses_per_s": npt / (ms / 1e3), "ms_per_step": ms, "steps": steps,
             "warmup": warm,
             "latency_ms": {"p50": percentile(r["lat"], 50), "p99": percentile(r["lat"], 99),
                            "max": float(np.max(r["lat"]))},
             "gpu_event_ms": {"p50": percentile(r["gpu_ms"], 50),
                              "p99": percentile(r["gpu_ms"], 99)},
             "set_obs_ms": r["obs_ms"],
             "latency_with_set_obs_ms": {"p50": percentile(r["lat"], 50) + r["obs_ms"],
                                         "p99": percentile(r["lat"], 99) + r["obs_ms"]},
             "kernels": kernel_table(r["kt"], steps)}
        if prec != "f64":
            f["fp64_redo_points"] = nfp64  # condition number above the mixed_kappa threshold
        forms[prec] = f
        if prec == "f64":
            res64 = r
        log(f"radar {prec}: {ms:.0f} ms/cycle (p99 {f['latency_ms']['p99']:.0f} ms), "
            f"set_obs {r['obs_ms']:.0f} ms")

    # reference Fortran letkf_core + apply on the sample: the FP64 oracle and the CPU baseline
    import oracle
    perm, _ = oracle.obs_bins_regional(p["ri"], p["rj"], p["nx"], p["ny"])
    off = np.cumsum([0] + [l[0].size for l in lists])
    hd = np.ascontiguousarray(np.asarray(p["hdxf"])[perm])
    rd = np.asarray(p["err"])[perm] ** 2
    dep = np.asarray(p["dep"])[perm]
    t0 = time.perf_counter()
    xa_ref, _ = oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                                       np.concatenate([l[1] for l in lists]), hd, rd, dep, xb_s,
                                       np.ones(sample.size), fixed_infl=1.0)
    t_ref = time.perf_counter() - t0
    dev = np.abs(xa_ref - xa_ref.mean(axis=2, keepdims=True)).max(axis=(0, 2))
    for prec, f in forms.items():
        d = np.abs(xa_s[prec] - xa_ref)
        f["err_vs_oracle"] = {
            "max_over_bg_spread": float((d.max(axis=2) / spread_s).max()),
            "max_over_perturbation_scale": float((d.max(axis=(0, 2)) / dev).max()),
            "points": int(sample.size)}
        if prec != "f64":
            f["err_within_budget"] = bool(f["err_vs_oracle"]["max_over_bg_spread"] <= MIXED_BUDGET)
            f["speedup_vs_f64"] = f["local_analyses_per_s"] / forms["f64"]["local_analyses_per_s"]
    cores = host_cores()
    cpu = {"value": sample.size / t_ref, "unit": "local analyses/s",
           "cores": oracle.ref_threads(), "kind": "reference",
           "sample": (f"{sample.size} seeded points of the radar 256x256x60 k=50 grid on their "
                      f"capped local lists (mean {np.mean(np.diff(off)):.0f} obs): reference "
                      f"Fortran letkf_core + EISPACK + reference DGEMM + apply to 6 variables, "
                      f"OpenMP over points, {t_ref:.2f} s; selection not timed"),
           "host": cores}
    log(f"reference Fortran on {sample.size} radar points: {t_ref:.1f}s "
        f"({oracle.ref_threads()} threads, nproc {cores['nproc']})")

    f64 = forms["f64"]
    rl = roofline(res64["kt"], nobsl, k, nv, args.steps, "radar_256x256x60_f64.md",
                  scan_records=radar_scan_records(p))
    out = {
        "metric": METRIC, "value": f64["local_analyses_per_s"], "unit": "local analyses/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded radar fields and obs, letkf_amd.synth.radar)",
        "config": {"workload": ("regional radar 256x256x60, k=50, u,v,w,T,qv,qr, pseudo-radar "
                                "obs on 60% of columns at every level, sigma_h = sigma_v = 2 km, "
                                "nobs_max 1000 per bucket, plain FP64"),
                   "local_analyses_per_cycle": npt, "k": k, "nobs": int(p["ri"].size),
                   "mean_nobsl": float(nobsl.mean()), "max_nobsl": int(nobsl.max())},
        "latency_ms": f64["latency_ms"],
        "set_obs_ms": f64["set_obs_ms"],
        "forms": forms,
        "roofline": rl,
        "cpu_baseline": cpu,
        "gen_s": gen_s,
    }
    return out, p


# ---------------------------------------------------------------------------
# other configurations
# ---------------------------------------------------------------------------
def speedy_bench(steps=5, warmup=2, precs=("f64", "f32", "bf16x3"), cpu=True):
    """SPEEDY-sized global config: T30 96x48x7, u,v,T,q + ps, k=20, 415 stations,
    sigma_h 500 km, sigma_v 0.4 ln p, adaptive inflation; errors against the full
    FP64 oracle cycle (the CPU restatement of das_letkf)."""
    import oracle
    from letkf_amd import das, synth
    from letkf_amd.device import Event
    p = synth.speedy()
    npt = p["nij"] * p["nlev"]
    ref = oracle.das_global_problem(p)
    spread = np.maximum(p["gues3d"].std(axis=2, ddof=1), 1e-300)
    dev = np.abs(ref["anal3d"] - ref["anal3d"].mean(axis=2, keepdims=True)).max(axis=(0, 1, 2))
    res = {}
    for prec in precs:
        h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                            ips=0, iq=p["iq"], q_top_lev=p["q_top_lev"], sigma_obs=p["sigma_obs"],
                            sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"], prec=prec,
                            infl_value=p["infl"])
        h.set_ensemble(p["gues3d"], p["gues2d"])
        o = p["obs"]
        h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
        st = h.stream()
        for _ in range(warmup):
            h.run()
        h.sync()
        lat = []
        for _ in range(steps):
            e0, e1 = Event(), Event()
            t0 = time.perf_counter()
            e0.record(st)
            h.run()
            e1.record(st)
            h.sync()
            lat.append((time.perf_counter() - t0) * 1e3)
        a3, _ = h.analysis()
        nl = h.nobsl()
        h.close()
        d = np.abs(a3 - ref["anal3d"])
        res[prec] = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3),
                     "cycle_ms": float(np.mean(lat)), "p50_ms": percentile(lat, 50),
                     "p99_ms": percentile(lat, 99), "mean_nobsl": float(nl.mean()),
                     "err_vs_oracle": {"max_over_bg_spread": float((d.max(axis=2) / spread).max()),
                                       "max_over_perturbation_scale":
                                           float((d.max(axis=(0, 1, 2)) / dev).max())}}
    for prec in res:
        res[prec]["speedup_vs_f64"] = (res[prec]["local_analyses_per_s"] /
                                       res["f64"]["local_analyses_per_s"])
    out = {"grid": [96, 48, 7], "k": p["ne"], "local_analyses_per_cycle": npt,
           "nobs": int(p["obs"]["elm"].size), "forms": res}
    if cpu and oracle.ref_available():
        out["cpu_baseline"] = speedy_cpu(p, ref)
    log(f"speedy: f64 {res['f64']['cycle_ms']:.2f} ms/cycle")
    return out


def speedy_cpu(p, ref):
    """Reference Fortran letkf_core + apply over the full SPEEDY-sized cycle's local
    problems (lists from the oracle's obs_local restatement)."""
    import oracle
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
    xb = np.ascontiguousarray(p["gues3d"].transpose(1, 0, 3, 2)).reshape(nlev * nij, 4, ne)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.ref_solve_apply(off, np.concatenate(idx_all), np.concatenate(rho_all),
                               np.ascontiguousarray(s["hdxf"].T), s["err"] ** 2, s["dep"], xb,
                               np.full(nlev * nij, p["infl"]))
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": nlev * nij / t, "unit": "local analyses/s", "cores": oracle.ref_threads(),
            "kind": "reference",
            "sample": (f"full SPEEDY-sized cycle, {nlev * nij} local analyses, median of 3 "
                       f"runs ({t:.3f} s)")}


def gfs_bench(steps=3, precs=("f64", "f32", "bf16x3")):
    """GFS-sized global config: 192x94x64, u,v,T,q + ps, k=40, sigma_h 700 km,
    sigma_v 0.4 ln p, nobs_max = 1000 nearest first.  Each form's analysis is
    compared with the FP64 GPU analysis (held to 1e-10 of the reference Fortran
    by the tests)."""
    from letkf_amd import das, synth
    from letkf_amd.device import Event, ktime_all, ktime_enable
    t0 = time.perf_counter()
    p = synth.gfs()
    gen_s = time.perf_counter() - t0
    npt = p["nij"] * p["nlev"]
    spread = np.maximum(p["gues3d"].std(axis=2, ddof=1), 1e-300)
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
        lat = []
        ktime_enable(True)
        for _ in range(steps):
            e0, e1 = Event(), Event()
            t1 = time.perf_counter()
            e0.record(st)
            h.run()
            e1.record(st)
            h.sync()
            lat.append((time.perf_counter() - t1) * 1e3)
        kt = ktime_all()
        ktime_enable(False)
        a3, _ = h.analysis()
        nl = h.nobsl()
        nfp64 = h.fp64_count()
        h.close()
        r = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3), "cycle_ms": float(np.mean(lat)),
             "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
             "mean_nobsl": float(nl.mean()), "kernels": kernel_table(kt, steps)}
        if prec != "f64":
            r["fp64_redo_points"] = nfp64
        if ref is None:
            ref = a3
        else:
            r["err_vs_f64_over_bg_spread"] = float((np.abs(a3 - ref).max(axis=2) / spread).max())
        res[prec] = r
        log(f"gfs {prec}: {r['cycle_ms']:.0f} ms/cycle")
    for prec in res:
        res[prec]["speedup_vs_f64"] = (res[prec]["local_analyses_per_s"] /
                                       res["f64"]["local_analyses_per_s"])
    return {"grid": [p["lon"].size, p["lat"].size, p["sig"].size], "k": p["ne"],
            "local_analyses_per_cycle": npt, "nobs": int(p["obs"]["elm"].size),
            "nobs_max": 1000, "gen_s": gen_s, "forms": res}


def speedy_k256_bench(steps=3, precs=("f64", "f32", "bf16x3")):
    """The SPEEDY-sized grid with k=256 members (points with fewer local obs than
    members take the dual p x p form).  Errors against the FP64 GPU analysis."""
    from letkf_amd import das, synth
    from letkf_amd.device import Event
    p = synth.speedy(ne=256)
    npt = p["nij"] * p["nlev"]
    spread = np.maximum(p["gues3d"].std(axis=2, ddof=1), 1e-300)
    res, ref = {}, None
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
        h.close()
        r = {"local_analyses_per_s": npt / (np.mean(lat) / 1e3), "cycle_ms": float(np.mean(lat)),
             "mean_nobsl": float(nl.mean()),
             "frac_points_dual": float(np.mean((nl > 0) & (nl < 256)))}
        if ref is None:
            ref = a3
        else:
            r["err_vs_f64_over_bg_spread"] = float((np.abs(a3 - ref).max(axis=2) / spread).max())
        res[prec] = r
        log(f"speedy k=256 {prec}: {r['cycle_ms']:.0f} ms/cycle")
    for prec in res:
        res[prec]["speedup_vs_f64"] = (res[prec]["local_analyses_per_s"] /
                                       res["f64"]["local_analyses_per_s"])
    return {"grid": [96, 48, 7], "k": 256, "local_analyses_per_cycle": npt, "forms": res}


def serve_bench(radar_p, duration_s=60.0, prec="bf16x3", seed=5):
    """Served analysis stream: Poisson arrivals, each job a LETKF cycle with a 30-s
    deadline.  The radar 256x256x60 k=50 domain cycles at 0.1 jobs/s (a fresh
    background ensemble and obs set each cycle, the instance resident between
    cycles); beside it 0.9 jobs/s of small instances: Lorenz-96 (N 40..960,
    k 20..200) and SPEEDY-sized (k 20, 64).  The arrival rates are chosen so the
    measured service times keep the expected utilisation below 1
    (utilisation_est = sum over kinds of rate x alone service time)."""
    from letkf_amd import serve, synth
    rng = np.random.default_rng(seed)
    small = {"l96_40_20": (serve.l96_job, synth.l96(nx=40, ne=20), 0.15),
             "l96_240_40": (serve.l96_job, synth.l96(nx=240, ne=40, seed=2), 0.15),
             "l96_960_80": (serve.l96_job, synth.l96(nx=960, ne=80, seed=3), 0.15),
             "l96_960_200": (serve.l96_job, synth.l96(nx=960, ne=200, seed=6), 0.1),
             "speedy_20": (serve.global_job, synth.speedy(), 0.25),
             "speedy_64": (serve.global_job, synth.speedy(ne=64, seed=8), 0.1)}
    radar_rate = 0.1
    small_rate = 0.9
    srv = serve.Server(streams=2)
    alone = {}
    # warm-up and alone service time of each kind
    for n, (fn, pr, _) in small.items():
        t0 = time.perf_counter()
        srv.submit(fn(f"warm-{n}", pr, prec)).result(timeout=300)
        alone[n] = time.perf_counter() - t0
    t0 = time.perf_counter()
    srv.submit(serve.regional_job("radar", radar_p, prec)).result(timeout=300)
    alone["radar_256_50_first"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    srv.submit(serve.regional_job("radar", radar_p, prec)).result(timeout=300)
    alone["radar_256_50"] = time.perf_counter() - t0
    srv.completed.clear()
    names = list(small)
    w = np.array([small[n][2] for n in names])
    w = w / w.sum()
    util = radar_rate * alone["radar_256_50"] + small_rate * float(
        sum(wi * alone[n] for wi, n in zip(w, names)))
    log(f"serve: radar alone {alone['radar_256_50']:.2f}s, expected utilisation {util:.2f}")
    ar = serve.poisson_arrivals(radar_rate, duration_s, seed)
    if ar.size == 0:
        ar = np.array([0.0])
    asml = serve.poisson_arrivals(small_rate, duration_s, seed + 1)
    kinds = rng.choice(len(names), size=asml.size, p=w)
    events = sorted([(t, "radar") for t in ar] + [(t, names[k]) for t, k in zip(asml, kinds)])
    futs = []
    t0 = time.perf_counter()
    for i, (ta, n) in enumerate(events):
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        if n == "radar":
            futs.append((n, srv.submit(serve.regional_job("radar", radar_p, prec))))
        else:
            fn, pr, _ = small[n]
            futs.append((n, srv.submit(fn(f"{n}-{i}", pr, prec))))
    srv.drain(timeout=900)
    wall = time.perf_counter() - t0
    res = [(n, f.result()) for n, f in futs]
    st = srv.stats()
    srv.close()
    lat = np.array([r.latency_s for _, r in res])
    rl = np.array([r.latency_s for n, r in res if n == "radar"])
    sl = np.array([r.latency_s for n, r in res if n != "radar"])
    out = {"form": prec, "duration_s": duration_s, "arrivals_per_s": {"radar_256_50": radar_rate,
                                                                       "small": small_rate},
           "small_mix": dict(zip(names, w.tolist())), "alone_service_s": alone,
           "utilisation_est": util, "jobs": len(res), "wall_s": wall,
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "radar": {"jobs": int(rl.size), "p50_s": float(np.percentile(rl, 50)),
                     "p99_s": float(np.percentile(rl, 99)), "max_s": float(rl.max()),
                     "deadline_met": float(np.mean(rl <= 30.0))},
           "small": {"jobs": int(sl.size),
                     "p50_s": float(np.percentile(sl, 50)) if sl.size else None,
                     "p99_s": float(np.percentile(sl, 99)) if sl.size else None},
           "deadline_met": float(np.mean(lat <= 30.0)),
           "local_analyses_per_s": st["local_analyses"] / wall,
           "max_jobs_per_batch": int(max(r.batch_jobs for _, r in res))}
    log(f"serve {prec}: {len(res)} jobs, radar p99 {out['radar']['p99_s']:.2f}s, "
        f"all p99 {out['p99_s']:.2f}s")
    return out


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
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mixed-steps", type=int, default=6,
                    help="timed cycles of the FP32 / BF16x3 forms (at most --steps)")
    ap.add_argument("--precs", default="f64,f32,bf16x3")
    ap.add_argument("--cpu-points", type=int, default=16000,
                    help="radar points of the reference Fortran sample")
    ap.add_argument("--profile", action="store_true",
                    help="only the radar timed cycles of the first form (for rocprofv3)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only a 1 GiB device-to-device copy, 5 times (HBM counter calibration)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the SPEEDY, GFS, k=256 and served-stream measurements")
    ap.add_argument("--serve-s", type=float, default=60.0, help="served-stream duration")
    ap.add_argument("--only", choices=["speedy", "gfs", "k256"],
                    help="run one of the other configurations alone (with --steps, --precs "
                         "and, for speedy, --warmup) and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the LETKF local analyses of a job run on a single "
              "MI355X (the multi-GPU design shards spatial blocks; it is not built). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    args.precs = tuple(args.precs.split(","))
    if args.precs[0] != "f64" and not args.profile:
        print("the first form must be f64 (the headline value)", file=sys.stderr)
        sys.exit(2)

    from letkf_amd._lib import lib
    if lib().lekf_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calib:
        copy_calibration()
        return
    if args.only:
        kw = {"steps": args.steps, "precs": args.precs}
        if args.only == "speedy":
            kw["warmup"] = args.warmup
        r = {"speedy": speedy_bench, "gfs": gfs_bench, "k256": speedy_k256_bench}[args.only](**kw)
        print(json.dumps(r), flush=True)
        return
    if args.profile:
        from letkf_amd import synth
        p = synth.radar()
        log("generated radar inputs")
        r = time_radar_form(p, args.precs[0], args.steps, args.warmup, ktime=False)
        ms = r["wall_ms"] / args.steps
        print(json.dumps({"form": args.precs[0], "ms_per_step": ms,
                          "p99_ms": percentile(r["lat"], 99)}), flush=True)
        r["h"].close()
        return

    out, radar_p = radar_headline(args)
    if not args.no_extra:
        extra = {}
        extra["serve"] = serve_bench(radar_p, duration_s=args.serve_s)
        del radar_p
        extra["speedy"] = speedy_bench()
        extra["gfs"] = gfs_bench()
        extra["speedy_k256"] = speedy_k256_bench()
        out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
