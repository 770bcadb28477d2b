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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
