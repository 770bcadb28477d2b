"""Served analysis stream: continuous batching of LETKF cycles on one GPU.

A job is one analysis cycle of a model instance (its ``key``): a background
ensemble (or none, to continue from the instance's resident analysis) and an
observation set.  The server

* admits pending jobs as device memory allows (an HBM budget; idle resident
  instances are evicted least-recently-used when a new one needs room),
* keeps every instance's handle -- grid, ensemble, obs table and work buffers
  -- resident on the GPU between its cycles (the analysis becomes the next
  cycle's background unless the next job brings a new ensemble),
* launches the cycles of all admitted jobs as one batch on a stream
  (``lekf_das_run_batch``: every handle's selection, then shared local-solve
  launches grouped by kernel class), and keeps several batches in flight on
  different streams, so jobs that arrive while others run start as soon as a
  stream and memory free up,
* completes jobs as their batch's event fires and reports per-job latency.

A local analysis depends only on its own inputs, so a job's analysis is bit
for bit the same whether it runs alone or in any mix (checked by the tests).
Cycles of one instance run in submission order, one at a time.
"""
from __future__ import annotations

import collections
import concurrent.futures
import dataclasses
import threading
import time

import numpy as np

from . import _lib, das
from .device import Event, Stream, mem_info


@dataclasses.dataclass
class Job:
    key: str                      # model instance (its cycles share one resident handle)
    kind: str                     # "l96" | "global" | "regional"
    cfg: dict                     # handle constructor arguments (used on the first cycle)
    obs: dict                     # set_obs arguments of this cycle
    ensemble: dict | None = None  # set_ensemble arguments; None: continue from the analysis
    fetch: bool = True            # copy the analysis back to the host
    deadline_s: float = 30.0


@dataclasses.dataclass
class Result:
    key: str
    analysis: tuple | None
    submit_t: float
    start_t: float
    end_t: float
    batch_jobs: int
    local_analyses: int

    @property
    def latency_s(self):
        return self.end_t - self.submit_t

    @property
    def queued_s(self):
        return self.start_t - self.submit_t


_KINDS = {"l96": das.L96Letkf, "global": das.GlobalLetkf, "regional": das.RegionalLetkf}


def estimate_bytes(job: Job) -> int:
    """Device bytes a new instance needs (ensemble in and out, obs table, lists)."""
    c, o = job.cfg, job.obs
    ne = c["ne"]
    if job.kind == "l96":
        npt, nv, nobs = c["nx"], 1, len(o["y"])
        lists = npt * min(nobs, 64)
    elif job.kind == "global":
        npt = len(c["lon"]) * len(c["lat"]) * len(c["sig"])
        nv = c.get("nv3d", 4) + c.get("nv2d", 1)
        nobs = len(o["elm"])
        lists = npt * min(nobs, 2000)
    else:
        npt = c["nx"] * c["ny"] * len(c["zlev"])
        nv = c["nv3d"]
        nobs = len(o["ri"])
        lists = npt * min(nobs, 2 * max(c.get("nobs_max", 1000), 1))
    extra = 64 << 20
    # the local solve's workspace (k > 32) under the current tuning
    extra += int(_lib.lib().lekf_solve_workspace_bytes(npt, ne))
    if job.kind == "regional":
        extra += 1024 * min(nobs, 1 << 16) * 20  # selection scratch slots
    return int(2 * npt * nv * ne * 8 + nobs * (ne * 12 + 160) + lists * 12 + extra)


class _Instance:
    def __init__(self, key, handle, nbytes):
        self.key = key
        self.h = handle
        self.bytes = nbytes
        self.busy = False
        self.cycles = 0
        self.last_use = time.perf_counter()


class Server:
    """Continuous-batching LETKF server on the current GPU."""

    def __init__(self, hbm_budget: int | None = None, streams: int = 2, max_batch: int = 64):
        if hbm_budget is None:
            free, _ = mem_info()
            hbm_budget = int(free * 0.8)
        self.budget = hbm_budget
        self.max_batch = max_batch
        self._streams = [Stream() for _ in range(streams)]
        self._free_streams = list(self._streams)
        self._pending = collections.deque()
        self._inst: dict[str, _Instance] = {}
        self._inflight = []
        self._cv = threading.Condition()
        self._stop = False
        self._error = None
        self._outstanding = 0  # submitted and not yet resolved
        self.completed: list[Result] = []
        self._worker = threading.Thread(target=self._loop, name="letkf-serve", daemon=True)
        self._worker.start()

    # -- public -----------------------------------------------------------------
    def submit(self, job: Job) -> concurrent.futures.Future:
        if job.kind not in _KINDS:
            raise ValueError(f"unknown job kind {job.kind!r}")
        fut = concurrent.futures.Future()
        with self._cv:
            if self._error is not None:
                raise RuntimeError("server stopped after an error") from self._error
            self._pending.append((job, fut, time.perf_counter()))
            self._outstanding += 1
            self._cv.notify()
        return fut

    def drain(self, timeout: float | None = None):
        """Wait until every submitted job has completed."""
        t0 = time.perf_counter()
        with self._cv:
            while self._outstanding > 0:
                if self._error is not None:
                    raise RuntimeError("server stopped after an error") from self._error
                left = None if timeout is None else timeout - (time.perf_counter() - t0)
                if left is not None and left <= 0:
                    raise TimeoutError("jobs still running")
                self._cv.wait(0.05 if left is None else min(0.05, left))

    def close(self):
        with self._cv:
            self._stop = True
            self._cv.notify()
        self._worker.join()
        for inst in self._inst.values():
            inst.h.close()
        self._inst.clear()

    def resident_bytes(self) -> int:
        return sum(i.bytes for i in self._inst.values())

    def stats(self) -> dict:
        lat = np.array([r.latency_s for r in self.completed]) if self.completed else np.zeros(1)
        return {"jobs": len(self.completed),
                "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
                "max_s": float(lat.max()),
                "local_analyses": int(sum(r.local_analyses for r in self.completed))}

    # -- worker -----------------------------------------------------------------
    def _loop(self):
        try:
            while True:
                with self._cv:
                    if self._stop and self._outstanding == 0:
                        return
                self._complete_ready()
                started = self._start_batch()
                if not started:
                    with self._cv:
                        if not self._stop or self._inflight or self._pending:
                            self._cv.wait(0.0005 if self._inflight else 0.01)
        except BaseException as e:  # surface to callers rather than hang them
            with self._cv:
                self._error = e
                for job, fut, _ in self._pending:
                    fut.set_exception(e)
                self._pending.clear()
                for b in self._inflight:
                    for ent in b["entries"]:
                        if not ent["fut"].done():
                            ent["fut"].set_exception(e)
                self._inflight.clear()
                self._outstanding = 0
                self._cv.notify_all()

    def _complete_ready(self):
        done = [b for b in self._inflight if b["event"].query()]
        for b in done:
            t_end = time.perf_counter()
            for ent in b["entries"]:
                inst = ent["inst"]
                a = inst.h.analysis() if ent["job"].fetch else None
                inst.h.cycle_swap()  # the analysis is the next cycle's background
                inst.busy = False
                inst.cycles += 1
                inst.last_use = t_end
                r = Result(ent["job"].key, a, ent["submit_t"], b["start_t"], time.perf_counter(),
                           len(b["entries"]), inst.h.nij * inst.h.nlev)
                self.completed.append(r)
                ent["fut"].set_result(r)
            with self._cv:
                self._outstanding -= len(b["entries"])
                self._inflight.remove(b)
                self._free_streams.append(b["stream"])
                self._cv.notify_all()

    def _evict_for(self, need: int, protect: set) -> bool:
        while self.resident_bytes() + need > self.budget:
            idle = [i for k, i in self._inst.items() if not i.busy and k not in protect]
            if not idle:
                return False
            victim = min(idle, key=lambda i: i.last_use)
            victim.h.close()
            del self._inst[victim.key]
        return True

    def _start_batch(self) -> bool:
        with self._cv:
            if not self._free_streams or not self._pending:
                return False
            stream = self._free_streams[-1]
            take, rest = [], collections.deque()
            keys = set()
            for item in self._pending:
                job = item[0]
                inst = self._inst.get(job.key)
                if len(take) < self.max_batch and job.key not in keys and not (inst and inst.busy):
                    take.append(item)
                    keys.add(job.key)
                else:
                    rest.append(item)
            if not take:
                return False
            self._pending = rest  # later submissions append here
        entries = []
        back = []  # jobs waiting for memory go back to the front, in order
        for job, fut, t_sub in take:
            inst = self._inst.get(job.key)
            if inst is None:
                need = estimate_bytes(job)
                if not self._evict_for(need, keys):
                    back.append((job, fut, t_sub))
                    continue
                h = _KINDS[job.kind](**job.cfg)
                inst = _Instance(job.key, h, need)
                self._inst[job.key] = inst
            inst.busy = True
            try:
                inst.h.set_stream(stream.ptr)
                if job.ensemble is not None:
                    inst.h.set_ensemble(**job.ensemble)
                inst.h.set_obs(**job.obs)
                inst.bytes = inst.h.device_bytes()
            except BaseException as e:
                inst.busy = False
                fut.set_exception(e)
                with self._cv:
                    self._outstanding -= 1
                    self._cv.notify_all()
                continue
            entries.append({"job": job, "fut": fut, "submit_t": t_sub, "inst": inst})
        if back:
            with self._cv:
                for item in reversed(back):
                    self._pending.appendleft(item)
        if not entries:
            return False
        t_start = time.perf_counter()
        das.run_batch([e["inst"].h for e in entries], stream.ptr)
        ev = Event()
        ev.record(stream.ptr)
        with self._cv:
            self._free_streams.remove(stream)
            self._inflight.append({"event": ev, "stream": stream, "entries": entries,
                                   "start_t": t_start})
        return True


def poisson_arrivals(rate_per_s: float, duration_s: float, seed: int) -> np.ndarray:
    """Arrival times of a Poisson stream over [0, duration_s)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate_per_s)
        if t >= duration_s:
            return np.array(out)
        out.append(t)


# ---------------------------------------------------------------------------
# jobs from the synthetic problems of letkf_amd.synth
# ---------------------------------------------------------------------------
def l96_job(key, p, prec="f64", with_ensemble=True):
    nx, ne = p["nx"], p["ne"]
    cfg = dict(nx=nx, ne=ne, xlocal=p["xlocal"], prec=prec, infl_mode="fixed",
               infl_value=p["infl"])
    ens = dict(gues3d=np.asfortranarray(p["xb"]).reshape(nx, 1, ne, 1, order="F"))
    obs = dict(gidx=p["gidx"], y=p["y"], err=p["err"], hxf=p["hxf"])
    return Job(key, "l96", cfg, obs, ens if with_ensemble else None)


def global_job(key, p, prec="f64", with_ensemble=True):
    cfg = dict(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4,
               nv2d=p["gues2d"].shape[2], ips=p.get("ips", 0), iq=p.get("iq", 3),
               q_top_lev=p.get("q_top_lev", 5), sigma_obs=p["sigma_obs"],
               sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"], prec=prec,
               infl_value=p["infl"])
    o = p["obs"]
    obs = dict(elm=o["elm"], lon=o["lon"], lat=o["lat"], lev=o["lev"], dat=o["dat"],
               err=o["err"], hxf=o["hxf"])
    ens = dict(gues3d=p["gues3d"], gues2d=p["gues2d"])
    return Job(key, "global", cfg, obs, ens if with_ensemble else None)


def regional_job(key, p, prec="f64", with_ensemble=True, list_budget=0):
    cfg = dict(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"], ne=p["ne"],
               nv3d=p["nv3d"], sigma_h=p["sigma_h"], sigma_v=p["sigma_v"],
               nobs_max=p["nobs_max"], prec=prec, infl_value=1.0, list_budget=list_budget)
    obs = dict(ri=p["ri"], rj=p["rj"], z=p["z"], err=p["err"], dep=p["dep"],
               bucket=p["bucket"], hdxf=p["hdxf"])
    ens = dict(gues3d=p["gues3d"])
    return Job(key, "regional", cfg, obs, ens if with_ensemble else None)
