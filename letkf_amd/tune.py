"""Launch autotuner of the analysis cycle.

Three launch parameters of the library are process-wide and can be chosen at
run time (``lekf_tune_set`` / ``lekf_tune_get`` in include/lekf.h):

``gram_waves``
    waves per point of the k > 32 Gram kernel (2, 4, 8).  Changes only the
    order in which the waves' partial Grams are summed.
``gram_ws_mb``
    Gram workspace per launch in MiB; sets how many points one Gram/eigen
    launch pair covers.  Results are bit-identical for every value.
``sel_per_cu``
    workgroups per CU of the capped selection kernels (1..4).  Results are
    bit-identical for every value.

:func:`autotune` times the whole cycle of a live handle (``lekf_das_run`` on
the handle's stream, bracketed by device events) for each candidate value,
one parameter at a time, keeps the fastest, and records the choice in a JSON
cache keyed by device, grid kind, ensemble size and precision.  :func:`load`
applies a cached choice.  The tuner changes launch shapes only; it never
changes which points are analysed or which observations they use.
"""
from __future__ import annotations

import json
import os
import time

from . import _lib
from .device import Event

PARAMS = {
    "gram_waves": (_lib.LEKF_TUNE_GRAM_WAVES, (2, 4, 8)),
    "gram_ws_mb": (_lib.LEKF_TUNE_GRAM_WS_MB, (1024, 4096, 16384)),
    "sel_per_cu": (_lib.LEKF_TUNE_SEL_PER_CU, (1, 2, 3, 4)),
    "eig_split": (_lib.LEKF_TUNE_EIG_SPLIT, (1, 0)),
}

# knobs that are set by hand, never searched: the rotations a point's QL stream
# may hold in the three-kernel eigensolver before the point takes the fused
# kernel's steps instead (0: the default 2k^2 + 8k), and the condition number
# above which a mixed-precision point is solved again in FP64 (an accuracy
# choice, not a speed one)
FIXED = {"split_rcap": _lib.LEKF_TUNE_SPLIT_RCAP, "mixed_kappa": _lib.LEKF_TUNE_MIXED_KAPPA}


def _pid(name):
    return PARAMS[name][0] if name in PARAMS else FIXED[name]


DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "letkf_amd", "tune.json")


def get(name):
    """Current value of a tuning parameter."""
    v = _lib.lib().lekf_tune_get(_pid(name))
    if v < 0:
        raise _lib.LetkfError(-v, "tune_get")
    return int(v)


def set(name, value):  # noqa: A001 - mirrors lekf_tune_set
    """Set a tuning parameter (raises LetkfError for a value out of range)."""
    _lib.check(_lib.lib().lekf_tune_set(_pid(name), int(value)), "tune_set")


def current():
    return {n: get(n) for n in PARAMS}


def apply(values):
    for n, v in values.items():
        if n in PARAMS:
            set(n, v)


def cycle_ms(handle, runs=3):
    """Median device time (ms) of ``runs`` analysis cycles of a handle whose
    ensemble and observations are set, after one untimed cycle."""
    st = handle.stream()
    handle.run()
    handle.sync()
    times = []
    for _ in range(runs):
        e0, e1 = Event(), Event()
        e0.record(st)
        handle.run()
        e1.record(st)
        times.append(Event.elapsed_ms(e0, e1))
    times.sort()
    return times[len(times) // 2]


def cache_key(handle, prec):
    dev = os.environ.get("LEKF_TUNE_DEVICE", "gfx950")
    kind = type(handle).__name__
    return f"{dev}|{kind}|k{handle.ne}|{prec}"


def _read(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load(key, path=None):
    """Apply the cached choice for ``key``; returns it, or None if absent."""
    entry = _read(path or os.environ.get("LEKF_TUNE_CACHE", DEFAULT_CACHE)).get(key)
    if entry:
        apply(entry["params"])
        return entry["params"]
    return None


def autotune(handle, *, params=tuple(PARAMS), candidates=None, runs=3, key=None, path=None,
             log=None):
    """Coordinate search over ``params`` on a live handle.

    Each parameter in turn takes every candidate value (``candidates[name]``
    or the defaults above) with the others held at their best so far; the
    fastest median cycle wins.  The winning values stay set on return.  With
    ``key`` the result is stored in the JSON cache at ``path``.
    Returns ``{"params": {...}, "ms": best, "trials": [(values, ms), ...]}``.
    """
    best = current()
    trials = []
    best_ms = cycle_ms(handle, runs)
    trials.append((dict(best), best_ms))
    for name in params:
        cands = (candidates or {}).get(name, PARAMS[name][1])
        for v in cands:
            if v == best[name]:
                continue
            set(name, v)
            ms = cycle_ms(handle, runs)
            trials.append(({**best, name: v}, ms))
            if log:
                log(f"tune {name}={v}: {ms:.3f} ms")
            if ms < best_ms:
                best_ms, best[name] = ms, v
        set(name, best[name])
    out = {"params": best, "ms": best_ms, "trials": trials}
    if key:
        path = path or os.environ.get("LEKF_TUNE_CACHE", DEFAULT_CACHE)
        data = _read(path)
        data[key] = {"params": best, "ms": best_ms, "time": time.time()}
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    return out
