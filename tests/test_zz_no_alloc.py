"""lekf_core_batch_noalloc (the batched solve over the library's fixed device
scratch, no allocation) against lekf_core_batch: two calls in a row, the
second with a smaller input, each bit-identical to the allocating entry's
result, for k = 20 (no workspace), 40 and 50 (the split eigensolver's
workspace) and 160 (the large-k workspace).  Sizes beyond the scratch are
rejected with LEKF_ERR_SIZE."""
import ctypes

import numpy as np
import pytest

from letkf_amd import _lib, core
from letkf_amd.device import DeviceArray, sync

pytestmark = pytest.mark.gpu


def _problem(k, ps, seed):
    rng = np.random.default_rng(seed)
    nobs = 3000
    Y = rng.standard_normal((nobs, k))
    Y -= Y.mean(axis=1, keepdims=True)
    R = rng.uniform(0.5, 2.0, nobs)
    D = rng.standard_normal(nobs)
    idx, rho, off = [], [], [0]
    for p in ps:
        idx.append(rng.choice(nobs, p, replace=False))
        rho.append(rng.uniform(0.01, 1.0, p))
        off.append(off[-1] + p)
    t = DeviceArray.from_numpy
    npt = len(ps)
    bufs = dict(Y=t(Y), R=t(R), D=t(D), off=t(np.array(off), np.int64),
                idx=t(np.concatenate(idx).astype(np.int32), np.int32), rho=t(np.concatenate(rho)),
                infl=t(np.full(npt, 1.02)))
    return bufs, npt


def _solve(bufs, npt, k, noalloc):
    out = dict(infl_out=DeviceArray.zeros(npt), trans=DeviceArray.zeros(npt * k * k),
               transm=DeviceArray.zeros(npt * k))
    job = core.make_job(k=k, npt=npt, hdxf=bufs["Y"], rdiag=bufs["R"], dep=bufs["D"],
                        loc_off=bufs["off"], loc_idx=bufs["idx"], loc_rho=bufs["rho"],
                        infl_in=bufs["infl"], infl_out=out["infl_out"], trans=out["trans"])
    L = _lib.lib()
    arr = (_lib.LekfJob * 1)(job)
    fn = L.lekf_core_batch_noalloc if noalloc else L.lekf_core_batch
    _lib.check(fn(arr, 1, _lib.LEKF_PREC_F64, None), "core_batch")
    sync()
    return out["trans"].numpy(), out["infl_out"].numpy()


@pytest.mark.parametrize("k", [20, 40, 50, 160])
def test_noalloc_equals_core_batch(gpu, k):
    for call, ps in enumerate(([0, 1, 7, 64, 333, 1200, 50, 900] * 40, [5, 300, 2])):
        bufs, npt = _problem(k, ps, seed=10 * k + call)
        print(f"noalloc k={k} call {call}: {npt} points", flush=True)
        a_tr, a_in = _solve(bufs, npt, k, noalloc=False)
        b_tr, b_in = _solve(bufs, npt, k, noalloc=True)
        assert np.array_equal(a_tr, b_tr)
        assert np.array_equal(a_in, b_in)


def test_noalloc_rejects_too_many_jobs(gpu):
    L = _lib.lib()
    n = _lib.LEKF_NOALLOC_MAX_JOBS + 1
    arr = (_lib.LekfJob * n)()
    for j in arr:
        j.k = 20
        j.npt = 0
    assert L.lekf_core_batch_noalloc(arr, n, _lib.LEKF_PREC_F64, None) == _lib.LEKF_ERR_SIZE
    assert b"at most" in L.lekf_last_error()
