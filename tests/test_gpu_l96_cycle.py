"""Served Lorenz-96 cycling on the GPU: the on-device forecast
(lekf_l96_forecast, lorenz96/model/lorenz96.f90:93-128) and the on-device
H(x) (lorenz96/obs/h_ope.f90:14-39).

* The forecast of every member equals the reference model (oracle/_ref at
  nx = 40, the oracle's restatement, pinned to it, elsewhere) bit for bit.
* A three-cycle OSSE run entirely on the device (analysis -> forecast ->
  obs with device H(x) -> analysis) equals the same cycles done with the
  oracle's forecast and host H(x), within 1e-10 of the ensemble spread.
"""
import ctypes

import numpy as np
import pytest

import oracle
from letkf_amd import _lib, das, synth

pytestmark = pytest.mark.gpu


def _background(h, nx, ne):
    """The handle's background buffer (device -> host)."""
    ptrs = [ctypes.c_void_p() for _ in range(4)]
    _lib.check(_lib.lib().lekf_device_state(h.h, *[ctypes.byref(q) for q in ptrs]), "state")
    h.sync()
    out = np.zeros(nx * ne)
    _lib.check(_lib.lib().lekf_memcpy(out.ctypes.data, ptrs[0].value, out.nbytes, 1, None),
               "memcpy")
    return out.reshape(nx, ne)


@pytest.mark.parametrize("nx,ne,kt", [(40, 20, 10), (40, 64, 1), (960, 200, 5), (250, 33, 7)])
def test_forecast_bitwise(gpu, nx, ne, kt):
    p = synth.l96(nx=nx, ne=ne, seed=nx + ne)
    h = das.L96Letkf(nx=nx, ne=ne, xlocal=3.0)
    h.set_ensemble(p["xb"].reshape(nx, 1, ne, 1, order="F"))
    h.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h.run()
    xa = h.analysis()[0][:, 0, :, 0]
    h.forecast(kt, dt=0.005, force=8.0)
    xf = _background(h, nx, ne)
    impl = "ref" if (nx == 40 and oracle.ref_available()) else "c"
    for m in range(ne):
        np.testing.assert_array_equal(xf[:, m], oracle.l96_rk4(xa[:, m], kt, 0.005, 8.0, impl=impl))


def test_device_cycles_equal_host_cycles(gpu):
    nx, ne, kt = 40, 20, 10
    p = synth.l96(nx=nx, ne=ne, seed=12)
    rng = np.random.default_rng(2)
    truth = p["truth"].copy()
    gidx = np.arange(0, nx, 2, dtype=np.int32)
    err = np.ones(gidx.size)
    dev = das.L96Letkf(nx=nx, ne=ne, xlocal=3.0)
    host = das.L96Letkf(nx=nx, ne=ne, xlocal=3.0)
    xb = p["xb"].copy()
    dev.set_ensemble(xb.reshape(nx, 1, ne, 1, order="F"))
    host.set_ensemble(xb.reshape(nx, 1, ne, 1, order="F"))
    for cyc in range(3):
        y = truth[gidx] + rng.standard_normal(gidx.size)
        dev.set_obs(gidx, y, err)                      # H(x) on the device
        host.set_obs(gidx, y, err, xb[gidx, :])
        dev.run()
        host.run()
        xa_d = dev.analysis()[0][:, 0, :, 0]
        xa_h = host.analysis()[0][:, 0, :, 0]
        spread = xa_h.std(axis=1, ddof=1).mean()
        assert np.abs(xa_d - xa_h).max() <= 1e-10 * spread, cyc
        truth = oracle.l96_rk4(truth, kt)
        dev.forecast(kt)
        xb = np.stack([oracle.l96_rk4(xa_h[:, m], kt) for m in range(ne)], 1)
        host.cycle_swap()
        host.set_ensemble(xb.reshape(nx, 1, ne, 1, order="F"))


def test_forecast_arguments(gpu):
    import letkf_amd
    h = das.L96Letkf(nx=40, ne=20)
    with pytest.raises(letkf_amd.LetkfError):
        h.forecast(5)                                  # no ensemble yet
    with pytest.raises(letkf_amd.LetkfError):
        h.set_obs(np.array([0], np.int32), np.zeros(1), np.ones(1))   # device H(x) needs it too
