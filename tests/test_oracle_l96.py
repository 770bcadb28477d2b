"""Pins of the oracle's Lorenz-96 model (orc_l96_rk4,
lorenz96/model/lorenz96.f90:93-128,290-309): the reference model itself
(compiled from its sources into oracle/_ref, nx = 40) bit for bit, the fixed
point x_i = F, fourth-order convergence in the step size, and the numpy
generator of letkf_amd.synth (the same scheme, dt applied in another order,
so equal to rounding)."""
import numpy as np
import pytest

import oracle
from letkf_amd import synth


def _x0(nx, seed=1):
    rng = np.random.default_rng(seed)
    return 8.0 + rng.standard_normal(nx)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("kt", [1, 10, 400])
def test_matches_reference_model_bitwise(kt):
    x = _x0(40)
    np.testing.assert_array_equal(oracle.l96_rk4(x, kt, 0.005, 8.0, impl="c"),
                                  oracle.l96_rk4(x, kt, 0.005, 8.0, impl="ref"))
    np.testing.assert_array_equal(oracle.l96_rk4(x, kt, 0.01, 6.5, impl="c"),
                                  oracle.l96_rk4(x, kt, 0.01, 6.5, impl="ref"))


def test_fixed_point():
    x = np.full(96, 8.0)
    np.testing.assert_array_equal(oracle.l96_rk4(x, 50, 0.01, 8.0), x)


def test_fourth_order_convergence():
    x = _x0(60, seed=3)
    t = 0.2
    ref = oracle.l96_rk4(x, 3200, t / 3200)
    e1 = np.abs(oracle.l96_rk4(x, 40, t / 40) - ref).max()
    e2 = np.abs(oracle.l96_rk4(x, 80, t / 80) - ref).max()
    assert 12.0 < e1 / e2 < 20.0          # 2^4 = 16


def test_matches_synth_generator():
    x = _x0(40, seed=5)
    np.testing.assert_allclose(oracle.l96_rk4(x, 25, 0.01, 8.0),
                               synth.l96_rk4(x.copy(), 8.0, 0.01, 25), rtol=0, atol=1e-12)
