"""Invariants of the SPEEDY level rules (speedy/letkf/letkf_tools.f90:172-199),
checked on the GPU analysis without the oracle:

* q at ilev >= 5 is not analysed: its analysis is the background (mean plus
  perturbation, so equal to rounding), while q below does change;
* the 2-D variable is updated at ilev = 1 with the transform of the 3-D
  variable its localisation row maps to (var_local_n2n): with a uniform
  localisation that is u's level-1 transform, so a ps background that is an
  affine function of level-1 u, member by member, gets the same affine
  function of u's analysis.
"""
import numpy as np
import pytest

from letkf_amd import das, synth

pytestmark = pytest.mark.gpu


def test_upper_q_untouched_and_ps_follows_level_one(gpu):
    p = synth.speedy()
    g3 = p["gues3d"]
    g2 = p["gues2d"].copy()
    c, s = 1.0e5, 50.0
    g2[:, :, 0] = c + s * g3[:, 0, :, 0]
    h = das.GlobalLetkf(lon=p["lon"], lat=p["lat"], sig=p["sig"], ne=p["ne"], nv3d=4, nv2d=1,
                        ips=0, iq=3, q_top_lev=5, sigma_obs=p["sigma_obs"],
                        sigma_obsv=p["sigma_obsv"], gross_error=p["gross_error"],
                        infl_mode="fixed", infl_value=1.01)
    h.set_ensemble(g3, g2)
    o = p["obs"]
    h.set_obs(o["elm"], o["lon"], o["lat"], o["lev"], o["dat"], o["err"], o["hxf"])
    h.run()
    a3, a2 = h.analysis()
    h.close()
    q_b, q_a = g3[:, :, :, 3], a3[:, :, :, 3]
    eps = np.finfo(np.float64).eps
    upper = np.abs(q_a[:, 4:] - q_b[:, 4:])
    assert (upper <= 4 * eps * np.abs(q_b[:, 4:]).max()).all()
    spread_q = q_b[:, :4].std(axis=2, ddof=1).max()
    assert np.abs(q_a[:, :4] - q_b[:, :4]).max() > 1e-3 * spread_q   # analysed below
    # ps = c + s u(level 1), member by member, before and after
    ps_pred = c + s * a3[:, 0, :, 0]
    dev = np.abs(a2[:, :, 0] - a2[:, :, 0].mean(axis=1, keepdims=True)).max()
    assert np.abs(a2[:, :, 0] - ps_pred).max() <= 1e-9 * dev
    assert np.abs(a2[:, :, 0] - g2[:, :, 0]).max() > 1e-3 * dev       # ps was analysed
