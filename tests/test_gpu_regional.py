"""GPU parity of the regional (SCALE-style) path: index, capped selection, solve, apply.

Reference: the oracle's restatement of the SCALE obs index and obs_local
(pinned against brute force in tests/test_oracle_regional.py) for the local
lists, and the reference Fortran letkf_core + apply (oracle/_ref) for the
analysis.  Selection sets must be identical (bit-exact cutoff arithmetic);
the analysis within 1e-10 of each variable's analysis perturbation scale.
"""
import numpy as np
import pytest

import oracle
from letkf_amd import das, synth

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def _problem(ne, seed, nobs_max):
    p = synth.radar(sub=(24, 20, 8), ne=ne, seed=seed)
    p["nobs_max"] = nobs_max
    return p


def _oracle(p):
    nx, ny, nlev, nij = p["nx"], p["ny"], p["nlev"], p["nij"]
    perm, grd = oracle.obs_bins_regional(p["ri"], p["rj"], nx, ny)
    so = {k: np.asarray(p[k])[perm] for k in ("ri", "rj", "z", "err", "bucket")}
    idx_all, rho_all, off, lists = [], [], [0], []
    for b in range(nij * nlev):
        lev, ij = divmod(b, nij)
        idx, rl, _ = oracle.obs_local_regional(float(ij % nx + 1), float(ij // nx + 1),
                                               p["zlev"][lev], nx, ny, grd, so, p["dx"], p["dy"],
                                               p["sigma_h"], p["sigma_v"], p["nobs_max"], 2)
        lists.append((idx, rl))
        idx_all.append(idx)
        rho_all.append(rl)
        off.append(off[-1] + idx.size)
    hd = np.asarray(p["hdxf"])[perm]
    g3 = p["gues3d"]
    xb = np.ascontiguousarray(g3.transpose(1, 0, 3, 2)).reshape(nlev * nij, p["nv3d"], p["ne"])
    xa, infl = oracle.ref_solve_apply(np.array(off), np.concatenate(idx_all),
                                      np.concatenate(rho_all), hd, so["err"] ** 2,
                                      np.asarray(p["dep"])[perm], xb, np.ones(nlev * nij))
    a3 = xa.reshape(nlev, nij, p["nv3d"], p["ne"]).transpose(1, 0, 3, 2)
    return dict(perm=perm, lists=lists, anal3d=a3, infl=infl.reshape(nlev, nij).T, so=so)


def _handle(p, budget=0, prec="f64"):
    h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                          ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"], sigma_v=p["sigma_v"],
                          nobs_max=p["nobs_max"], infl_value=1.0, list_budget=budget, prec=prec)
    h.set_ensemble(p["gues3d"])
    h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
    return h


def _scaled(a, ref):
    dev = ref - ref.mean(axis=2, keepdims=True)
    out = 0.0
    for v in range(ref.shape[3]):
        out = max(out, np.abs(a[..., v] - ref[..., v]).max() / np.abs(dev[..., v]).max())
    return out


@pytest.mark.parametrize("ne,cap", [(12, 25), (12, 0), (50, 40)])
def test_regional_matches_reference(gpu, ne, cap):
    p = _problem(ne, seed=7 + ne, nobs_max=cap)
    ref = _oracle(p)
    h = _handle(p)
    h.run()
    h.sync()
    nl = h.nobsl()
    rng = np.random.default_rng(1)
    for b in rng.choice(p["nij"] * p["nlev"], 60, replace=False):
        idx, rho = h.local(b)
        ridx, rrho = ref["lists"][b]
        assert idx.size == ridx.size
        bk = ref["so"]["bucket"]
        for t in range(2):
            np.testing.assert_array_equal(np.sort(idx[bk[idx] == t]), np.sort(ridx[bk[ridx] == t]))
        if cap > 0:
            assert np.all(np.diff(bk[idx]) >= 0)
            for t in range(2):
                assert np.all(np.diff(idx[bk[idx] == t]) > 0)  # scan order in a bucket
        else:
            np.testing.assert_array_equal(idx, ridx)
            np.testing.assert_allclose(rho, rrho, rtol=1e-13)
        lev, ij = divmod(int(b), p["nij"])
        assert nl[ij, lev] == ridx.size
    a3, _ = h.analysis()
    assert _scaled(a3, ref["anal3d"]) < RTOL
    i3, _ = h.infl()
    np.testing.assert_allclose(i3, ref["infl"], rtol=1e-11)


def test_regional_obs_index(gpu):
    p = _problem(12, seed=11, nobs_max=0)
    p["ri"][:4] = [0.5, p["nx"] + 0.5, 0.51, 1.5]  # edge cells clamp
    qc = np.ones(p["ri"].size, dtype=np.int32)
    qc[5:40] = 0
    perm, grd = oracle.obs_bins_regional(p["ri"], p["rj"], p["nx"], p["ny"], qc=qc)
    h = _handle(p)
    h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"], qc=qc)
    gperm, cum = h.obs_index()
    np.testing.assert_array_equal(gperm, perm)
    np.testing.assert_array_equal(cum, grd[:, 1:])


def test_regional_chunks_equal_single_launch(gpu):
    p = _problem(12, seed=3, nobs_max=20)
    h1 = _handle(p)
    h1.run()
    a1, _ = h1.analysis()
    h2 = _handle(p, budget=3000)  # many chunks
    h2.run()
    a2, _ = h2.analysis()
    assert np.array_equal(a1, a2)
    for b in (0, 5, p["nij"] * p["nlev"] // 2, p["nij"] * p["nlev"] - 1):
        i1, r1 = h1.local(b)
        i2, r2 = h2.local(b)  # re-selects the chunk holding b
        np.testing.assert_array_equal(i1, i2)
        np.testing.assert_array_equal(r1, r2)


def test_regional_exact_ties(gpu):
    """Gridded obs with identical positions: ties go to the earlier obs (strict <)."""
    p = _problem(12, seed=5, nobs_max=15)
    p["ri"] = np.round(p["ri"])
    p["rj"] = np.round(p["rj"])
    p["z"] = np.round(p["z"] / 500.0) * 500.0
    p["err"] = np.full_like(p["err"], 4.0)
    ref = _oracle(p)
    h = _handle(p)
    h.run()
    h.sync()
    for b in range(0, p["nij"] * p["nlev"], 97):
        idx, _ = h.local(b)
        ridx, _ = ref["lists"][b]
        np.testing.assert_array_equal(np.sort(idx), np.sort(ridx))
    a3, _ = h.analysis()
    assert _scaled(a3, ref["anal3d"]) < RTOL


@pytest.mark.parametrize("nsame,split", [(700, False), (500, True)])
def test_regional_many_equal_keys(gpu, nsame, split):
    """More than 256 co-located obs in one bucket.  700 bit-identical rdiag keys:
    the threshold is found by the tie scan in scan order.  500 keys in two
    groups of 250 that differ only in the lowest 3 bits: the last digit pass
    leaves a sortable bin whose every key bit is fixed."""
    p = _problem(12, seed=9, nobs_max=30)
    n = p["ri"].size
    sel = np.arange(0, n, 3)[:nsame]
    p["ri"][sel] = 12.25
    p["rj"][sel] = 10.25
    p["z"][sel] = p["zlev"][3]
    p["bucket"][sel] = 1
    p["err"] = np.where(p["bucket"] == 0, 3.0, 5.0)
    if split:
        p["err"][sel[1::2]] = np.nextafter(5.0, 6.0)
    ref = _oracle(p)
    h = _handle(p)
    h.run()
    h.sync()
    for b in [3 * p["nij"] + 9 * p["nx"] + 11] + list(range(3 * p["nij"], 4 * p["nij"], 13)):
        idx, _ = h.local(b)
        ridx, _ = ref["lists"][b]
        np.testing.assert_array_equal(np.sort(idx), np.sort(ridx))
    a3, _ = h.analysis()
    assert _scaled(a3, ref["anal3d"]) < RTOL
