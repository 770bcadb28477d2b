"""The FP64 redo of ill-conditioned points in the mixed forms.

The FP32 and split-BF16x3 Grams perturb the analysis roughly in proportion to
the square of the condition number of C + (k-1)/alpha I
(tests/test_mixed_emulation.py).  On the three-kernel eigensolver
(32 < k <= 64) the QL kernel flags a point whose condition number exceeds the
mixed_kappa tuning value (16 times it for FP32, whose Gram rounds finer) and
lists it; after the launch's last sub-batch one FP64 pass solves the listed
points again from an FP64 Gram, and their status gets LEKF_PT_FP64.  The redo runs the FP64
form's own kernels, so a redone point equals the FP64 form bit for bit and
every other point the plain mixed form bit for bit (the reference has one
precision; common/common_letkf.f90:101-118 is the product the mixed Gram
replaces).
"""
import numpy as np
import pytest

from letkf_amd import _lib, core, das, synth, tune

pytestmark = pytest.mark.gpu
K = 50
P = 600


@pytest.fixture
def knobs():
    old = {n: tune.get(n) for n in ("mixed_kappa", "split_rcap", "gram_ws_mb")}
    yield
    for n, v in old.items():
        tune.set(n, v)


def _points(rng, npt):
    """Half the points with rows spanned by 4 ensemble modes and scaled up
    (ill-conditioned C), half with independent rows (condition number about 3)."""
    Y, R, rho, kappa = [], [], [], []
    for b in range(npt):
        if b % 2:
            y = rng.standard_normal((P, 4)) @ rng.standard_normal((4, K))
            y += 0.01 * rng.standard_normal((P, K))
            y *= 10.0 ** rng.uniform(0.0, 1.0)  # condition numbers about 400 .. 40000
        else:
            y = rng.standard_normal((P, K))
        y -= y.mean(axis=1, keepdims=True)
        r = rng.uniform(0.5, 2.0, P)
        w = rng.uniform(0.05, 1.0, P)
        c = (y * (w / r)[:, None]).T @ y + (K - 1) * np.eye(K)
        lam = np.linalg.eigvalsh(c)
        Y.append(y)
        R.append(r)
        rho.append(w)
        kappa.append(lam[-1] / lam[0])
    return np.concatenate(Y), np.concatenate(R), np.concatenate(rho), np.array(kappa)


def _batch(prec, Y, R, D, rho, npt):
    from letkf_amd.device import DeviceArray, sync
    t = DeviceArray.from_numpy
    dY, dY32, dR, dD = t(Y), t(Y, np.float32), t(R), t(D)
    dOff = t(np.arange(npt + 1) * P, np.int64)
    dIdx, dRho = t(np.arange(npt * P), np.int32), t(rho)
    trans = DeviceArray.zeros(npt * K * K)
    infl_out = DeviceArray.zeros(npt)
    status = DeviceArray.zeros(npt, np.int32)
    job = core.make_job(k=K, npt=npt, hdxf=dY, hdxf32=dY32 if prec != "f64" else None,
                        rdiag=dR, dep=dD, loc_off=dOff, loc_idx=dIdx, loc_rho=dRho,
                        infl_fixed=1.0, infl_out=infl_out, trans=trans, status=status)
    core.letkf_core_batch([job], prec=prec)
    sync()
    return trans.numpy().reshape(npt, K * K), infl_out.numpy(), status.numpy()


@pytest.mark.parametrize("prec,thr,eff", [("f32", 125, 2000), ("bf16x3", 2000, 2000)])
def test_batch_redo_follows_the_condition_number(gpu, knobs, prec, thr, eff):
    """trans output (the eigenvector kernel): the points above the threshold are
    the FP64 solve's bits, the others the plain mixed solve's."""
    npt = 32
    rng = np.random.default_rng(17)
    Y, R, rho, kappa = _points(rng, npt)
    D = rng.standard_normal(npt * P)
    tune.set("mixed_kappa", 0)
    t64, i64, s64 = _batch("f64", Y, R, D, rho, npt)
    tmix, imix, smix = _batch(prec, Y, R, D, rho, npt)
    tune.set("mixed_kappa", thr)
    tg, ig, sg = _batch(prec, Y, R, D, rho, npt)
    assert not (s64 & _lib.LEKF_PT_FP64).any() and not (smix & _lib.LEKF_PT_FP64).any()
    red = (sg & _lib.LEKF_PT_FP64) != 0
    sure = np.abs(kappa / eff - 1.0) > 0.01
    print(f"{prec}: kappa {kappa.min():.0f} .. {kappa.max():.0f}, {int(red.sum())} of {npt} redone",
          flush=True)
    np.testing.assert_array_equal(red[sure], (kappa > eff)[sure])
    assert red.any() and (~red).any()
    assert np.array_equal(tg[red], t64[red]) and np.array_equal(ig[red], i64[red])
    assert np.array_equal(tg[~red], tmix[~red]) and np.array_equal(ig[~red], imix[~red])
    # what the redo removes: the mixed error on the ill-conditioned points
    scale = np.abs(t64).max(axis=1)
    e_red = (np.abs(tmix - t64).max(axis=1) / scale)[red].max()
    e_keep = (np.abs(tmix - t64).max(axis=1) / scale)[~red].max()
    print(f"{prec}: mixed error of the redone points {e_red:.2e}, of the others {e_keep:.2e}",
          flush=True)
    assert e_red > e_keep


@pytest.fixture(scope="module")
def radar40():
    return synth.radar(sub=(40, 40, 60), ne=50, seed=21)


def _radar_run(p, prec):
    h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                          ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"],
                          sigma_v=p["sigma_v"], nobs_max=1000, infl_mode="fixed",
                          infl_value=1.0, prec=prec)
    h.set_ensemble(p["gues3d"])
    h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
    h.run()
    h.sync()
    a3 = h.analysis()[0]
    n = h.fp64_count()
    nobs_pts = int((h.nobsl() > 0).sum())
    h.close()
    return a3.reshape(a3.shape[0] * a3.shape[1], -1), n, nobs_pts


@pytest.mark.parametrize("rcap", [0, 8])
def test_radar_redo_is_the_fp64_analysis(gpu, knobs, radar40, rcap):
    """Threshold 1 redoes every point with local obs: the BF16x3 handle's analysis
    is then the FP64 handle's, bit for bit, on the eigenvector-free apply and
    (rotation streams capped at 8) on the overflow fallback."""
    tune.set("split_rcap", rcap)
    tune.set("mixed_kappa", 0)
    a64, n64, _ = _radar_run(radar40, "f64")
    tune.set("mixed_kappa", 1)
    a, n, nobs_pts = _radar_run(radar40, "bf16x3")
    assert n64 == 0 and n == nobs_pts > 0
    assert np.array_equal(a, a64)


def test_radar_partial_redo(gpu, knobs, radar40):
    """Thresholds inside the radar lists' range of condition numbers: each point is
    either the FP64 or the plain BF16x3 analysis, bit for bit, the count of FP64
    points is the count of points the redo changed, and some threshold splits
    the points (waves of the apply kernel with both kinds)."""
    tune.set("mixed_kappa", 0)
    a64, _, _ = _radar_run(radar40, "f64")
    amix, _, nobs_pts = _radar_run(radar40, "bf16x3")
    split = False
    for thr in (10, 30, 100, 300, 1000):
        tune.set("mixed_kappa", thr)
        a, n, _ = _radar_run(radar40, "bf16x3")
        eq64 = (a == a64).all(axis=1)
        eqm = (a == amix).all(axis=1)
        print(f"threshold {thr}: {n} of {nobs_pts} points redone", flush=True)
        assert (eq64 | eqm).all()
        assert int((eq64 & ~eqm).sum()) == n
        split |= 0 < n < nobs_pts
    assert split


def test_radar_redo_list_overflow(gpu, knobs, radar40):
    """More flagged points than the list holds (one sub-batch: 1024 points at a
    16 MiB workspace): exactly that many are solved again in FP64, the others
    keep the BF16x3 solution, and each point is one or the other bit for bit."""
    L = _lib.lib()
    tune.set("mixed_kappa", 0)
    a64, _, _ = _radar_run(radar40, "f64")
    amix, _, nobs_pts = _radar_run(radar40, "bf16x3")
    tune.set("gram_ws_mb", 16)
    tune.set("mixed_kappa", 1)
    a, n, _ = _radar_run(radar40, "bf16x3")
    per = L.lekf_solve_workspace_bytes(a.shape[0], K) // L.lekf_solve_workspace_bytes(1, K)
    print(f"list capacity {per}, {n} of {nobs_pts} points redone", flush=True)
    assert per < nobs_pts and n == per
    eq64 = (a == a64).all(axis=1)
    eqm = (a == amix).all(axis=1)
    assert (eq64 | eqm).all()
    assert int((eq64 & ~eqm).sum()) == n


def test_redo_in_graph_replay(gpu, knobs, radar40):
    """The FP64 pass captured in the cycle's HIP graph (the list emptied by a
    kernel, not a memset): every replay equals the eager cycle bit for bit,
    with the same points redone (about 40% of them at threshold 30)."""
    p = radar40

    def make():
        h = das.RegionalLetkf(nx=p["nx"], ny=p["ny"], zlev=p["zlev"], dx=p["dx"], dy=p["dy"],
                              ne=p["ne"], nv3d=p["nv3d"], sigma_h=p["sigma_h"],
                              sigma_v=p["sigma_v"], nobs_max=1000, infl_mode="fixed",
                              infl_value=1.0, prec="bf16x3")
        h.set_ensemble(p["gues3d"])
        h.set_obs(p["ri"], p["rj"], p["z"], p["err"], p["dep"], p["bucket"], p["hdxf"])
        return h

    n0 = 0
    for thr in (30, 10, 1):
        tune.set("mixed_kappa", thr)
        ref = make()
        ref.run()
        ref.sync()
        a0, n0 = ref.analysis()[0], ref.fp64_count()
        ref.close()
        if n0 > 0:
            break
    print(f"threshold {thr}: {n0} points redone", flush=True)
    assert n0 > 0
    h = make()
    h.set_graph(True)
    h.run()                      # eager + capture
    h.sync()
    assert np.array_equal(h.analysis()[0], a0)
    for _ in range(3):
        h.run()                  # replay
        h.sync()
        assert np.array_equal(h.analysis()[0], a0)
        assert h.fp64_count() == n0
    h.close()
