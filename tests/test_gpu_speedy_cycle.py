"""The SPEEDY programs on the GPU (letkf_amd.speedy_cycle) with the reference's
files, at the T30 size (96 x 48 x 7, u, v, T, q + ps, rain, k = 20).

Inputs are written as the reference's files (little-endian here so that the
reference's own readers, compiled into oracle/_ref, read them back).  Checks:

* ``letkf``: its analysis files equal, after the single-precision write, a
  direct GlobalLetkf run on what the reference readers (read_grd, read_obs2)
  return with qc the minimum over members; a sample of points of that run
  against the reference letkf_core (1e-10 of the analysis spread); the mean
  and spread files; infl_mul.grd as das_letkf writes it (every 3-D variable
  the updated field, every 2-D variable the first 2-D input);
* a second cycle reads the inflation field back;
* ``cycle`` (obsope + letkf in one pass) writes the same bytes as ``obsope``
  followed by ``letkf``;
* cov_infl_mul > 0 (top level 1.01, no inflation file), and no accepted
  obs (analysis = background, no inflation file).
"""
import os

import numpy as np
import pytest

import oracle
from letkf_amd import das, synth
from letkf_amd import speedy_cycle as sc
from letkf_amd import speedy_io as sio

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")]
NE = 20
RTOL = 1e-10


def _problem(seed=41):
    p = synth.global_problem(nlon=96, lat=synth.SPEEDY_LAT, sig=synth.SPEEDY_SIG, ne=NE,
                             nstation=415, seed=seed)
    rng = np.random.default_rng(seed)
    rain = np.abs(rng.standard_normal((p["nij"], NE, 1))) * 1e-3
    p["gues2d"] = np.asfortranarray(np.concatenate([p["gues2d"], rain], axis=2))
    return p


def _write_inputs(d, p, member_qc=True):
    f = sc.Files(d)
    for m in range(NE):
        v3 = p["gues3d"][:, :, m, :].reshape(96, 48, 7, 4, order="F")
        v2 = p["gues2d"][:, m, :].reshape(96, 48, 2, order="F")
        sio.write_grd(f.gues(m), v3, v2, big_endian=False)
    o = p["obs"]
    n = o["elm"].size
    rng = np.random.default_rng(3)
    bad = rng.choice(n, 40, replace=False)
    for m in range(NE):
        qc = np.ones(n, dtype=np.int32)
        if member_qc:
            qc[bad[m::NE]] = 0           # a member rejects a few records
        rec = dict(elem=o["elm"], rlon=o["lon"], rlat=o["lat"], rlev=o["lev"], odat=o["dat"],
                   oerr=o["err"], ohx=o["hxf"][:, m], oqc=qc)
        sio.write_obs2(f.obs(m), rec, big_endian=False)
    return f


def _ref_inputs(f):
    """What the reference's readers return."""
    g3 = np.zeros((sc.NIJ, 7, NE, 4), order="F")
    g2 = np.zeros((sc.NIJ, NE, 2), order="F")
    for m in range(NE):
        v3, v2 = oracle.ref_read_grd(f.gues(m))
        g3[:, :, m, :] = v3.reshape(sc.NIJ, 7, 4, order="F")
        g2[:, m, :] = v2.reshape(sc.NIJ, 2, order="F")
    n = oracle.ref_get_nobs(f.obs(0), 8)
    hx = np.zeros((n, NE), order="F")
    qc = np.zeros((n, NE), dtype=np.int32)
    for m in range(NE):
        r = oracle.ref_read_obs2(f.obs(m), n)
        hx[:, m] = r["ohx"]
        qc[:, m] = r["oqc"]
    obs = dict(elm=r["elem"], lon=r["rlon"], lat=r["rlat"], lev=r["rlev"], dat=r["odat"],
               err=r["oerr"], hxf=hx)
    return g3, g2, obs, qc.min(axis=1)


def _direct(g3, g2, obs, qc, i3, i2):
    h = das.GlobalLetkf(lon=sc.LON, lat=sc.LAT, sig=sc.SIG, ne=NE, nv3d=4, nv2d=2, ips=0, iq=3,
                        q_top_lev=5, sigma_obs=sc.SIGMA_OBS, sigma_obsv=sc.SIGMA_OBSV,
                        gross_error=sc.GROSS_ERROR, infl_mode="adaptive", infl_value=1.0)
    h.set_ensemble(g3, g2)
    h.set_obs(obs["elm"], obs["lon"], obs["lat"], obs["lev"], obs["dat"], obs["err"], obs["hxf"],
              qc)
    h.set_infl(i3, i2)
    h.run()
    return h


def _ref_anal(f):
    a3 = np.zeros((sc.NIJ, 7, NE, 4), order="F")
    a2 = np.zeros((sc.NIJ, NE, 2), order="F")
    for m in range(NE):
        v3, v2 = oracle.ref_read_grd(f.anal(m))
        a3[:, :, m, :] = v3.reshape(sc.NIJ, 7, 4, order="F")
        a2[:, m, :] = v2.reshape(sc.NIJ, 2, order="F")
    return a3, a2


def _f32(a):
    return a.astype(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def first_cycle(tmp_path_factory):
    d = tmp_path_factory.mktemp("speedy")
    p = _problem()
    f = _write_inputs(str(d), p)
    print(f"letkf program, T30 k={NE}, dir {d}", flush=True)
    r = sc.letkf(str(d), ne=NE, big_endian=False)
    return p, f, r


def test_letkf_program_matches_direct_run(gpu, first_cycle):
    p, f, r = first_cycle
    g3, g2, obs, qc = _ref_inputs(f)
    assert r["nobs"] == obs["elm"].size and 0 < r["nobs_assim"] < r["nobs"]
    i3 = np.full((sc.NIJ, 7), 1.01, order="F")
    i2 = np.full((sc.NIJ, 2), 1.01, order="F")
    h = _direct(g3, g2, obs, qc, i3, i2)
    d3, d2 = h.analysis()
    a3, a2 = _ref_anal(f)
    np.testing.assert_array_equal(a3, _f32(d3))
    np.testing.assert_array_equal(a2, _f32(d2))
    assert not np.array_equal(a3, g3)

    # the direct run against the reference letkf_core at sampled points (u, v, T)
    sobs, _, _ = oracle.set_letkf_obs_global(obs, sc.LON, sc.LAT, sc.GROSS_ERROR, qc=qc)
    assert sobs["elm"].size == r["nobs_assim"]
    rng = np.random.default_rng(9)
    sample = rng.choice(sc.NIJ * 7, 12, replace=False)
    lists = [h.local(int(b)) for b in sample]
    off = np.cumsum([0] + [l[0].size for l in lists])
    xb = np.stack([g3[int(b) % sc.NIJ, int(b) // sc.NIJ, :, :3].T for b in sample])
    infl = np.full(len(sample), 1.01)
    ref, infl_ref = oracle.ref_solve_apply(off, np.concatenate([l[0] for l in lists]),
                                           np.concatenate([l[1] for l in lists]),
                                           np.ascontiguousarray(sobs["hdxf"].T), sobs["err"] ** 2,
                                           sobs["dep"], np.ascontiguousarray(xb), infl)
    hi3, _ = h.infl()
    for s, b in enumerate(sample):
        ij, lev = int(b) % sc.NIJ, int(b) // sc.NIJ
        got = d3[ij, lev, :, :3].T
        dev = ref[s] - ref[s].mean(axis=1, keepdims=True)
        for v in range(3):
            assert np.abs(got[v] - ref[s, v]).max() <= RTOL * np.abs(dev[v]).max(), (int(b), v)
        assert abs(hi3[ij, lev] - infl_ref[s]) <= 1e-12 * infl_ref[s]

    # mean and spread files (float32; the analysis ones from the float64 ensemble)
    for kind, x3, x2 in (("gues", g3, g2), ("anal", a3, a2)):
        m3, m2 = oracle.ref_read_grd(f.mean(kind))
        s3, s2 = oracle.ref_read_grd(f.sprd(kind))
        mean3 = x3.mean(axis=2).reshape(96, 48, 7, 4, order="F")
        sp3 = x3.std(axis=2, ddof=1).reshape(96, 48, 7, 4, order="F")
        for v in range(4):
            scale = np.abs(x3[:, :, :, v]).max()
            assert np.abs(m3[..., v] - mean3[..., v]).max() <= 4e-7 * scale, (kind, v)
            assert np.abs(s3[..., v] - sp3[..., v]).max() <= 4e-7 * scale, (kind, v)
        mean2 = x2.mean(axis=1).reshape(96, 48, 2, order="F")
        sp2 = x2.std(axis=1, ddof=1).reshape(96, 48, 2, order="F")
        for v in range(2):
            scale = np.abs(x2[:, :, v]).max()
            assert np.abs(m2[..., v] - mean2[..., v]).max() <= 4e-7 * scale, (kind, v)
            assert np.abs(s2[..., v] - sp2[..., v]).max() <= 4e-7 * scale, (kind, v)

    # infl_mul.grd: every 3-D variable = the updated field, 2-D = first 2-D input
    v3, v2 = oracle.ref_read_grd(f.infl)
    want = _f32(hi3).reshape(96, 48, 7, order="F")
    for v in range(4):
        np.testing.assert_array_equal(v3[:, :, :, v], want)
    np.testing.assert_array_equal(v2, _f32(np.full((96, 48, 2), 1.01)))
    assert np.any(hi3 != 1.01)


def test_second_cycle_reads_inflation(gpu, first_cycle, tmp_path):
    p, f0, _ = first_cycle
    f = _write_inputs(str(tmp_path), p)
    v3, v2 = oracle.ref_read_grd(f0.infl)
    # a distinct field per 3-D variable: only the first one enters the solve
    v3 = v3.copy()
    v3[:, :, :, 1:] = 1.3
    v2 = v2.copy()
    v2[:, :, 1] = 1.7
    oracle.ref_write_grd(f.infl, v3, v2)
    sc.letkf(str(tmp_path), ne=NE, big_endian=False)
    g3, g2, obs, qc = _ref_inputs(f)
    i3 = np.asfortranarray(v3[:, :, :, 0].reshape(sc.NIJ, 7, order="F"))
    i2 = np.asfortranarray(v2.reshape(sc.NIJ, 2, order="F"))
    h = _direct(g3, g2, obs, qc, i3, i2)
    d3, _ = h.analysis()
    a3, _ = _ref_anal(f)
    np.testing.assert_array_equal(a3, _f32(d3))
    w3, w2 = oracle.ref_read_grd(f.infl)
    np.testing.assert_array_equal(w3[:, :, :, 3], _f32(h.infl()[0]).reshape(96, 48, 7, order="F"))
    np.testing.assert_array_equal(w2[:, :, 1], v2[:, :, 0])       # rain takes ps's field


def _obsin(p, d):
    """obsin.dat in file units, with records obsope rejects: above the model top,
    below the surface (T), and a missing ps."""
    o = p["obs"]
    lev = o["lev"].copy()
    dat = o["dat"].copy()
    e = o["elm"].astype(int)
    idx3 = np.flatnonzero(e == synth.ID_T)
    lev[idx3[:5]] = 500.0                     # 5 hPa: above sigma 0.08
    lev[idx3[5:10]] = 1.2e5                   # below the lowest level
    idx_ps = np.flatnonzero(e == synth.ID_PS)
    dat[idx_ps[:3]] = -999.0e2                # missing ps (odat < -100 after reading)
    wk = np.stack([o["elm"], o["lon"], o["lat"], np.where(e == synth.ID_PS, lev, lev * 0.01),
                   np.where(e == synth.ID_PS, dat * 0.01, dat),
                   np.where(e == synth.ID_PS, o["err"] * 0.01, o["err"])], axis=1)
    sio.write_records(os.path.join(d, "obsin.dat"), wk.astype(np.float32), big_endian=False)
    return np.concatenate([idx3[:10], idx_ps[:3]])


def test_cycle_equals_obsope_then_letkf(gpu, tmp_path):
    p = _problem(seed=43)
    da, db = tmp_path / "a", tmp_path / "b"
    da.mkdir()
    db.mkdir()
    fa = _write_inputs(str(da), p)
    fb = sc.Files(str(db))
    for m in range(NE):                        # the same background, no member obs files
        os.link(fa.gues(m), fb.gues(m))
        os.remove(fa.obs(m))
    planted = _obsin(p, str(da))
    _obsin(p, str(db))
    ro = sc.obsope(str(da), ne=NE, big_endian=False)
    assert 0 < ro["accepted"] <= ro["nobs"] - 13
    ra = sc.letkf(str(da), ne=NE, big_endian=False)
    rb = sc.cycle(str(db), ne=NE, big_endian=False)
    assert ra["nobs_assim"] == rb["nobs_assim"] > 0
    for m in range(NE):
        assert open(fa.anal(m), "rb").read() == open(fb.anal(m), "rb").read(), m
    for name in ("infl_mul.grd", "anal_me.grd", "anal_sp.grd", "gues_me.grd", "gues_sp.grd"):
        assert open(fa.path(name), "rb").read() == open(fb.path(name), "rb").read(), name
    # obsope's files read by the reference: H(x) 0 and qc 0 on the records every
    # member rejects, H(x) set on the records every member accepts
    for m in (0, NE - 1):
        r = oracle.ref_read_obs2(fa.obs(m), ro["nobs"])
        assert np.all(r["oqc"][planted] == 0) and np.all(r["ohx"][planted] == 0.0)
        assert np.all(r["ohx"][r["oqc"] == 1] != 0.0)
        assert (r["oqc"] == 1).sum() == ro["accepted"]


def test_fixed_inflation_and_no_obs(gpu, tmp_path):
    p = _problem(seed=45)
    f = _write_inputs(str(tmp_path), p, member_qc=False)
    sc.letkf(str(tmp_path), ne=NE, cov_infl_mul=1.05, big_endian=False)
    assert not os.path.exists(f.infl)
    g3, g2, obs, qc = _ref_inputs(f)
    i3 = np.full((sc.NIJ, 7), 1.05, order="F")
    i3[:, 6] = 1.01
    h = _direct(g3, g2, obs, qc, i3, np.full((sc.NIJ, 2), 1.05, order="F"))
    d3, d2 = h.analysis()
    a3, a2 = _ref_anal(f)
    np.testing.assert_array_equal(a3, _f32(d3))
    np.testing.assert_array_equal(a2, _f32(d2))

    # every record rejected by some member: analysis = background, no inflation file
    o = p["obs"]
    for m in range(NE):
        rec = dict(elem=o["elm"], rlon=o["lon"], rlat=o["lat"], rlev=o["lev"], odat=o["dat"],
                   oerr=o["err"], ohx=o["hxf"][:, m],
                   oqc=np.full(o["elm"].size, 0 if m == 7 else 1, dtype=np.int32))
        sio.write_obs2(f.obs(m), rec, big_endian=False)
    r = sc.letkf(str(tmp_path), ne=NE, big_endian=False)
    assert r["nobs_assim"] == 0 and not os.path.exists(f.infl)
    for m in range(NE):
        assert open(f.anal(m), "rb").read() == open(f.gues(m), "rb").read()
