"""The Fortran host side (ISO_C_BINDING drop-in common_letkf + letkf_tools_gpu).

letkf_amd/fortran/l96_driver runs the reference's Lorenz-96 per-point loop
(lorenz96/letkf/letkf.f90:216-265) with ``USE common_letkf`` resolved to the
GPU drop-in, then the same step through das_letkf_l96; both must equal the
reference Fortran's analysis (golden fixture) to 1e-10.
"""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "letkf_amd", "fortran", "l96_driver")


def test_fortran_l96_drop_in(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("Fortran host side not built (amdflang missing)")
    z = np.load(os.path.join(ROOT, "tests", "golden", "l96_step.npz"), allow_pickle=False)
    nx, ne = z["xb"].shape
    ny = z["gidx"].size
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([nx, ne, ny], dtype=np.int32).tofile(f)
        np.asfortranarray(z["xb"]).T.astype(np.float64).tofile(f)  # column-major (nx, ne)
        z["gidx"].astype(np.int32).tofile(f)
        z["y"].astype(np.float64).tofile(f)
        z["err"].astype(np.float64).tofile(f)
        np.asfortranarray(z["hxf"]).T.astype(np.float64).tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.float64)
    xa1 = out[:nx * ne].reshape(ne, nx).T
    xa2 = out[nx * ne:].reshape(ne, nx).T
    scale = np.abs(z["xa"] - z["xa"].mean(axis=1, keepdims=True)).max()
    assert np.abs(xa1 - z["xa"]).max() / scale < 1e-10
    assert np.abs(xa2 - z["xa"]).max() / scale < 1e-10


SPD = os.path.join(ROOT, "letkf_amd", "fortran", "speedy_driver")


@pytest.fixture(scope="module")
def speedy_prob():
    import oracle
    from letkf_amd import synth
    p = synth.speedy()
    return p, oracle


def _run_speedy_driver(p, cov_infl_mul, infl3d, tmp_path):
    o = p["obs"]
    nij, nlev, ne, nv3d = p["gues3d"].shape
    nv2d = p["gues2d"].shape[2]
    fin = tmp_path / "spd_in.bin"
    fout = tmp_path / "spd_out.bin"
    col = lambda a: np.asfortranarray(a).ravel(order="F").astype(np.float64)  # noqa: E731
    with open(fin, "wb") as f:
        np.array([p["nlon"], p["nlat"], nlev, ne, nv3d, nv2d, o["elm"].size],
                 dtype=np.int32).tofile(f)
        for a in (p["lon"], p["lat"], p["sig"], col(p["gues3d"]), col(p["gues2d"]), o["elm"],
                  o["lon"], o["lat"], o["lev"], o["dat"], o["err"], col(o["hxf"]),
                  np.array([p["sigma_obs"], p["sigma_obsv"], p["gross_error"], cov_infl_mul]),
                  col(infl3d)):
            np.asarray(a, dtype=np.float64).tofile(f)
    r = subprocess.run([SPD, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.float64)
    n3, n2 = nij * nlev * ne * nv3d, nij * ne * nv2d
    a3 = out[:n3].reshape((nij, nlev, ne, nv3d), order="F")
    a2 = out[n3:n3 + n2].reshape((nij, ne, nv2d), order="F")
    infl = out[n3 + n2:].reshape((nij, nlev), order="F")
    return a3, a2, infl


def _scaled(a, ref, ens_axis, var_axis):
    dev = ref - ref.mean(axis=ens_axis, keepdims=True)
    out = 0.0
    for v in range(ref.shape[var_axis]):
        sl = tuple(v if i == var_axis else slice(None) for i in range(ref.ndim))
        out = max(out, np.abs(a[sl] - ref[sl]).max() / max(np.abs(dev[sl]).max(), 1e-300))
    return out


@pytest.mark.parametrize("cov_infl_mul", [-1.01, 1.05])
def test_fortran_speedy_das_letkf(gpu, tmp_path, speedy_prob, cov_infl_mul):
    """das_letkf(gues3d,gues2d,anal3d,anal2d) (speedy/letkf/letkf_tools.f90:45-51)
    through the drop-in modules letkf_obs / letkf_tools, on the SPEEDY-sized
    config, against the oracle's das_letkf restatement: the adaptive field
    (cov_infl_mul <= 0, updated in place) and a constant factor with 1.01 at
    the top level (:113-121)."""
    if not os.path.exists(SPD):
        pytest.skip("Fortran host side not built (amdflang missing)")
    p, oracle = speedy_prob
    nij, nlev = p["nij"], p["nlev"]
    if cov_infl_mul > 0:
        field = np.full((nij, nlev), cov_infl_mul)
        field[:, nlev - 1] = 1.01
        infl_in = np.full((nij, nlev), 7.0)  # ignored for a constant factor
    else:
        field = np.full((nij, nlev), -cov_infl_mul)
        infl_in = field
    a3, a2, infl = _run_speedy_driver(p, cov_infl_mul, infl_in, tmp_path)
    ref = oracle.das_global_problem(p, infl3d=field)
    assert _scaled(a3, ref["anal3d"], 2, 3) < 1e-10
    assert _scaled(a2, ref["anal2d"], 1, 2) < 1e-10
    if cov_infl_mul > 0:
        np.testing.assert_array_equal(infl, infl_in)
    else:
        np.testing.assert_allclose(infl, ref["infl3d"], rtol=1e-12)


OMP = os.path.join(ROOT, "letkf_amd", "fortran", "omp_core_driver")


def _omp_cases(ne, ncase, nobs, seed):
    rng = np.random.default_rng(seed)
    cases = []
    for c in range(ncase):
        p = int(rng.integers(0, nobs + 1)) if c % 7 else 0
        h = rng.normal(size=(nobs, ne))
        h -= h.mean(axis=1, keepdims=True)
        cases.append(dict(p=p, h=h, rd=rng.uniform(0.5, 2.0, nobs), rl=rng.uniform(0.05, 1.0, nobs),
                          dep=rng.normal(size=nobs), parm=1.0 + 0.1 * rng.uniform()))
    return cases


def _run_omp(cases, ne, nobs, threads, tmp_path):
    fin, fout = tmp_path / f"omp_in_{threads}.bin", tmp_path / f"omp_out_{threads}.bin"
    with open(fin, "wb") as f:
        np.array([len(cases), ne, nobs], dtype=np.int32).tofile(f)
        for c in cases:
            np.array([c["p"]], dtype=np.int32).tofile(f)
            for a in (c["h"].ravel(order="F"), c["rd"], c["rl"], c["dep"], [c["parm"]]):
                np.asarray(a, dtype=np.float64).tofile(f)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([OMP, str(fin), str(fout)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, dtype=np.float64).reshape(len(cases), 2 * ne * ne + ne + 1)
    return out


@pytest.mark.parametrize("ne", [20, 40])
def test_fortran_letkf_core_openmp(gpu, tmp_path, ne):
    """letkf_core through the drop-in common_letkf from 8 OpenMP threads at once
    (the reference's PARALLEL DO over grid points, scale/letkf/letkf_tools.f90:175):
    every case equals the single-thread run bit for bit, and the reference
    Fortran letkf_core (oracle/_ref) to 1e-10; k = 20 (fused kernel) and 40
    (three-kernel eigensolver)."""
    if not os.path.exists(OMP):
        pytest.skip("Fortran host side not built (amdflang missing)")
    import oracle
    nobs = 300
    cases = _omp_cases(ne, 96, nobs, seed=ne)
    o8 = _run_omp(cases, ne, nobs, 8, tmp_path)
    o1 = _run_omp(cases, ne, nobs, 1, tmp_path)
    np.testing.assert_array_equal(o8, o1)
    if not oracle.ref_available():
        return
    for c, row in zip(cases, o8):
        p = c["p"]
        ref = oracle.letkf_core(c["h"][:p], c["rd"][:p], c["rl"][:p], c["dep"][:p], c["parm"],
                                transm=True, pao=True, impl="ref")
        tr = row[:ne * ne].reshape(ne, ne, order="F")
        tm = row[ne * ne:ne * ne + ne]
        pa = row[ne * ne + ne:2 * ne * ne + ne].reshape(ne, ne, order="F")
        sc = max(np.abs(ref["trans"]).max(), 1.0)
        assert np.abs(tr - ref["trans"]).max() <= 1e-10 * sc
        assert np.abs(tm - ref["transm"]).max() <= 1e-10 * max(np.abs(ref["transm"]).max(), 1e-3)
        assert np.abs(pa - ref["pao"]).max() <= 1e-10 * max(np.abs(ref["pao"]).max(), 1e-3)
        assert abs(row[-1] - ref["parm_infl"]) <= 1e-10 * abs(ref["parm_infl"])
