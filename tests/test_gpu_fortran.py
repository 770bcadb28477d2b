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
