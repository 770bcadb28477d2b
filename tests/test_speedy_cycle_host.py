"""Host side of the SPEEDY programs (letkf_amd.speedy_cycle), no GPU:

* the grid layout between files and the handle's (nij, nlev, ne, nv) arrays,
  checked through the reference's read_grd on the files written;
* set_letkf_obs's reading: member files -> one record set, H(x) per member,
  qc the minimum over members;
* the in-memory obs path of ``cycle`` equals writing obs01MMM.dat with the
  reference's write_obs2 and reading it with read_obs2;
* the inflation field entering and leaving das_letkf;
* the ensemble mean and spread against numpy.
"""
import os

import numpy as np
import pytest

import oracle
from letkf_amd import speedy_cycle as sc
from letkf_amd import speedy_io as sio

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
NE = 5


def _ens(seed=0):
    rng = np.random.default_rng(seed)
    g3 = np.asfortranarray(rng.standard_normal((sc.NIJ, 7, NE, 4)))
    g2 = np.asfortranarray(rng.standard_normal((sc.NIJ, NE, 2)))
    return g3, g2


@needs_ref
def test_ensemble_layout_through_reference_reader(tmp_path):
    g3, g2 = _ens()
    f = sc.Files(str(tmp_path))
    sc.write_ensemble(f, g3, g2, big_endian=False)
    for m in (0, NE - 1):
        v3, v2 = oracle.ref_read_grd(f.anal(m))
        i, j, k, v = 17, 31, 4, 2                   # grid point (i, j) is ij = i + 96 j
        assert v3[i, j, k, v] == np.float32(g3[i + 96 * j, k, m, v])
        assert v2[i, j, 1] == np.float32(g2[i + 96 * j, m, 1])
        np.testing.assert_array_equal(v3.reshape(sc.NIJ, 7, 4, order="F"),
                                      g3[:, :, m, :].astype(np.float32))
    for m in range(NE):
        os.rename(f.anal(m), f.gues(m))
    r3, r2 = sc.read_ensemble(f, NE, big_endian=False)
    np.testing.assert_array_equal(r3, g3.astype(np.float32))
    np.testing.assert_array_equal(r2, g2.astype(np.float32))


def _member_records(n, seed):
    rng = np.random.default_rng(seed)
    elem = rng.choice([sio.ID_U, sio.ID_T, sio.ID_Q, sio.ID_PS, sio.ID_RH], n).astype(float)
    return dict(elem=elem, rlon=rng.uniform(0, 360, n), rlat=rng.uniform(-80, 80, n),
                rlev=rng.uniform(2e4, 1e5, n), odat=rng.standard_normal(n),
                oerr=rng.uniform(0.5, 2, n))


@needs_ref
def test_member_obs_and_in_memory_path(tmp_path):
    n = 200
    base = _member_records(n, 1)
    rng = np.random.default_rng(2)
    hx = rng.standard_normal((n, NE)) * 3.0
    qc = np.ones((n, NE), dtype=np.int32)
    qc[rng.choice(n, 15), rng.integers(0, NE, 15)] = 0
    f = sc.Files(str(tmp_path))
    for m in range(NE):
        oracle.ref_write_obs2(f.obs(m), dict(base, ohx=hx[:, m], oqc=qc[:, m]))
    got = sc.read_member_obs(f, NE, big_endian=False)
    ref = [oracle.ref_read_obs2(f.obs(m), n) for m in range(NE)]
    for k in sio.OBS_KEYS:
        np.testing.assert_array_equal(got[k], ref[0][k], err_msg=k)
    np.testing.assert_array_equal(got["hxf"], np.stack([r["ohx"] for r in ref], axis=1))
    np.testing.assert_array_equal(got["qc"], qc.min(axis=1))

    # cycle's path: the values read_obs2 returns for files written by write_obs2,
    # with the all-member qc written to every member file
    qall = qc.min(axis=1).astype(np.int32)
    for m in range(NE):
        oracle.ref_write_obs2(f.obs(m), dict(base, ohx=hx[:, m], oqc=qall))
    ref = [oracle.ref_read_obs2(f.obs(m), n) for m in range(NE)]
    mem = sc._obs_through_files(base, np.asfortranarray(hx), qall)
    for k in sio.OBS_KEYS:
        np.testing.assert_array_equal(mem[k], ref[0][k], err_msg=k)
    np.testing.assert_array_equal(mem["hxf"], np.stack([r["ohx"] for r in ref], axis=1))
    np.testing.assert_array_equal(mem["qc"], qall)


def test_missing_obs_files(tmp_path):
    assert sc.read_member_obs(sc.Files(str(tmp_path)), NE) is None


def test_member_files_must_agree(tmp_path):
    f = sc.Files(str(tmp_path))
    for m, n in enumerate([10, 10, 9, 10, 10]):
        rec = _member_records(n, 5)
        sio.write_obs2(f.obs(m), dict(rec, ohx=np.zeros(n), oqc=np.ones(n, np.int32)))
    with pytest.raises(ValueError):
        sc.read_member_obs(f, NE)


def test_inflation_field(tmp_path):
    f = sc.Files(str(tmp_path))
    i3, i2 = sc.initial_infl(f, -1.01)                     # no file: 1.01 everywhere
    assert np.all(i3 == 1.01) and np.all(i2 == 1.01)
    i3, i2 = sc.initial_infl(f, 1.05)                      # fixed: top level 1.01
    assert np.all(i3[:, :6] == 1.05) and np.all(i3[:, 6] == 1.01) and np.all(i2 == 1.05)
    rng = np.random.default_rng(3)
    upd = np.asfortranarray(1.0 + rng.uniform(0, 0.5, (sc.NIJ, 7)))
    inp2 = np.asfortranarray(np.stack([np.full(sc.NIJ, 1.2), np.full(sc.NIJ, 1.4)], axis=1))
    sc.write_infl(f, upd, inp2)
    v3, v2 = sio.read_grd(f.infl)
    for v in range(4):
        np.testing.assert_array_equal(v3[:, :, :, v].reshape(sc.NIJ, 7, order="F"),
                                      upd.astype(np.float32))
    assert np.all(v2 == np.float32(1.2))
    r3, r2 = sc.initial_infl(f, -1.01)                     # read back: first variable
    np.testing.assert_array_equal(r3, upd.astype(np.float32))
    assert np.all(r2 == np.float32(1.2))


def test_mean_sprd():
    g3, g2 = _ens(4)
    m3, m2, s3, s2 = sc.mean_sprd(g3, g2)
    np.testing.assert_allclose(m3, g3.mean(axis=2), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(s3, g3.std(axis=2, ddof=1), rtol=1e-12)
    np.testing.assert_allclose(m2, g2.mean(axis=1), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(s2, g2.std(axis=1, ddof=1), rtol=1e-12)


def test_command_line_rejects_unknown_program():
    with pytest.raises(SystemExit):
        sc.main(["assimilate", "--ne", "20"])
