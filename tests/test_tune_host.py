"""Host side of the launch tuning: range checks, get/set and the JSON cache
(no kernel runs; the C ABI calls only touch the process-wide tuning table)."""
import json

import pytest

from letkf_amd import _lib, tune


@pytest.fixture(autouse=True)
def _restore():
    saved = tune.current()
    yield
    tune.apply(saved)


def test_defaults_and_roundtrip():
    cur = tune.current()
    assert cur["gram_waves"] in (2, 4, 8)
    assert cur["sel_per_cu"] in (1, 2, 3, 4)
    assert cur["eig_split"] in (0, 1)
    tune.set("gram_waves", 8)
    tune.set("gram_ws_mb", 64)
    tune.set("sel_per_cu", 3)
    tune.set("eig_split", 0)
    assert tune.current() == {"gram_waves": 8, "gram_ws_mb": 64, "sel_per_cu": 3,
                              "eig_split": 0}


@pytest.mark.parametrize("name,bad", [("gram_waves", 3), ("gram_waves", 16), ("gram_ws_mb", 8),
                                      ("sel_per_cu", 0), ("sel_per_cu", 5), ("eig_split", 2),
                                      ("eig_split", -1)])
def test_out_of_range_rejected(name, bad):
    before = tune.get(name)
    with pytest.raises(_lib.LetkfError):
        tune.set(name, bad)
    assert tune.get(name) == before


def test_unknown_param():
    L = _lib.lib()
    assert L.lekf_tune_set(7, 1) == _lib.LEKF_ERR_ARG
    assert L.lekf_tune_get(7) == -_lib.LEKF_ERR_ARG


def test_cache_load(tmp_path):
    path = tmp_path / "t.json"
    path.write_text(json.dumps({"k": {"params": {"gram_waves": 2, "gram_ws_mb": 256,
                                                 "sel_per_cu": 1}, "ms": 1.0}}))
    assert tune.load("missing", str(path)) is None
    assert tune.load("k", str(path)) == {"gram_waves": 2, "gram_ws_mb": 256, "sel_per_cu": 1}
    cur = tune.current()
    assert {n: cur[n] for n in ("gram_waves", "gram_ws_mb", "sel_per_cu")} == \
        {"gram_waves": 2, "gram_ws_mb": 256, "sel_per_cu": 1}


def test_solve_workspace_bytes_follows_the_tuning():
    """lekf_solve_workspace_bytes: none for k <= 32, one record per point up to
    the tuned workspace (at least 1024 points on the split path), then flat."""
    L = _lib.lib()
    assert L.lekf_solve_workspace_bytes(1000, 20) == 0
    assert L.lekf_solve_workspace_bytes(0, 50) == 0
    assert L.lekf_solve_workspace_bytes(1000, 300) == 0        # above LEKF_KMAX
    rec = L.lekf_solve_workspace_bytes(1, 50)
    assert rec > 50 * 50 * 8                                    # a record holds C at least
    assert L.lekf_solve_workspace_bytes(10, 50) == 10 * rec
    tune.set("gram_ws_mb", 1024)
    small = L.lekf_solve_workspace_bytes(10 ** 7, 50)
    assert small // rec == max(1024, (1024 << 20) // rec)
    tune.set("gram_ws_mb", 4096)
    assert L.lekf_solve_workspace_bytes(10 ** 7, 50) == (4096 << 20) // rec * rec
    assert L.lekf_solve_workspace_bytes(10 ** 7, 50) == L.lekf_solve_workspace_bytes(10 ** 8, 50)
