"""Launch tuning leaves the analysis unchanged.

The Gram workspace size and the selection slot count only re-partition the
same per-point work, so the analysis must be bit-identical for every value;
the Gram wave count re-orders the waves' partial-sum reduction and the
three-kernel eigensolver applies the eigenvectors without forming them, so
those agree to rounding (1e-12 of the analysis perturbation scale).
A graph captured before a change is retaken after it, and the autotuner's
choice is one of the candidates, is applied, and lands in its cache.
"""
import json

import numpy as np
import pytest

from letkf_amd import das, serve, synth, tune

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore():
    saved = tune.current()
    yield
    tune.apply(saved)


def _make(job, graph=False):
    h = serve._KINDS[job.kind](**job.cfg)
    h.set_ensemble(**job.ensemble)
    h.set_obs(**job.obs)
    if graph:
        h.set_graph(True)
    return h


def _analysis(job):
    h = _make(job)
    h.run()
    a = h.analysis()
    h.close()
    return a


def _radar():
    return serve.regional_job("r", synth.radar(sub=(32, 28, 8), ne=50, seed=5))


def _global_capped():
    p = synth.global_problem(nlon=48, lat=synth.gaussian_lats(24), sig=synth.SPEEDY_SIG, ne=40,
                             nstation=900, seed=7)
    p.update(sigma_obs=700.0e3, sigma_obsv=0.4, gross_error=10.0, infl=1.0)
    job = serve.global_job("g", p)
    job.cfg["nobs_max"] = 30
    return job


def _eq(a, b):
    for x, y in zip(a, b):
        assert (x is None and y is None) or np.array_equal(x, y)


def _close(a, b, rel):
    for x, y in zip(a, b):
        if x is None:
            continue
        scale = max(float(np.abs(x - x.mean(axis=-2, keepdims=True)).max()), 1e-300)
        assert float(np.abs(x - y).max()) <= rel * scale


@pytest.mark.parametrize("mk", [_radar, _global_capped])
def test_slots_and_workspace_bit_identical(gpu, mk):
    job = mk()
    tune.set("sel_per_cu", 4)
    tune.set("gram_ws_mb", 1024)
    a0 = _analysis(job)
    tune.set("sel_per_cu", 1)
    _eq(_analysis(job), a0)
    tune.set("sel_per_cu", 2)
    tune.set("gram_ws_mb", 16)   # many sub-batches of the Gram/eigen pair
    _eq(_analysis(job), a0)


@pytest.mark.parametrize("mk", [_radar, _global_capped])
def test_eig_split_agrees(gpu, mk):
    """k = 50 and k = 40: the three-kernel eigensolver (Householder, QL with one
    lane per point recording its rotations, then the apply through Q and the
    rotations without forming the eigenvectors) agrees with the fused kernel's
    explicit eigenvectors to rounding; its sub-batching leaves it bit-identical."""
    job = mk()
    tune.set("eig_split", 1)
    a1 = _analysis(job)
    tune.set("eig_split", 0)
    _close(_analysis(job), a1, 1e-12)
    tune.set("eig_split", 1)
    tune.set("gram_ws_mb", 16)  # many sub-batches
    _eq(_analysis(job), a1)


@pytest.mark.parametrize("mk", [_radar, _global_capped])
def test_split_overflow_fallback(gpu, mk):
    """A point whose QL rotation stream outgrows its cap is solved by the fused
    kernel's steps (lekf_eig_fix_kernel) from the Gram still in its record:
    with a cap of one rotation every point with obs takes that path (then the
    analysis is the fused kernel's, bit for bit), with 900 some do; both agree
    with the default cap to rounding."""
    job = mk()
    tune.set("eig_split", 1)
    tune.set("split_rcap", 0)
    a1 = _analysis(job)
    tune.set("eig_split", 0)
    a0 = _analysis(job)
    tune.set("eig_split", 1)
    try:
        tune.set("split_rcap", 1)
        a = _analysis(job)
        _eq(a, a0)
        _close(a, a1, 1e-12)
        tune.set("split_rcap", 900)
        _close(_analysis(job), a1, 1e-12)
    finally:
        tune.set("split_rcap", 0)


@pytest.mark.parametrize("waves", [2, 8])
def test_gram_waves_agree(gpu, waves):
    job = _radar()
    tune.set("gram_waves", 4)
    a0 = _analysis(job)
    tune.set("gram_waves", waves)
    _close(_analysis(job), a0, 1e-12)


def test_graph_retaken_after_tune_change(gpu):
    job = _radar()
    tune.set("gram_ws_mb", 1024)
    a0 = _analysis(job)
    h = _make(job, graph=True)
    h.run()
    h.run()
    _eq(h.analysis(), a0)
    tune.set("gram_ws_mb", 16)   # the graph's launches no longer match
    for _ in range(2):
        h.run()
        _eq(h.analysis(), a0)
    h.close()


def test_autotune_picks_candidate_and_caches(gpu, tmp_path):
    job = _radar()
    h = _make(job)
    path = str(tmp_path / "tune.json")
    tune.apply({"gram_waves": 4, "gram_ws_mb": 1024, "sel_per_cu": 4})
    cands = {"gram_waves": (2, 4), "gram_ws_mb": (64, 1024), "sel_per_cu": (2, 4),
             "eig_split": (1, 0)}
    out = tune.autotune(h, candidates=cands, runs=2, key="test|radar", path=path)
    for n, v in out["params"].items():
        assert v in cands[n]
        assert tune.get(n) == v
    assert out["ms"] > 0 and len(out["trials"]) >= 1
    data = json.load(open(path))
    assert data["test|radar"]["params"] == out["params"]
    tune.set("gram_waves", 8 if out["params"]["gram_waves"] != 8 else 2)
    assert tune.load("test|radar", path) == out["params"]
    assert tune.current() == out["params"]
    h.close()
