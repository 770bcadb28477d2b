"""The served analysis stream: continuous batching, residency, budget, identity.

A job's analysis must be bit for bit the same alone, in any mix of other
jobs, and as a plain handle run (a local analysis depends only on its own
inputs).  Reference for the numbers themselves: the handle path, which the
other GPU tests pin against the reference letkf_core.
"""
import threading
import time

import numpy as np
import pytest

from letkf_amd import das, serve, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def problems():
    return dict(l96=synth.l96(nx=40, ne=20), l96b=synth.l96(nx=320, ne=40, seed=5),
                speedy=synth.speedy(), radar=synth.radar(sub=(24, 20, 8), ne=50, seed=3))


def _direct(job):
    h = serve._KINDS[job.kind](**job.cfg)
    h.set_ensemble(**job.ensemble)
    h.set_obs(**job.obs)
    h.run()
    a = h.analysis()
    h.close()
    return a


def _equal(a, b):
    for x, y in zip(a, b):
        if x is None or y is None:
            assert x is None and y is None
        else:
            assert np.array_equal(x, y)


def _jobs(problems, tag, prec="f64"):
    return [serve.l96_job(f"l96-{tag}", problems["l96"], prec),
            serve.l96_job(f"l96b-{tag}", problems["l96b"], prec),
            serve.global_job(f"speedy-{tag}", problems["speedy"], prec),
            serve.regional_job(f"radar-{tag}", problems["radar"], prec)]


@pytest.mark.parametrize("prec", ["f64", "bf16x3"])
def test_alone_equals_mix_equals_direct(gpu, problems, prec):
    jobs = _jobs(problems, "a", prec)
    direct = [_direct(j) for j in jobs]
    alone = []
    for j in jobs:
        s = serve.Server(streams=1)
        alone.append(s.submit(j).result(timeout=120).analysis)
        s.close()
    s = serve.Server(streams=2)
    futs = [s.submit(j) for j in _jobs(problems, "m", prec) + _jobs(problems, "n", prec)]
    mixed = [f.result(timeout=120) for f in futs]
    s.close()
    assert max(r.batch_jobs for r in mixed) > 1  # really batched together
    for i, j in enumerate(jobs):
        _equal(alone[i], direct[i])
        _equal(mixed[i].analysis, direct[i])
        _equal(mixed[i + len(jobs)].analysis, direct[i])


def test_resident_cycles_continue_from_analysis(gpu, problems):
    p = problems["l96"]
    s = serve.Server()
    r1 = s.submit(serve.l96_job("m", p)).result(timeout=60)
    r2 = s.submit(serve.l96_job("m", p, with_ensemble=False)).result(timeout=60)
    s.close()
    h = das.L96Letkf(nx=p["nx"], ne=p["ne"], xlocal=p["xlocal"], infl_value=p["infl"])
    h.set_ensemble(np.asfortranarray(p["xb"]).reshape(p["nx"], 1, p["ne"], 1, order="F"))
    h.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h.run()
    a1 = h.analysis()
    h.cycle_swap()
    h.set_obs(p["gidx"], p["y"], p["err"], p["hxf"])
    h.run()
    a2 = h.analysis()
    _equal(r1.analysis, a1)
    _equal(r2.analysis, a2)
    assert not np.array_equal(a1[0], a2[0])


def test_continuous_arrivals_and_budget(gpu, problems):
    """Jobs arrive while others run; a small HBM budget forces evictions."""
    tmpl = _jobs(problems, "x")
    need = max(serve.estimate_bytes(j) for j in tmpl)
    s = serve.Server(hbm_budget=int(need * 2.5), streams=2)
    futs = []

    def feeder():
        for i in range(16):
            j = tmpl[i % len(tmpl)]
            # a fresh instance per job (a repeated key would continue its adaptive inflation)
            futs.append(s.submit(serve.Job(f"{j.key}-{i}", j.kind, j.cfg, j.obs, j.ensemble)))
            time.sleep(0.01)

    t = threading.Thread(target=feeder)
    t.start()
    t.join()
    s.drain(timeout=300)
    res = [f.result() for f in futs]
    assert s.resident_bytes() <= s.budget
    st = s.stats()
    s.close()
    assert st["jobs"] == 16
    assert st["p99_s"] >= st["p50_s"] > 0
    direct = {j.key: _direct(j) for j in tmpl}
    for i, r in enumerate(res):
        _equal(r.analysis, direct[tmpl[i % len(tmpl)].key])
