"""HIP-graph replay of the analysis cycle equals the eager cycle.

Each replay is checked against the eager result, with other eager GPU work
and synchronisations between replays; a change of inputs (cycle swap and
new obs) rebuilds the graph.
"""
import numpy as np
import pytest

from letkf_amd import serve, synth

pytestmark = pytest.mark.gpu


def _problems():
    return {"l96": serve.l96_job("a", synth.l96(nx=80, ne=20)),
            "l96k40": serve.l96_job("b", synth.l96(nx=60, ne=40, seed=4)),
            "speedy": serve.global_job("c", synth.speedy()),
            "radar": serve.regional_job("d", synth.radar(sub=(24, 20, 8), ne=50, seed=3))}


def _make(job):
    h = serve._KINDS[job.kind](**job.cfg)
    h.set_ensemble(**job.ensemble)
    h.set_obs(**job.obs)
    return h


def _same(a, b):
    for x, y in zip(a, b):
        assert (x is None and y is None) or np.array_equal(x, y)


@pytest.mark.parametrize("name", ["l96", "l96k40", "speedy", "radar"])
def test_graph_replays_equal_eager(gpu, name):
    jobs = _problems()
    job = jobs[name]
    ref = _make(job)
    ref.run()
    a0 = ref.analysis()
    other = _make(jobs["l96"])
    h = _make(job)
    h.set_graph(True)
    h.run()                      # eager + capture
    _same(h.analysis(), a0)
    for _ in range(4):
        h.run()                  # graph replay
        other.run()              # eager work between replays
        other.sync()
        _same(h.analysis(), a0)
    # a new cycle: swap + new obs invalidate the graph; it is captured again
    ref.cycle_swap()
    ref.set_obs(**job.obs)
    ref.run()
    a1 = ref.analysis()
    h.cycle_swap()
    h.set_obs(**job.obs)
    for _ in range(3):
        h.run()
        _same(h.analysis(), a1)
    for x in (ref, other, h):
        x.close()
