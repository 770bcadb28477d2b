"""Host-side pieces of the serving loop (no GPU)."""
import numpy as np

from letkf_amd import serve, synth


def test_poisson_arrivals_rate():
    t = serve.poisson_arrivals(2.0, 500.0, seed=1)
    assert np.all(np.diff(t) > 0) and t[-1] < 500.0
    assert abs(t.size / 500.0 - 2.0) < 0.25


def test_estimate_bytes_covers_ensemble():
    p = synth.l96(nx=40, ne=20)
    j = serve.l96_job("a", p)
    assert serve.estimate_bytes(j) >= 2 * 40 * 20 * 8
