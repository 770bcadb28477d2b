"""Pins of the oracle's regional (SCALE) index and capped selection (CPU only).

* binning (scale/letkf/letkf_obs.f90:740-830) equals a stable numpy lexsort by
  the clamped cell (ceiling(ri-0.5), ceiling(rj-0.5));
* uncapped selection equals a brute-force scan with the same rectangle and
  normalised-distance cutoff, in scan order;
* capped selection keeps, per bucket, exactly the nobs_max smallest
  (rdiag, scan position) pairs of the uncapped list (a full sort, not the
  reference's insertion sort), so the insertion-sort restatement is pinned
  by an independent algorithm; exact ties go to the earlier obs;
* the order of a capped bucket is the reference's storage-slot order
  (scale/letkf/letkf_tools.f90:1063-1150): checked against hand-worked
  cases and against an independent max-heap simulation of the slot process
  (a newcomer below the set's largest entry evicts it and takes its slot).
"""
import heapq

import numpy as np
import pytest

import oracle

NX, NY = 30, 26
DX = DY = 500.0


def _obs(n, seed, dup=False):
    rng = np.random.default_rng(seed)
    ri = rng.uniform(0.5, NX + 0.5, n)
    rj = rng.uniform(0.5, NY + 0.5, n)
    z = rng.uniform(0.0, 8000.0, n)
    if dup:  # gridded obs with exact duplicates in position
        ri = np.round(ri)
        rj = np.round(rj)
        z = np.round(z / 500.0) * 500.0
    err = np.where(rng.uniform(size=n) < 0.5, 3.0, 5.0)
    bkt = (err == 5.0).astype(np.int32)
    return dict(ri=ri, rj=rj, z=z, err=err, bucket=bkt)


def test_bins_equal_stable_lexsort():
    o = _obs(3000, 1)
    o["ri"][:3] = [0.5, NX + 0.5, 0.6]  # edges clamp into the grid
    qc = np.ones(3000, dtype=np.int32)
    qc[10:20] = 0
    perm, grd = oracle.obs_bins_regional(o["ri"], o["rj"], NX, NY, qc=qc)
    ci = np.clip(np.ceil(o["ri"] - 0.5).astype(int), 1, NX)
    cj = np.clip(np.ceil(o["rj"] - 0.5).astype(int), 1, NY)
    cell = (ci - 1) + NX * (cj - 1)
    idx = np.flatnonzero(qc == 1)
    order = idx[np.lexsort((idx, cell[idx]))]
    np.testing.assert_array_equal(perm, order)
    counts = np.bincount(cell[idx], minlength=NX * NY).reshape(NY, NX)
    cum = np.cumsum(counts.ravel()).reshape(NY, NX)
    np.testing.assert_array_equal(grd[:, 1:], cum)


def _sorted(o, perm):
    return {k: np.asarray(v)[perm] for k, v in o.items()}


def _brute(so, ri, rj, rz, sh, sv):
    dlz = sh * oracle.DIST_ZERO_FAC_SCALE / DX
    imin, imax = max(1, int(np.floor(ri - dlz))), min(NX, int(np.ceil(ri + dlz)))
    jmin, jmax = max(1, int(np.floor(rj - dlz))), min(NY, int(np.ceil(rj + dlz)))
    ci = np.clip(np.ceil(so["ri"] - 0.5).astype(int), 1, NX)
    cj = np.clip(np.ceil(so["rj"] - 0.5).astype(int), 1, NY)
    inrect = (ci >= imin) & (ci <= imax) & (cj >= jmin) & (cj <= jmax)
    rdx = (ri - so["ri"]) * DX
    rdy = (rj - so["rj"]) * DY
    ndh = np.sqrt(rdx * rdx + rdy * rdy) / sh
    ndv = np.abs(so["z"] - rz) / sv
    nd = ndh * ndh + ndv * ndv
    ok = inrect & ~(nd > 13.33333333)
    idx = np.flatnonzero(ok)
    rho = np.exp(-0.5 * nd[idx])
    return idx, rho, so["err"][idx] ** 2 / rho


@pytest.mark.parametrize("pt", [(5.0, 7.0, 1000.0), (1.0, 1.0, 0.0), (30.0, 26.0, 7000.0),
                                (15.0, 13.0, 4000.0)])
@pytest.mark.parametrize("dup", [False, True])
def test_selection_uncapped_and_capped(pt, dup):
    o = _obs(4000, 2 + dup, dup=dup)
    perm, grd = oracle.obs_bins_regional(o["ri"], o["rj"], NX, NY)
    so = _sorted(o, perm)
    ri, rj, rz = pt
    sh, sv = 2000.0, 2000.0
    idx, rl, rd = oracle.obs_local_regional(ri, rj, rz, NX, NY, grd, so, DX, DY, sh, sv, 0, 2)
    bidx, brho, brd = _brute(so, ri, rj, rz, sh, sv)
    np.testing.assert_array_equal(idx, bidx)  # scan order = sorted-obs order here
    np.testing.assert_allclose(rl, brho, rtol=1e-15)
    for cap in (1, 7, 40, 100000):
        cidx, crl, crd = oracle.obs_local_regional(ri, rj, rz, NX, NY, grd, so, DX, DY, sh, sv,
                                                   cap, 2)
        want = []
        for b in range(2):
            sel = np.flatnonzero(so["bucket"][idx] == b)
            order = sel[np.lexsort((sel, rd[sel]))][:cap]  # (rdiag, scan position)
            want.append(np.sort(idx[order]))
        got_b = so["bucket"][cidx]
        assert np.all(np.diff(got_b) >= 0)  # bucket by bucket
        for b in range(2):
            np.testing.assert_array_equal(np.sort(cidx[got_b == b]), want[b])
        np.testing.assert_allclose(crd, so["err"][cidx] ** 2 / crl, rtol=1e-15)


def _heap_slots(key, cap):
    """Storage slots of the reference cap by a max-heap over (key, position):
    an independent algorithm (no sorted index, no shifting)."""
    heap = []  # (-key, -pos, slot): the top is the largest (key, pos)
    slot_src = []
    for i, k in enumerate(key):
        if len(slot_src) < cap:
            heapq.heappush(heap, (-k, -i, len(slot_src)))
            slot_src.append(i)
        elif (k, i) < (-heap[0][0], -heap[0][1]):
            _, _, sl = heapq.heapreplace(heap, (-k, -i, heap[0][2]))
            slot_src[sl] = i
    return np.array(slot_src, dtype=np.int64)


def test_cap_stream_known_answers():
    # a falling key sequence: 3 evicts 5 (slot 0), 2 evicts 4 (slot 1), 1 evicts 3 (slot 0)
    np.testing.assert_array_equal(oracle.cap_stream([5.0, 4.0, 3.0, 2.0, 1.0], 2), [4, 3])
    # ties: an equal key never evicts, so the earlier obs stay
    np.testing.assert_array_equal(oracle.cap_stream([1.0, 1.0, 1.0], 2), [0, 1])
    np.testing.assert_array_equal(oracle.cap_stream([2.0, 1.0, 1.0], 1), [1])
    # within the cap nothing moves: scan order
    np.testing.assert_array_equal(oracle.cap_stream([3.0, 1.0, 2.0], 5), [0, 1, 2])
    # a rising sequence keeps the first cap in scan order
    np.testing.assert_array_equal(oracle.cap_stream(np.arange(10.0), 3), [0, 1, 2])
    assert oracle.cap_stream([], 3).size == 0


@pytest.mark.parametrize("seed", range(12))
def test_cap_stream_equals_heap_simulation_and_sort(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 3000))
    cap = int(rng.integers(1, 400))
    kind = seed % 3
    if kind == 0:
        key = rng.integers(0, 6, n).astype(float)  # many exact ties
    elif kind == 1:
        key = rng.standard_normal(n)
    else:  # a radar-like sweep: far, near, far again
        key = np.abs(np.arange(n) - n / 2.0) + rng.uniform(0.0, 40.0, n)
    got = oracle.cap_stream(key, cap)
    np.testing.assert_array_equal(got, _heap_slots(key, cap))
    keep = np.sort(np.lexsort((np.arange(n), key))[:cap])  # the set: a full sort
    np.testing.assert_array_equal(np.sort(got), keep)


def test_capped_buckets_in_slot_order():
    o = _obs(6000, 7, dup=True)
    perm, grd = oracle.obs_bins_regional(o["ri"], o["rj"], NX, NY)
    so = _sorted(o, perm)
    ri, rj, rz = 14.0, 12.0, 3000.0
    sh = sv = 2000.0
    idx, rl, rd = oracle.obs_local_regional(ri, rj, rz, NX, NY, grd, so, DX, DY, sh, sv, 0, 2)
    for cap in (1, 13, 150):
        cidx, _, _ = oracle.obs_local_regional(ri, rj, rz, NX, NY, grd, so, DX, DY, sh, sv, cap,
                                               2)
        want = []
        for b in range(2):
            sel = np.flatnonzero(so["bucket"][idx] == b)
            assert sel.size > cap  # the cap binds
            want.append(idx[sel][_heap_slots(rd[sel], cap)])
        np.testing.assert_array_equal(cidx, np.concatenate(want))
