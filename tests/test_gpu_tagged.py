"""The row walk's tagged heap keys at their edges, bit for bit against the
oracle (the reference's KHeap walk, src/knearest.rs:351-486), per-source
reach included.

A tagged key is a d2 with its low 5 bits (6 for the 32-slot heap) replaced by
a slot tag, so two distances within 32 (64) ulps order by tag inside the
merge network. The probes below put such a pair exactly at the k-th slot
(the source must go to the exact-order walk) and just inside the kept k
(the emit must restore d2 order), for heaps of 16 and 32 slots. Boxes of
1e40 and 1e60 put every coordinate beyond float range, where the fill's
FP32 tile tests see infinite or NaN gaps and must keep every tile, while
the distances (up to ~1e121) still fit the tagged keys. (Cells with entries
of 1e150 and more, which the walk hands to the exact-order walk, are
refused by the reference's cell check before any walk: the norms of their
face vectors overflow, src/cell.rs:97-123.)"""
import numpy as np
import pytest

from gpu_util import run_device, same_bits
from oracle import lcoracle as orc

pytestmark = pytest.mark.gpu

BOX = 16.0


def near_tie_frame(n_close, seed):
    """Probe sources on a coarse grid, each with n_close neighbours strictly
    nearer than 0.75, then A at distance 0.75 along x (d2 = 0.5625) and B at
    0.75 + 2^-50 along y: d2 = 0.5625 + 12 ulps, the same 32-ulp tag bucket
    as A's. Coordinates are in [4, 8), where 2^-50 is representable, so both
    differences are exact; probes are 2 apart, so every other point is at
    least 0.9 from a probe."""
    rng = np.random.default_rng(seed)
    pts = []
    probes = []
    for gx in (4.25, 6.25):
        for gy in (4.25, 6.25):
            for gz in (4.25, 6.25):
                p = np.array([gx, gy, gz]) + rng.integers(0, 4, 3) * 0.125
                probes.append(len(pts))
                pts.append(p)
                # close neighbours: distinct radii in (0.2, 0.7), random
                # directions (their d2 are distinct from one another and
                # far below A's)
                r = np.sort(rng.uniform(0.2, 0.7, n_close))
                u = rng.normal(size=(n_close, 3))
                u /= np.linalg.norm(u, axis=1, keepdims=True)
                for q in p + r[:, None] * u:
                    pts.append(q)
                a = p + np.array([0.75, 0.0, 0.0])
                b = p + np.array([0.0, 0.75 + 2.0 ** -50, 0.0])
                # the order of A and B in the input varies between probes
                if rng.random() < 0.5:
                    pts += [a, b]
                else:
                    pts += [b, a]
    return np.array(pts), probes


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


@pytest.mark.parametrize("kc,straddle", [(16, True), (16, False), (32, True), (32, False)])
@pytest.mark.parametrize("split", [0, -2])
def test_near_tied_pair_at_and_inside_the_kth(ws, kc, straddle, split):
    # k = kc. kc - 1 close neighbours: A is the k-th, B the first beyond;
    # kc - 2: A and B are the last two kept
    n_close = kc - 1 if straddle else kc - 2
    k = kc
    xyz, probes = near_tie_frame(n_close, seed=kc * 10 + straddle)
    # the pair's exact d2, as the walk computes them
    p = xyz[probes[0]]
    dxa = (p[0] + 0.75) - p[0]
    dyb = (p[1] + (0.75 + 2.0 ** -50)) - p[1]
    assert dxa * dxa == 0.5625 and dyb * dyb != 0.5625
    assert (np.float64(dyb * dyb).view(np.uint64) >> np.uint64(5)) == \
        (np.float64(0.5625).view(np.uint64) >> np.uint64(5))
    on, od, orr, _ = orc.knearest_many(xyz, [BOX, BOX, BOX], k, cell_hint=2.0, stats=True)
    # the oracle keeps A at the k-th slot (straddle) or A then B (inside)
    for i in probes:
        if straddle:
            assert od[i, k - 1] == 0.5625
        else:
            assert od[i, k - 2] == 0.5625 and od[i, k - 1] == dyb * dyb
    ws.set_dense_min(split)
    try:
        gn, gd, gr = run_device(ws, xyz, [BOX, BOX, BOX], k, hint=2.0, reach=True)
    finally:
        ws.set_dense_min(0)
    assert np.array_equal(gn, on)
    assert same_bits(gd, od)
    assert np.array_equal(gr, orr)


@pytest.mark.parametrize("k", [8, 16, 32])
@pytest.mark.parametrize("L", [1e40, 1e60])
def test_coordinates_beyond_float_range(ws, k, L):
    rng = np.random.default_rng(k)
    xyz = rng.random((1500, 3)) * L
    on, od, orr, _ = orc.knearest_many(xyz, [L, L, L], k, cell_hint=L / 6, stats=True)
    for split in (0, -2):
        ws.set_dense_min(split)
        try:
            gn, gd, gr = run_device(ws, xyz, [L, L, L], k, hint=L / 6, reach=True)
        finally:
            ws.set_dense_min(0)
        assert np.array_equal(gn, on), split
        assert same_bits(gd, od), split
        assert np.array_equal(gr, orr), split
