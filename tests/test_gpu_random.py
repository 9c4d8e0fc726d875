"""Randomized parity: small seeded inputs of every kind the walk has to
handle (orthorhombic and triclinic boxes down to the size of a cell, uniform
and clustered points, lattices with exact distance ties, coincident points,
points several images outside the box, masks, several frames, k from 1 to
above the point count, coordinate scales from 1e-3 to 1e45, beyond float
range) through every walk and entry point, each against the oracle (the
reference's KHeap walk, src/knearest.rs:351-486) bit for bit, per-source
reach included."""
import numpy as np
import pytest

from gpu_util import cell12, run_device, same_bits
from oracle import lcoracle as orc

pytestmark = pytest.mark.gpu

MODES = ("row", "persistent", "lane", "dense", "exact", "dcell")


def make_case(seed):
    rng = np.random.default_rng(seed)
    ortho = rng.random() < 0.5
    edges = rng.uniform(1.5, 30.0, size=3)
    if rng.random() < 0.25:
        edges = rng.uniform(1.0, 6.0, size=3)
    H = np.diag(edges)
    if not ortho:
        H[1, 0] = rng.uniform(-0.5, 0.5) * edges[0]
        H[2, 0] = rng.uniform(-0.5, 0.5) * edges[0]
        H[2, 1] = rng.uniform(-0.5, 0.5) * edges[1]
    cell = edges.copy() if ortho else H.copy()
    n = int(rng.choice([1, 2, 3, 17, 64, 65, 200, 700, 1500]))
    kind = str(rng.choice(["uniform", "cluster", "lattice", "dup", "outside"]))
    s = rng.random((n, 3))
    if kind == "cluster":
        nc = int(rng.integers(1, 6))
        cen = rng.random((nc, 3))
        s = cen[rng.integers(0, nc, n)] + rng.normal(0, rng.uniform(0.01, 0.1), (n, 3))
    elif kind == "lattice":
        g = int(rng.integers(2, 9))
        s = rng.integers(0, g, (n, 3)) / g
    elif kind == "dup":
        m = max(1, n // 3)
        s = rng.random((m, 3))[rng.integers(0, m, n)]
    elif kind == "outside":
        s = s + rng.integers(-3, 4, (n, 3))
    xyz = s @ H
    if kind == "lattice" and ortho:
        xyz = np.round(xyz * 2) / 2  # multiples of 0.5: exact ties in d2
    k = int(rng.choice([1, 2, 4, 5, 8, 12, 16, 17, 24, 32, 33, 48]))
    mask = (rng.random(n) < 0.8).astype(np.int32) if rng.random() < 0.3 else None
    hint = float(rng.choice([0.0, 0.0, 1.0, 2.0, 3.0, 5.0]))
    frames = int(rng.choice([1, 1, 2]))
    if frames > 1:
        xyz = np.concatenate([xyz, xyz + rng.normal(0, 0.05, xyz.shape)])
    scale = float(rng.choice([1.0, 1.0, 1.0, 1e-3, 1e30, 1e45]))
    if scale != 1.0:
        # the hint scales with the box, so the mesh stays within the
        # reference's cell limit
        xyz = xyz * scale
        cell = cell * scale
        hint = max(hint, 1.0) * scale
    return dict(xyz=xyz, cell=cell, k=k, mask=mask, hint=hint, frames=frames, n=n)


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


@pytest.mark.parametrize("case", range(192))
def test_random_inputs_match_oracle(ws, case):
    import torch
    seed = 7100 + case
    mode = MODES[case % len(MODES)]
    c = make_case(seed)
    on, od, orr, _ = orc.knearest_many(c["xyz"], c["cell"], c["k"], mask=c["mask"],
                                       cell_hint=c["hint"], n_frames=c["frames"], stats=True)
    where = f"seed {seed} mode {mode} n {c['n']} k {c['k']} frames {c['frames']}"
    if mode == "dcell":
        dev = torch.device("cuda:0")
        rows = c["frames"] * c["n"]
        x = torch.from_numpy(np.ascontiguousarray(c["xyz"])).to(dev)
        cd = torch.tensor(cell12(c["cell"]), dtype=torch.float64, device=dev)
        nn = torch.full((rows, c["k"]), -7, dtype=torch.int32, device=dev)
        d2 = torch.full((rows, c["k"]), -7.0, dtype=torch.float64, device=dev)
        dm = torch.from_numpy(c["mask"]).to(dev) if c["mask"] is not None else None
        ws.knearest_dcell_ptr(x.data_ptr(), c["n"], c["frames"], cd.data_ptr(), 12, c["k"],
                              nn.data_ptr(), d2.data_ptr(),
                              dm.data_ptr() if dm is not None else 0, c["hint"], wait=True)
        gn, gd = nn.cpu().numpy(), d2.cpu().numpy()
        assert np.array_equal(gn, on), where
        assert same_bits(gd, od), where
        return
    split = {"persistent": -2, "lane": 1 << 30, "dense": 1}.get(mode, 0)
    ws.set_dense_min(split)
    ws.set_mode(1 if mode == "exact" else 0)
    try:
        gn, gd, gr = run_device(ws, c["xyz"], c["cell"], c["k"], c["mask"], c["hint"],
                                frames=c["frames"], reach=True)
    finally:
        ws.set_dense_min(0)
        ws.set_mode(0)
    assert np.array_equal(gn, on), where
    assert same_bits(gd, od), where
    assert np.array_equal(gr, orr), where
