"""The device-cell path (lc_gpu_knearest_many_dcell): the cell is inverted
and the mesh planned on the device (no read-back), bit for bit the host
plan; results equal the host-cell search; failures are reported after the
wait; the whole search captures into a graph and replays to the same bits."""
import numpy as np
import pytest

from gpu_util import run_device, same_bits
from oracle import lcoracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


def _random_cells(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 3
        if kind == 0:
            out.append(rng.uniform(3.0, 90.0, size=3))
        else:
            L = rng.uniform(5.0, 60.0, size=3)
            t = rng.uniform(-0.45, 0.45, size=3)
            H = np.array([[L[0], 0.0, 0.0], [t[0] * L[0], L[1], 0.0],
                          [t[1] * L[0], t[2] * L[1], L[2]]])
            H = H + rng.normal(0.0, 0.05, size=(3, 3)) * (rng.random() < 0.5)
            if kind == 2:
                H = np.vstack([H, rng.uniform(-20.0, 20.0, size=3)])
            out.append(H)
    return out


def test_device_plan_matches_host_plan_bitwise(ws):
    import torch
    dev = torch.device("cuda:0")
    for i, cell in enumerate(_random_cells(7, 600)):
        c = np.ascontiguousarray(cell, dtype=np.float64)
        hint = (0.0, 2.0, 3.0, 4.7, -1.0)[i % 5]
        dc = torch.from_numpy(c.ravel()).to(dev)
        got = ws.dcell_plan(dc.data_ptr(), c.size, hint)
        c12 = orc.cell12_of(c)
        oc = orc.Cell(c12[0:3], c12[3:6], c12[6:9], c12[9:12])
        want = oc.plan(hint if hint > 0 else float("nan"))
        assert got["status"] == 0, i
        assert np.array_equal(np.array(got["widths"]), np.array(oc.widths)), (i, cell)
        assert np.array_equal(np.array(got["hinv"]), np.array(oc.hinv)), (i, cell)
        assert got["ortho"] == int(oc.is_ortho), i
        assert got["cell_min"] == want["cell_min"], i
        assert (got["nx"], got["ny"], got["nz"], got["ncell"], got["max_reach"]) == (
            want["nx"], want["ny"], want["nz"], want["ncell"], want["max_reach"]), i
        assert np.array_equal(np.array(got["origin"]), c12[9:12]), i


def _dcell_search(ws, xyz, cell, k, mask=None, hint=0.0, frames=1, wait=True):
    import torch
    dev = torch.device("cuda:0")
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(x) // frames
    dx = torch.from_numpy(x).to(dev)
    dc = torch.from_numpy(np.ascontiguousarray(cell, dtype=np.float64).ravel()).to(dev)
    dm = None if mask is None else torch.from_numpy(np.ascontiguousarray(mask, np.int32)).to(dev)
    nn = torch.full((frames * n, k), -7, dtype=torch.int32, device=dev)
    d2 = torch.full((frames * n, k), -7.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ws.knearest_dcell_ptr(dx.data_ptr(), n, frames, dc.data_ptr(), dc.numel(), k,
                          nn.data_ptr(), d2.data_ptr(), dm.data_ptr() if dm is not None else 0,
                          hint, wait=wait)
    return nn, d2, (dx, dc, dm)


def test_dcell_search_matches_host_cell_on_random_cells(ws):
    rng = np.random.default_rng(8)
    for i, cell in enumerate(_random_cells(9, 9)):
        c12 = orc.cell12_of(cell)
        H = c12[:9].reshape(3, 3)
        frames = 1 + i % 3
        n = 900
        xyz = (rng.random((frames, n, 3)) * 1.3 - 0.15) @ H + c12[9:12]
        mask = (rng.random(n) > 0.1).astype(np.int32) if i % 2 else None
        k = (4, 12, 20)[i % 3]
        nn, d2, _ = _dcell_search(ws, xyz, cell, k, mask, 2.5, frames)
        hn, hd = run_device(ws, xyz, c12, k, mask, 2.5, frames)
        assert np.array_equal(nn.cpu().numpy(), hn), i
        assert same_bits(d2.cpu().numpy(), hd), i
        on, od = orc.knearest_many(xyz, cell, k, mask=mask, cell_hint=2.5, nthreads=None)
        assert np.array_equal(hn, on), i


def test_dcell_failures_are_reported_after_the_wait(ws):
    import linkcell_amd as lc
    xyz = np.random.default_rng(1).random((50, 3)) * 5.0
    with pytest.raises(lc.Error, match="bad cell"):
        _dcell_search(ws, xyz, [0.0, 5.0, 5.0], 3)
    # the rows of a failed plan are -1 / NaN
    try:
        ws.set_dcell_capacity(1000)
        with pytest.raises(lc.Error, match="too many cells"):
            _dcell_search(ws, xyz, [60.0, 60.0, 60.0], 3, hint=3.0)  # 8000 cells
        nn, d2, keep = _dcell_search(ws, xyz, [60.0, 60.0, 60.0], 3, hint=3.0, wait=False)
        with pytest.raises(lc.Error, match="too many cells"):
            ws.wait()
        assert (nn.cpu().numpy() == -1).all() and np.isnan(d2.cpu().numpy()).all()
        # a good search afterwards succeeds and the error slot clears
        nn, d2, keep = _dcell_search(ws, xyz, [5.0, 5.0, 5.0], 3, hint=1.0)
        on, od = orc.knearest_many(xyz, [5.0, 5.0, 5.0], 3, cell_hint=1.0)
        assert np.array_equal(nn.cpu().numpy(), on)
        ws.wait()
    finally:
        ws.set_dcell_capacity(262144)


def test_dcell_search_replays_from_a_captured_graph(ws):
    import torch
    rng = np.random.default_rng(12)
    H = np.array([[16.0, 0, 0], [3.0, 15.0, 0], [-2.0, 1.0, 14.0], [0.5, -1.0, 2.0]])
    xyz = (rng.random((20000, 3)) * 1.2 - 0.1) @ H[:3] + H[3]
    ref = orc.knearest_many(xyz, H, 12, cell_hint=2.0, nthreads=None)
    nn, d2, keep = _dcell_search(ws, xyz, H, 12, hint=2.0)  # sizes the buffers
    assert np.array_equal(nn.cpu().numpy(), ref[0])
    queue = torch.cuda.ExternalStream(ws.queue)
    g = torch.cuda.CUDAGraph()
    dx, dc, _ = keep
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=queue, capture_error_mode="global"):
        ws.knearest_dcell_ptr(dx.data_ptr(), 20000, 1, dc.data_ptr(), dc.numel(), 12,
                              nn.data_ptr(), d2.data_ptr(), 0, 2.0, wait=False)
    for rep in range(3):
        nn.fill_(-3)
        d2.fill_(-3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(nn.cpu().numpy(), ref[0]), rep
        assert same_bits(d2.cpu().numpy(), ref[1]), rep
        y = (dx * 2.0).sum()  # eager work between replays
        torch.cuda.synchronize()
        del y
    ws.wait()
