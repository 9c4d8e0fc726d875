"""Parity of the HIP path with the CPU oracle: indices and d2 bits
bit-identical, -1 / NaN fill, tie order, per-source final reach.
Runs through the C-ABI (lc_gpu_knearest_ex and friends)."""
import json
from pathlib import Path

import numpy as np
import pytest

from gpu_util import run_device, same_bits
from oracle import lcoracle as orc
from linkcell_amd import workloads as wl

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"
CASES = json.loads((GOLD / "known_answers.json").read_text())["cases"]


@pytest.fixture(scope="module")
def ws(gpu):
    import linkcell_amd as lc
    w = lc.Workspace()
    yield w
    w.close()


def oracle(xyz, cell, k, mask=None, hint=0.0, frames=1, stats=False):
    x = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(x) // frames
    return orc.knearest_many(x.reshape(frames, n, 3) if frames > 1 else x, cell, k,
                             mask=mask, cell_hint=hint, nthreads=None, stats=stats)


@pytest.mark.parametrize("cs", CASES, ids=[c["name"] for c in CASES])
def test_known_answer_cases_bitwise(ws, cs):
    hint = cs["hint"] if cs["hint"] is not None else 0.0
    gn, gd = run_device(ws, cs["xyz"], cs["cell"], cs["k"], cs["mask"], hint, cs["frames"])
    on, od = oracle(cs["xyz"], cs["cell"], cs["k"], cs["mask"], hint, cs["frames"])
    assert np.array_equal(gn, on), cs["name"]
    assert same_bits(gd, od), cs["name"]
    if cs["nn"] is not None:
        assert gn.ravel().tolist() == cs["nn"]


def test_c1_uniform_matches_golden_fixture(ws):
    c1 = wl.uniform_c1()
    g = np.load(GOLD / "c1_uniform_oracle.npz")
    nn, d2, reach = run_device(ws, c1["xyz"], c1["cell"], 12, None, 3.0, reach=True)
    assert np.array_equal(nn, g["nn"])
    assert np.array_equal(d2.view(np.uint64), g["d2_bits"])
    assert np.array_equal(reach, g["reach"])


@pytest.mark.parametrize("mode", [0, 1])
def test_c1_modes_agree(ws, mode):
    c1 = wl.uniform_c1()
    g = np.load(GOLD / "c1_uniform_oracle.npz")
    ws.set_mode(mode)
    try:
        nn, d2 = run_device(ws, c1["xyz"], c1["cell"], 12, None, 3.0)
    finally:
        ws.set_mode(0)
    assert np.array_equal(nn, g["nn"])
    assert np.array_equal(d2.view(np.uint64), g["d2_bits"])


def test_c2_triclinic_full_size_sampled(ws):
    c = wl.triclinic_c2()
    nn, d2, reach = run_device(ws, c["xyz"], c["cell"], 16, c["mask"], 3.0, reach=True)
    # masked rows: -1 / NaN, reach 0
    off = c["mask"] == 0
    assert (nn[off] == -1).all()
    assert (d2[off].view(np.uint64) == 0x7FF8000000000000).all()
    assert (reach[off] == 0).all()
    rng = np.random.default_rng(0)
    src = np.sort(rng.choice(np.flatnonzero(c["mask"]), 3000, replace=False))
    on, od, orr, _, _, _ = orc.knearest_sample(c["xyz"], c["cell"], 16, src,
                                               mask=c["mask"], cell_hint=3.0)
    assert np.array_equal(nn[src], on[src])
    assert same_bits(d2[src], od[src])
    assert np.array_equal(reach[src], orr[src])


@pytest.mark.parametrize("split", [-1, -2])
def test_c1_c2_row_walk_launches_agree(ws, split):
    """The row walk launched one wave per tile group and as one persistent
    launch (waves taking groups from a device counter) give the oracle's
    rows, reaches included."""
    c1 = wl.uniform_c1()
    g = np.load(GOLD / "c1_uniform_oracle.npz")
    c = wl.triclinic_c2(n=65536)
    ws.set_dense_min(split)
    try:
        nn, d2, reach = run_device(ws, c1["xyz"], c1["cell"], 12, None, 3.0, reach=True)
        a = run_device(ws, c["xyz"], c["cell"], 16, c["mask"], 3.0)
    finally:
        ws.set_dense_min(0)
    assert np.array_equal(nn, g["nn"])
    assert np.array_equal(d2.view(np.uint64), g["d2_bits"])
    assert np.array_equal(reach, g["reach"])
    on, od = oracle(c["xyz"], c["cell"], 16, c["mask"], 3.0)
    assert np.array_equal(a[0], on) and same_bits(a[1], od)


def test_c2_deterministic_across_calls(ws):
    c = wl.triclinic_c2(n=65536)
    a = run_device(ws, c["xyz"], c["cell"], 16, c["mask"], 3.0)
    b = run_device(ws, c["xyz"], c["cell"], 16, c["mask"], 3.0)
    assert np.array_equal(a[0], b[0]) and same_bits(a[1], b[1])
    on, od = oracle(c["xyz"], c["cell"], 16, c["mask"], 3.0)
    assert np.array_equal(a[0], on) and same_bits(a[1], od)


def test_c5i_small_boxes_k100(ws):
    for fr in wl.small_c5i(n_frames=6):
        nn, d2 = run_device(ws, fr["xyz"], fr["cell"], 100, None, 3.0)
        on, od = oracle(fr["xyz"], fr["cell"], 100, None, 3.0)
        assert np.array_equal(nn, on)
        assert same_bits(d2, od)


def test_shared_cell_frames_with_mask(ws):
    rng = np.random.default_rng(4)
    H = np.array([[14.0, 0, 0], [3.0, 13.0, 0], [-2.0, 1.5, 12.0]])
    F, n = 3, 700
    xyz = (rng.random((F, n, 3)) * 1.2 - 0.1) @ H
    mask = (rng.random(n) > 0.2).astype(np.int32)
    for k in (1, 5, 16, 24, 40):
        nn, d2 = run_device(ws, xyz, H, k, mask, 2.2, frames=F)
        on, od = oracle(xyz, H, k, mask, 2.2, frames=F)
        assert np.array_equal(nn, on), k
        assert same_bits(d2, od), k


@pytest.mark.parametrize("k", [4, 6, 7, 18, 26])
def test_perfect_lattice_ties_follow_reference_order(ws, k):
    # knn_hot.rs-style lattice: every shell is a boundary tie, so the
    # reference's visit order decides which equidistant index it keeps
    g = np.arange(10, dtype=np.float64)
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    xyz = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1) * 1.1
    nn, d2 = run_device(ws, xyz, [11.0, 11.0, 11.0], k, None, 3.0)
    on, od = oracle(xyz, [11.0, 11.0, 11.0], k, None, 3.0)
    assert np.array_equal(nn, on)
    assert same_bits(d2, od)


@pytest.mark.parametrize("masked", [False, True])
def test_points_given_in_cell_order(ws, masked):
    # whole waves of consecutive points in one cell take their count slots
    # one run at a time; a mask leaves lanes idle and takes the per-lane path
    rng = np.random.default_rng(21)
    xyz = rng.random((3000, 3)) * 12.0
    c = np.floor(xyz / 3.0).astype(np.int64)
    xyz = np.ascontiguousarray(xyz[np.lexsort((c[:, 0], c[:, 1], c[:, 2]))])
    mask = (rng.random(3000) > 0.1).astype(np.int32) if masked else None
    nn, d2 = run_device(ws, xyz, [12.0, 12.0, 12.0], 10, mask, 3.0)
    on, od = oracle(xyz, [12.0, 12.0, 12.0], 10, mask, 3.0)
    assert np.array_equal(nn, on)
    assert same_bits(d2, od)


def test_coincident_points_keep_zero_distance(ws):
    rng = np.random.default_rng(9)
    base = rng.random((200, 3)) * 9.0
    xyz = np.concatenate([base, base[:50]])
    nn, d2 = run_device(ws, xyz, [9.0, 9.0, 9.0], 6, None, 2.0)
    on, od = oracle(xyz, [9.0, 9.0, 9.0], 6, None, 2.0)
    assert np.array_equal(nn, on) and same_bits(d2, od)
    assert (d2[:50, 0] == 0.0).all()


def test_points_far_outside_primary_cell(ws):
    rng = np.random.default_rng(10)
    H = np.array([[20.0, 0, 0], [7.0, 18.0, 0], [3.0, -4.0, 16.0]])
    s = rng.random((3000, 3)) + rng.integers(-50, 50, size=(3000, 3))
    xyz = s @ H + np.array([1.0, -2.0, 0.5])
    cell = np.vstack([H, [1.0, -2.0, 0.5]])
    nn, d2 = run_device(ws, xyz, cell, 12, None, 2.5)
    on, od = oracle(xyz, cell, 12, None, 2.5)
    assert np.array_equal(nn, on) and same_bits(d2, od)


@pytest.mark.parametrize("k", [4, 8, 16, 32])
def test_unaligned_output_rows_match_aligned(ws, k):
    """Output rows that are not 16-byte aligned (buffers offset by one
    element) take the element-wise emit; the results must equal the aligned
    run bit for bit, and nothing outside the rows may be written."""
    import torch
    from gpu_util import cell12
    rng = np.random.default_rng(40 + k)
    xyz = rng.random((6000, 3)) * 24.0
    cell = [24.0, 24.0, 24.0]
    n = len(xyz)
    nn_a, d2_a = run_device(ws, xyz, cell, k, None, 2.0)
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(np.ascontiguousarray(xyz)).to(dev)
    nn_buf = torch.full((n * k + 2,), -7, dtype=torch.int32, device=dev)
    d2_buf = torch.full((n * k + 2,), -7.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ws.knearest_ptr(dx.data_ptr(), n, 1, cell12(cell), k, nn_buf.data_ptr() + 4,
                    d2_buf.data_ptr() + 8, 0, 2.0, 0, True)
    nb, db = nn_buf.cpu().numpy(), d2_buf.cpu().numpy()
    assert nb[0] == -7 and nb[-1] == -7 and db[0] == -7.0 and db[-1] == -7.0
    assert np.array_equal(nb[1:-1].reshape(n, k), nn_a)
    assert same_bits(db[1:-1].reshape(n, k), d2_a)
    on, od = oracle(xyz, cell, k, None, 2.0)
    assert np.array_equal(nn_a, on) and same_bits(d2_a, od)


def test_host_pointer_c_api_matches_oracle(gpu):
    import ctypes as C
    from linkcell_amd import _lib
    lib = _lib.load()
    c = wl.triclinic_c2(n=20000)
    n, k = 20000, 16
    x = np.ascontiguousarray(c["xyz"])
    nn = np.full(n * k, -7, dtype=np.int32)
    d2 = np.full(n * k, -7.0)
    cell = _lib.lc_cell.from12(list(c["cell"].reshape(9)) + [0, 0, 0])
    m = np.ascontiguousarray(c["mask"])
    rc = lib.lc_knearest_d2(x.ctypes.data, n, C.byref(cell), k, m.ctypes.data, 3.0,
                            nn.ctypes.data, d2.ctypes.data)
    assert rc == 0 and lib.lc_last_error() is None
    on, od = oracle(x, c["cell"], k, m, 3.0)
    assert np.array_equal(nn.reshape(n, k), on) and same_bits(d2.reshape(n, k), od)


def test_python_knearest_host_and_device(gpu):
    import torch
    import linkcell_amd as lc
    # tests/python/test_knearest.py and test_knearest_cuda.py, on ROCm
    xyz = np.ascontiguousarray([[0.2, 0.0, 0.0], [9.4, 0.0, 0.0]], dtype=np.float64)
    cell = np.ascontiguousarray([10.0, 10.0, 10.0])
    nn, d2 = lc.knearest(xyz, cell, 1)
    nn, d2 = np.from_dlpack(nn), np.from_dlpack(d2)
    assert nn.shape == (2, 1) and nn[0, 0] == 1 and nn[1, 0] == 0
    assert abs(d2[0, 0] - 0.64) < 1e-12
    dn, dd = lc.knearest(torch.from_numpy(xyz).cuda(), cell, 1)
    assert dn.device.type == "cuda" and int(dn[0, 0].item()) == 1
    assert abs(float(dd[0, 0].item()) - 0.64) < 1e-12
    # device cell stays on device
    dn, dd = lc.knearest(torch.from_numpy(xyz).cuda(), torch.from_numpy(cell).cuda(), 1)
    assert int(dn[0, 0].item()) == 1 and abs(float(dd[0, 0].item()) - 0.64) < 1e-12
    # batched frames (F, n, 3) -> (F, n, k)
    fr = np.stack([xyz, xyz])
    bn, bd = lc.knearest(torch.from_numpy(fr).cuda(), cell, 1)
    assert tuple(bn.shape) == (2, 2, 1) and int(bn[0, 0, 0].item()) == 1
    hn, hd = lc.knearest(fr, cell, 1)
    assert hn.shape == (2, 2, 1)
    # mask drops a source
    m = np.array([1, 0, 1], dtype=np.int32)
    x3 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], dtype=np.float64)
    mn, _ = lc.knearest(x3, cell, 1, mask=m)
    assert mn[:, 0].tolist() == [2, -1, 0]
    # device masks follow the host rules (python/src/lib.rs:188-211): int32
    # as given, float64 nonzero keeps, anything else raises
    dx3 = torch.from_numpy(x3).cuda()
    dn, _ = lc.knearest(dx3, cell, 1, mask=torch.from_numpy(m).cuda())
    assert dn[:, 0].tolist() == [2, -1, 0]
    dn, _ = lc.knearest(dx3, cell, 1, mask=torch.tensor([1.0, 0.0, 2.5]).double().cuda())
    assert dn[:, 0].tolist() == [2, -1, 0]
    with pytest.raises(ValueError, match="int32 or float64"):
        lc.knearest(dx3, cell, 1, mask=torch.tensor([1, 0, 1], dtype=torch.int64).cuda())
    with pytest.raises(ValueError, match="int32 or float64"):
        lc.knearest(x3, cell, 1, mask=np.array([1, 0, 1], dtype=np.int64))
    # a device cell must be float64, as the reference binding requires
    with pytest.raises(ValueError, match="float64"):
        lc.knearest(dx3, torch.tensor([10.0, 10.0, 10.0], dtype=torch.float32).cuda(), 1)
    # the search is ordered on torch's current stream without a host wait:
    # a producer kernel right before it and a consumer right after it see
    # the right data
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        big = torch.rand((20000, 3), dtype=torch.float64, device="cuda") * 20.0
        big.mul_(1.0)
        bn1, bd1 = lc.knearest(big, [20.0, 20.0, 20.0], 6)
        cn, cd = bn1.clone(), bd1.clone()
    s.synchronize()
    hn, hd = lc.knearest(big.cpu().numpy(), [20.0, 20.0, 20.0], 6)
    assert np.array_equal(cn.cpu().numpy(), hn)
    assert same_bits(cd.cpu().numpy(), hd)
    # one workspace per thread
    import threading
    seen = []

    def worker():
        seen.append(lc._workspace())  # kept alive past the thread's end

    th = [threading.Thread(target=worker) for _ in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mine = lc._workspace()
    assert seen[0] is not seen[1] and all(w is not mine for w in seen)


@pytest.mark.parametrize("shape", ["3", "9", "12"])
def test_device_cell_entry_matches_host_cell(ws, shape):
    import torch
    rng = np.random.default_rng(12)
    if shape == "3":
        cd = np.array([13.0, 11.0, 12.5]); cell = cd
    elif shape == "9":
        cd = np.array([[13.0, 0, 0], [2.0, 11.0, 0], [1.0, -1.0, 12.5]]); cell = cd
    else:
        cd = np.array([[13.0, 0, 0], [2.0, 11.0, 0], [1.0, -1.0, 12.5], [0.5, 0.25, -1.0]])
        cell = cd
    xyz = rng.random((1500, 3)) * 14.0
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(xyz).to(dev)
    dc = torch.from_numpy(np.ascontiguousarray(cd, dtype=np.float64).ravel()).to(dev)
    nn = torch.empty((1500, 8), dtype=torch.int32, device=dev)
    d2 = torch.empty((1500, 8), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    ws.knearest_dcell_ptr(dx.data_ptr(), 1500, 1, dc.data_ptr(), dc.numel(), 8,
                          nn.data_ptr(), d2.data_ptr(), 0, 2.0)
    on, od = oracle(xyz, cell, 8, None, 2.0)
    assert np.array_equal(nn.cpu().numpy(), on)
    assert same_bits(d2.cpu().numpy(), od)


def test_device_errors(ws):
    import linkcell_amd as lc
    with pytest.raises(lc.Error, match="bad cell"):
        run_device(ws, np.zeros((4, 3)), [0.0, 1.0, 1.0], 2)
    with pytest.raises(lc.Error, match="too many cells"):
        run_device(ws, np.zeros((4, 3)), [1e4, 1e4, 1e4], 2, None, 1e-3)


@pytest.mark.parametrize("k", [3, 16, 40])
def test_non_finite_coordinates_follow_the_oracle(ws, k):
    # A frame with NaN / +-inf coordinates: every source takes the exact-order
    # walk (src/cell.rs:176-207 folds NaN into cell 0, src/knearest.rs:106-158
    # keeps NaN distances by the same rules). The kept set is the oracle's;
    # finite entries match in order and bits. A NaN distance's sign is not
    # the reference's to fix (IEEE leaves NaN propagation to the platform:
    # x86's invalid-operation NaN is negative, a - b on the GPU is a + (-b),
    # and with two NaN operands x86 keeps the one the compiler put in the
    # destination register), and f64::total_cmp sorts negative NaN first and
    # positive NaN last: NaN entries match as a set of indices.
    rng = np.random.default_rng(12)
    xyz = rng.random((400, 3)) * 9.0
    xyz[5, 0] = np.nan
    xyz[17, 1] = np.inf
    xyz[33, 2] = -np.inf
    xyz[90] = np.nan
    nn, d2 = run_device(ws, xyz, [9.0, 9.5, 10.0], k, None, 3.0)
    on, od = oracle(xyz, [9.0, 9.5, 10.0], k, None, 3.0)
    for i in range(len(nn)):
        gf, of = ~np.isnan(d2[i]), ~np.isnan(od[i])
        assert np.array_equal(nn[i][gf], on[i][of]), i
        assert same_bits(d2[i][gf], od[i][of]), i
        assert sorted(nn[i][~gf].tolist()) == sorted(on[i][~of].tolist()), i


TIES = json.loads((GOLD / "tie_answers.json").read_text())["cases"]


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("cs", TIES, ids=[c["name"] for c in TIES])
def test_tie_answers_on_device(ws, cs, mode):
    """The hand-derived tie rows (tests/golden/make_golden.py) on the GPU:
    the fast walk must see each boundary tie and hand the source to the
    exact-order walk (mode 0), or run it there directly (mode 1)."""
    ws.set_mode(mode)
    ws.set_timing(True)
    try:
        nn, d2, reach = run_device(ws, cs["xyz"], cs["cell"], cs["k"], None, cs["hint"],
                                   reach=True)
        deferred = ws.last_timing()["deferred"]
    finally:
        ws.set_mode(0)
        ws.set_timing(False)
    for r, want in cs["rows"].items():
        assert nn[int(r)].tolist() == want, cs["derivation"]
        assert d2[int(r)].tolist() == cs["d2"][r], cs["derivation"]
    on, od, orr, _ = oracle(cs["xyz"], cs["cell"], cs["k"], None, cs["hint"], stats=True)
    assert np.array_equal(nn, on) and same_bits(d2, od)
    assert np.array_equal(reach, orr)
    if mode == 0 and cs["name"].startswith("tie_"):
        assert deferred >= 1


@pytest.mark.parametrize("k", [129, 200])
def test_exact_walk_beyond_wave_heaps(ws, k):
    """k above 128 runs the lane-per-source exact walk with its heap in HBM
    scratch: a small box whose stencil wraps (several images of one point),
    a masked frame, coincident points; indices, d2 bits and reach."""
    rng = np.random.default_rng(k)
    cases = []
    # small box, 300 points: k exceeds the other points for no source, the
    # stencil wraps (mesh 2 per axis at hint 3)
    cases.append((rng.random((300, 3)) * 7.0, [7.0, 7.0, 7.0], None, 3.0))
    # masked triclinic frame with coincident points
    H = np.array([[11.0, 0, 0], [2.0, 10.0, 0], [-1.0, 1.5, 9.0]])
    base = (rng.random((420, 3)) * 1.1 - 0.05) @ H
    xyz = np.concatenate([base, base[:60]])
    mask = (rng.random(len(xyz)) > 0.15).astype(np.int32)
    cases.append((xyz, H, mask, 2.5))
    # fewer active points than k: rows end in -1 / NaN
    cases.append((rng.random((150, 3)) * 9.0, [9.0, 9.0, 9.0], None, 2.0))
    for xyz, cell, m, hint in cases:
        nn, d2, reach = run_device(ws, xyz, cell, k, m, hint, reach=True)
        on, od, orr, _ = oracle(xyz, cell, k, m, hint, stats=True)
        assert np.array_equal(nn, on)
        assert same_bits(d2, od)
        assert np.array_equal(reach, orr)
    assert (nn[:, 149:] == -1).all()
