"""Pins for the CPU oracle (oracle/lc_oracle.c), the parity target of the
HIP path. Each check compares against something the oracle does not
compute itself: the reference tests' asserted answers, an independent
numpy 27-image brute force, closed forms, or a numpy restatement of the
fold and distance arithmetic evaluated bit for bit."""
import json
import math
import threading
from pathlib import Path

import numpy as np
import pytest

from oracle import lcoracle as orc
import refimpl

GOLD = Path(__file__).resolve().parent / "golden"
CASES = json.loads((GOLD / "known_answers.json").read_text())["cases"]


def almost(a, b):
    # tests/agree.rs:6-8
    return abs(a - b) <= 1e-12 * (1.0 + max(abs(a), abs(b)))


# ------------------------------------------------------------ cell (a1)

def test_ortho_flag_and_widths():
    # src/cell.rs:20-34, :56-61, :395-421
    c = orc.Cell.ortho(10.0, 11.0, 12.0)
    assert c.is_ortho
    assert orc.Cell.ortho(10.0, 10.0, 10.0).widths == [10.0, 10.0, 10.0]
    sheared = orc.Cell([10, 0, 0], [5, 8.66, 0], [0, 0, 10])
    assert not sheared.is_ortho


@pytest.mark.parametrize("edge", [10.0, 16.0, 120.0, 360.0, 0.5, 1024.0])
def test_widths_closed_form_on_exact_boxes(edge):
    # |det| / |b x c| is exact when all products are representable
    c = orc.Cell.ortho(edge, 2 * edge, edge)
    assert c.widths == [edge, 2 * edge, edge]


def test_rejects_bad_box():
    # tests/agree.rs:18-27
    with pytest.raises(orc.OracleError, match="singular or non-positive cell"):
        orc.Cell.ortho(0.0, 1.0, 1.0)
    with pytest.raises(orc.OracleError):
        orc.Cell([1, 0, 0], [2, 0, 0], [0, 0, 1])
    with pytest.raises(orc.OracleError):
        orc.Cell([1e-7, 0, 0], [0, 1e-7, 0], [0, 0, 1e-7])  # |det| < 1e-18


def test_triclinic_widths_match_geometry():
    a, b, c = np.array([40.0, 0, 0]), np.array([20.0, 34.64, 0]), np.array([8.0, 6, 38])
    cell = orc.Cell(a, b, c)
    vol = abs(np.dot(a, np.cross(b, c)))
    want = [vol / np.linalg.norm(np.cross(b, c)), vol / np.linalg.norm(np.cross(c, a)),
            vol / np.linalg.norm(np.cross(a, b))]
    assert all(abs(x - y) <= 1e-12 * y for x, y in zip(cell.widths, want))
    # widths differ from the edge lengths for a sheared box
    assert abs(cell.widths[0] - 40.0) > 1.0


# --------------------------------------------------------- fold (a2, a3)

def test_fractional_known_answer():
    # src/cell.rs:161-174
    c = orc.Cell.ortho(10.0, 10.0, 10.0)
    s = c.fractional([10.5, -1.0, 3.0])
    assert abs(s[0] - 0.05) < 1e-12 and abs(s[1] - 0.90) < 1e-12
    assert abs(s[2] - 0.30) < 1e-12
    assert abs(c.cartesian(s)[0] - 0.5) < 1e-12


def test_wrap01_folds_rounded_one_to_zero():
    # src/cell.rs:291-298: -tiny/10 - floor(...) rounds to 1.0, which the
    # reference maps to 0.0; without that rule s would be 1.0
    c = orc.Cell.ortho(10.0, 10.0, 10.0)
    s = c.fractional([-1e-17, 5.0, 5.0])
    assert s[0] == 0.0


def test_fold_bits_match_numpy_restatement():
    rng = np.random.default_rng(11)
    for cell in ([16.0, 16.0, 16.0], [[40.0, 0, 0], [20.0, 34.64, 0], [8.0, 6, 38]],
                 [[10.0, 0, 0], [3.0, 9.0, 0], [1.0, 2.0, 8.0], [1.5, -2.0, 0.25]]):
        H, o = refimpl.rows_of(cell)
        pts = (rng.random((500, 3)) * 3.0 - 1.0) @ H + o
        s_np, f_np, _, _ = refimpl.fold(pts, cell)
        c12 = orc.cell12_of(cell)
        oc = orc.Cell(c12[0:3], c12[3:6], c12[6:9], c12[9:12])
        for i in range(len(pts)):
            s = oc.fractional(pts[i])
            f = oc.cartesian(s)
            assert np.array_equal(np.array(s), s_np[i])
            assert np.array_equal(np.array(f), f_np[i])
            assert all(0.0 <= v < 1.0 for v in s)


# ------------------------------------------------ shift and dist2 (a6)

def test_lattice_shift_and_mic():
    c = orc.Cell.ortho(10.0, 10.0, 10.0)
    # src/cell.rs:218-225
    assert c.lattice_shift(-1, 0, 0) == [-10.0, 0.0, 0.0]
    # src/cell.rs:255-267
    left, right = [0.2, 0, 0], [9.4, 0, 0]
    mic = c.dist2(left, right)
    via = orc.dist2_shifted(left, right, c.lattice_shift(-1, 0, 0))
    assert abs(via - mic) < 1e-12 and abs(mic - 0.64) < 1e-12
    # src/cell.rs:395-412
    b = orc.Cell.ortho(10.0, 11.0, 12.0)
    assert abs(orc.dist2_shifted([1, 2, 3], [2, 3, 4], [0, 0, 0]) - b.dist2([1, 2, 3], [2, 3, 4])) <= 1e-12


def test_general_shift_is_integer_combination():
    H = np.array([[10.0, 0, 0], [3.0, 9.0, 0], [1.0, 2.0, 8.0]])
    c = orc.Cell(H[0], H[1], H[2])
    for na, nb, nc in [(1, 0, 0), (-1, 2, 1), (0, -1, -1), (3, -2, 5)]:
        got = c.lattice_shift(na, nb, nc)
        want = float(na) * H[0] + float(nb) * H[1] + float(nc) * H[2]
        assert np.array_equal(np.array(got), want)


def test_plan_matches_survey_meshes():
    # src/knearest.rs:377-397, :430 on the benchmark boxes
    p = orc.Cell.ortho(16, 16, 16).plan(3.0)
    assert (p["nx"], p["ny"], p["nz"], p["max_reach"]) == (5, 5, 5, 3)
    assert p["cell_min"] == 16.0 / 5
    p = orc.Cell([40, 0, 0], [20, 34.64, 0], [8, 6, 38]).plan(3.0)
    assert (p["nx"], p["ny"], p["nz"], p["max_reach"]) == (11, 11, 12, 7)
    p = orc.Cell.ortho(120, 120, 360).plan(3.0)
    assert (p["nx"], p["ny"], p["nz"], p["ncell"], p["max_reach"]) == (40, 40, 120, 192000, 61)
    with pytest.raises(orc.OracleError, match="too fine"):
        orc.Cell.ortho(1e4, 1e4, 1e4).plan(1e-3)


# ------------------------------------------------- walk + heap (a6-a8)

def _run_case(cs, nthreads=1):
    x = np.array(cs["xyz"], dtype=np.float64)
    F = cs["frames"]
    n = len(x) // F
    hint = cs["hint"] if cs["hint"] is not None else 0.0
    mask = cs["mask"]
    return orc.knearest_many(x.reshape(F, n, 3) if F > 1 else x, cs["cell"], cs["k"],
                             mask=mask, cell_hint=hint, nthreads=nthreads)


@pytest.mark.parametrize("cs", CASES, ids=[c["name"] for c in CASES])
def test_known_answers(cs):
    nn, d2 = _run_case(cs)
    flat_nn, flat_d2 = nn.ravel(), d2.ravel()
    k = cs["k"]
    if cs["nn"] is not None:
        assert flat_nn.tolist() == cs["nn"]
    if cs["rows"]:
        for r, want in cs["rows"].items():
            assert flat_nn[int(r) * k:int(r) * k + len(want)].tolist() == want
    for slot, want in cs["d2"].items():
        assert almost(flat_d2[int(slot)], want)
    # unused slots are exactly -1 / f64::NAN bits
    unused = flat_nn == -1
    assert np.all(flat_d2.view(np.uint64)[unused] == 0x7FF8000000000000)
    if cs["brute"]:
        bn, bd = refimpl.brute_knn(cs["xyz"], cs["cell"], k, cs["mask"])
        assert np.array_equal(nn, bn)
        ok = np.isnan(bd) == np.isnan(d2)
        assert ok.all()
        fin = ~np.isnan(bd)
        assert all(almost(a, b) for a, b in zip(d2[fin], bd[fin]))
    if cs["wrap"]:
        assert flat_d2[0] < 1.0  # tests/agree.rs:370-374


def test_restated_brute_matches_numpy_27_image():
    for cs in CASES:
        if not cs["brute"] or cs["frames"] != 1:
            continue
        bn, bd = orc.knearest_brute(np.array(cs["xyz"]), cs["cell"], cs["k"], cs["mask"])
        rn, rd = refimpl.brute_knn(cs["xyz"], cs["cell"], cs["k"], cs["mask"])
        assert np.array_equal(bn, rn), cs["name"]


@pytest.mark.parametrize("cell,n,k,hint,seed", [
    ([12.0, 11.0, 13.0], 300, 6, 2.5, 5),
    ([[10.0, 0, 0], [5.0, 8.660254037844386, 0], [0, 0, 10.0]], 250, 8, 2.0, 6),
    ([[10.0, 0, 0], [3.0, 9.0, 0], [1.0, 2.0, 8.0]], 200, 5, 0.0, 7),
    ([[9.0, 0, 0], [2.5, 8.0, 0], [-1.0, 1.5, 7.0], [0.5, -0.25, 3.0]], 220, 12, 1.7, 8),
    ([5.0, 6.0, 7.0], 90, 30, 3.0, 9),
])
def test_walk_matches_27_image_brute_and_d2_bits(cell, n, k, hint, seed):
    rng = np.random.default_rng(seed)
    H, o = refimpl.rows_of(cell)
    pts = (rng.random((n, 3)) * 1.4 - 0.2) @ H + o
    mask = (rng.random(n) > 0.1).astype(np.int32)
    nn, d2 = orc.knearest_many(pts, cell, k, mask=mask, cell_hint=hint)
    bn, bd = refimpl.brute_knn(pts, cell, k, mask)
    assert np.array_equal(nn, bn)
    # exact bits: every reported d2 is one image's (q + s) - p distance of
    # the folded points; check it against the numpy restatement on the
    # folded coordinates (27 images suffice in these boxes)
    _, f, _, _ = refimpl.fold(pts, cell)
    D = refimpl.mic27_matrix(f, H)
    for i in range(n):
        for t in range(k):
            j = nn[i, t]
            if j < 0:
                assert np.isnan(d2[i, t])
                continue
            assert d2[i, t].view(np.uint64) == D[i, j].view(np.uint64)


def test_threaded_equals_serial_bitwise():
    from linkcell_amd import workloads as wl
    c = wl.triclinic_c2(n=20000)
    a = orc.knearest_many(c["xyz"], c["cell"], 16, mask=c["mask"], cell_hint=3.0, nthreads=1)
    b = orc.knearest_many(c["xyz"], c["cell"], 16, mask=c["mask"], cell_hint=3.0, nthreads=8)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))


def test_golden_c1_fixture_reproduces():
    from linkcell_amd import workloads as wl
    g = np.load(GOLD / "c1_uniform_oracle.npz")
    c1 = wl.uniform_c1()
    nn, d2, reach, cand = orc.knearest_many(c1["xyz"], c1["cell"], 12, cell_hint=3.0,
                                            nthreads=4, stats=True)
    assert np.array_equal(nn, g["nn"])
    assert np.array_equal(d2.view(np.uint64), g["d2_bits"])
    assert np.array_equal(reach, g["reach"])
    # survey figure: ~885 candidates per source at reach 1
    assert 850 < cand.mean() < 920 and reach.max() == 1


def test_sample_walk_matches_full_walk():
    from linkcell_amd import workloads as wl
    c = wl.triclinic_c2(n=30000)
    full = orc.knearest_many(c["xyz"], c["cell"], 16, mask=c["mask"], cell_hint=3.0, nthreads=8)
    src = np.arange(0, 30000, 7)
    nn, d2, *_ = orc.knearest_sample(c["xyz"], c["cell"], 16, src, mask=c["mask"], cell_hint=3.0)
    act = src[c["mask"][src] != 0]
    assert np.array_equal(nn[act], full[0][act])
    assert np.array_equal(d2[act].view(np.uint64), full[1][act].view(np.uint64))


# ------------------------------------------------ error contract (a8, b)

def test_error_contract():
    # src/capi.rs:233-265 order; messages src/error.rs:27-33
    with pytest.raises(orc.OracleError, match="no points"):
        orc.knearest_many(np.zeros((0, 3)), [10, 10, 10], 1)
    with pytest.raises(orc.OracleError, match="k must be at least 1"):
        orc.knearest_many(np.zeros((1, 3)), [10, 10, 10], 0)
    with pytest.raises(orc.OracleError, match="singular"):
        orc.knearest_many(np.zeros((2, 3)), [0, 10, 10], 1)
    with pytest.raises(orc.OracleError, match="too fine"):
        orc.knearest_many(np.zeros((2, 3)), [1e4, 1e4, 1e4], 1, cell_hint=1e-3)
    # all masked: success with -1 / NaN rows and no mesh planning
    nn, d2 = orc.knearest_many(np.zeros((2, 3)), [1e4, 1e4, 1e4], 2, mask=[0, 0],
                               cell_hint=1e-3)
    assert (nn == -1).all() and np.isnan(d2).all()


def test_isolated_and_short_rows():
    # min(k, n_active - 1) entries per row
    x = np.array([[0, 0, 0], [1, 0, 0], [5, 5, 5]], dtype=np.float64)
    nn, d2 = orc.knearest_many(x, [10, 10, 10], 5, mask=[1, 0, 1])
    assert nn[0].tolist() == [2, -1, -1, -1, -1]
    assert nn[1].tolist() == [-1] * 5
    nn, _ = orc.knearest_many(x[:1], [10, 10, 10], 3)
    assert nn.tolist() == [[-1, -1, -1]]


# ------------------------------------- KHeap ties at the k-th boundary (a7)

TIES = json.loads((GOLD / "tie_answers.json").read_text())["cases"]
PLANTS = {"replace_on_equal": 1, "last_argmax": 2, "canonical_index": 3,
          "increasing_lists": 4, "finish_descending_index": 5}


def _tie_rows_hold(cs, lib_path=None):
    nn, d2 = orc.knearest_many(np.array(cs["xyz"], dtype=np.float64), cs["cell"],
                               cs["k"], cell_hint=cs["hint"], lib_path=lib_path)
    ok = True
    for r, want in cs["rows"].items():
        ok &= nn[int(r)].tolist() == want
        ok &= d2[int(r)].tolist() == cs["d2"][r]
    return ok


@pytest.mark.parametrize("cs", TIES, ids=[c["name"] for c in TIES])
def test_tie_answers(cs):
    # hand-derived rows (tests/golden/make_golden.py): exact equality, every
    # d2 is a sum of squares of multiples of 0.5
    assert _tie_rows_hold(cs), cs["derivation"]


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    """The oracle source built once per planted mistake (on the CPU)."""
    import subprocess
    src = Path(orc.__file__).resolve().parent / "lc_oracle.c"
    d = tmp_path_factory.mktemp("plants")
    out = {}
    for name, code in PLANTS.items():
        so = d / f"liblcoracle_{name}.so"
        subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fPIC", "-shared",
                        "-pthread", f"-DLC_ORACLE_PLANT={code}", "-o", str(so), str(src),
                        "-lm"], check=True)
        out[name] = so
    return out


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_tie_answers_reject_planted_mistakes(planted, plant):
    # every misreading of the tie rules fails the fixtures that name it,
    # and at least one fixture
    failed = [cs["name"] for cs in TIES if not _tie_rows_hold(cs, planted[plant])]
    named = [cs["name"] for cs in TIES if plant in cs["fails_if"]]
    assert named, plant
    assert set(named) <= set(failed), (plant, named, failed)
