"""Write the golden fixtures under tests/golden/.

known_answers.json
    Inputs and asserted answers transcribed from the reference's own tests
    and doc tests (linkcell 0.3.3): tests/agree.rs, tests/abi.rs,
    tests/gpu_agree.cpp, tests/python/test_knearest.py, src/lib.rs:53-99,
    src/knearest.rs:179-190, :321-334, :495-507. "nn" is the full packed
    index row set when the reference asserts all of it, "rows" lists only
    the rows it asserts; "d2" maps flat slot -> value (tolerance 1e-12 as
    in the reference). "brute" marks cases the reference also checks
    against knearest_brute / the 27-image oracle.

tie_answers.json
    Hand-derived rows for exact d2 ties at and inside the k-th slot, the
    one part of KHeap (src/knearest.rs:106-158) that random inputs never
    reach. Every coordinate is a multiple of 0.5 in a box of edge 16 or 32
    (fold, shift and d2 are then exact), so each asserted row follows from
    the visit order alone: stencil cells lexicographic in (dx, dy, dz)
    (:436-438), each cell's list in decreasing index (:423-428), pushes
    kept while the heap has room, then the first argmax replaced on a
    strict < (:120-128), finish by (d2, index) (:146-158). The derivation
    of every row is written out below and stored with the case. None of
    these answers comes from the oracle.

c1_uniform_oracle.npz
    The oracle's full output for the N=4096 uniform 16^3 box, k=12, hint
    3.0 (LCG seed 1): int32 indices and the float64 d2 bit patterns as
    uint64, plus per-source final reach. A regression fixture for the
    oracle and the direct parity target of the HIP path.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

HEX = [[10.0, 0.0, 0.0], [5.0, 8.660254037844386, 0.0], [0.0, 0.0, 10.0]]
TRI = [[10.0, 0.0, 0.0], [3.0, 9.0, 0.0], [1.0, 2.0, 8.0]]
O10 = [10.0, 10.0, 10.0]


def lcg_box(n, L, seed):
    from linkcell_amd.workloads import lcg_points_box
    return lcg_points_box(n, L, seed).tolist()


def lcg_cell(n, H, seed):
    from linkcell_amd.workloads import lcg_points_in_cell
    return lcg_points_in_cell(n, H[0], H[1], H[2], seed).tolist()


def cases():
    out = []

    def add(name, cell, xyz, k, hint=None, mask=None, nn=None, rows=None,
            d2=None, brute=False, wrap=False, frames=1):
        out.append(dict(name=name, cell=cell, xyz=xyz, k=k, hint=hint,
                        mask=mask, nn=nn, rows=rows, d2=d2 or {}, brute=brute,
                        wrap=wrap, frames=frames))

    # tests/agree.rs:30-37
    add("two_points_k1", O10, [[0, 0, 0], [1, 0, 0]], 1, 2.0, rows={"0": [1], "1": [0]},
        d2={"0": 1.0})
    # tests/agree.rs:40-47
    add("periodic_image_is_nearer", O10, [[0.2, 0, 0], [9.4, 0, 0]], 1, 2.0,
        rows={"0": [1]}, d2={"0": 0.64})
    # tests/agree.rs:50-58, src/knearest.rs:321-334
    add("mask_drops_sources_and_candidates", O10, [[0, 0, 0], [1, 0, 0], [0, 1, 0]], 1,
        None, mask=[1, 0, 1], nn=[2, -1, 0])
    # tests/agree.rs:61-82
    add("random_ortho_64_k4", [12.0, 11.0, 13.0], lcg_box(64, (12.0, 11.0, 13.0), 1), 4,
        2.5, brute=True)
    # tests/agree.rs:95-107
    add("hex_image_beats_cartesian", HEX, [[0.2, 0.1, 1.0], [9.7, 0.1, 1.0]], 1, 2.0,
        rows={"0": [1]}, brute=True)
    # tests/agree.rs:109-132 (r = 10u + 5v, 8.660254037844386 v, 10 w)
    add("hex_48_k4", HEX, lcg_cell(48, HEX, 7), 4, 2.0, brute=True)
    # tests/agree.rs:230-254
    for cname, cell in (("hex", HEX), ("triclinic", TRI)):
        for label, n, k, seed, hint in (("n5_k1", 5, 1, 1, 2.0), ("n5_k4", 5, 4, 1, 2.0),
                                        ("n8_k1", 8, 1, 11, None), ("n8_k4", 8, 4, 11, 2.0),
                                        ("n12_k1", 12, 1, 3, 1.5), ("n12_k4", 12, 4, 3, None)):
            add(f"sheared_{cname}_{label}", cell, lcg_cell(n, cell, seed), k, hint, brute=True)
    # tests/agree.rs:256-379, all with hint 2.0
    wraps = [
        ("ortho-face-x k=1", O10, [[0.12, 5.0, 5.0], [9.79, 5.0, 5.0], [5.0, 5.0, 5.0]], 1),
        ("ortho-face-y k=1", O10, [[5.0, 0.12, 5.0], [5.0, 9.79, 5.0], [5.0, 5.0, 5.0]], 1),
        ("ortho-face-z k=1", O10, [[5.0, 5.0, 0.12], [5.0, 5.0, 9.79], [5.0, 5.0, 5.0]], 1),
        ("ortho-edge-xy k=1", O10, [[0.12, 0.11, 5.0], [9.80, 9.81, 5.0], [5.0, 5.0, 5.0]], 1),
        ("ortho-corner k=1", O10, [[0.11, 0.12, 0.13], [9.82, 9.81, 9.80], [5.0, 5.0, 5.0]], 1),
        ("ortho-faces k=4", O10, [[0.10, 5.0, 5.0], [9.85, 5.0, 5.0], [5.0, 0.11, 5.0],
                                  [5.0, 9.84, 5.0], [5.0, 5.0, 0.12], [5.0, 5.0, 9.83],
                                  [5.0, 5.0, 5.0]], 4),
        ("ortho-same-bin two images k=2", O10, [[0.1, 0.0, 0.0], [1.0, 0.0, 0.0],
                                                [9.8, 0.0, 0.0]], 2),
        ("hex-face-a k=1", HEX, [[0.20, 0.10, 5.0], [9.70, 0.10, 5.0], [5.0, 4.0, 5.0]], 1),
        ("hex-corner k=1", HEX, [[0.25, 0.20, 0.25], [14.75, 8.46, 9.75], [5.0, 4.3, 5.0]], 1),
        ("hex-faces k=4", HEX, [[0.20, 0.15, 5.0], [9.70, 0.15, 5.0], [5.2, 0.20, 5.0],
                                [7.7, 8.40, 5.0], [5.0, 4.3, 0.20], [5.0, 4.3, 9.75],
                                [5.0, 4.3, 5.0]], 4),
        ("triclinic-face-a k=1", TRI, [[0.20, 0.20, 4.0], [9.75, 0.20, 4.0], [6.0, 5.0, 4.0]], 1),
        ("triclinic-corner k=1", TRI, [[0.20, 0.20, 0.20], [13.80, 10.80, 7.80],
                                       [7.0, 5.5, 4.0]], 1),
        ("triclinic-wrap k=4", TRI, [[0.18, 0.15, 0.20], [13.82, 10.78, 7.79],
                                     [9.80, 0.25, 0.30], [0.25, 8.80, 0.22],
                                     [0.22, 0.30, 7.75], [7.0, 5.5, 4.0]], 4),
    ]
    for name, cell, xyz, k in wraps:
        rows = {"0": [1]} if (len(xyz) == 3 and k == 1) else None
        add("wrap " + name, cell, xyz, k, 2.0, rows=rows, brute=True, wrap=True)
    # tests/agree.rs:431-459 (k in 1, 2, 3, 5; default hint)
    for k in (1, 2, 3, 5):
        add(f"packed_layout_k{k}", O10, [[0, 0, 0], [1, 0, 0], [0, 2, 0], [0, 0, 3]], k,
            None, brute=True)
    # tests/agree.rs:461-481, tests/abi.rs:195-240
    add("mask_none_matches_all_ones", O10, [[0, 0, 0], [1.2, 0, 0], [0, 1.4, 0],
                                            [0, 0, 1.6], [2, 2, 2]], 3, None,
        mask=[1, 1, 1, 1, 1], brute=True)
    # tests/agree.rs:511-518, src/lib.rs:86-99, src/knearest.rs:179-190
    add("packed_row_major", O10, [[0, 0, 0], [1, 0, 0]], 2, None, nn=[1, -1, 0, -1])
    # tests/agree.rs:520-530, tests/abi.rs:322-343
    add("into_d2_periodic_image", O10, [[0.2, 0, 0], [9.4, 0, 0]], 1, None, nn=[1, 0],
        d2={"0": 0.64, "1": 0.64})
    # tests/agree.rs:532-547, tests/abi.rs:345-367, test_knearest.py:45-54
    add("into_many_two_frames", O10, [[0.2, 0, 0], [9.4, 0, 0], [0.2, 0, 0], [9.4, 0, 0]],
        1, None, nn=[1, 0, 1, 0], d2={"0": 0.64, "3": 0.64}, frames=2)
    # tests/agree.rs:549-558
    add("corner_wrap", O10, [[0.1, 0.1, 0.1], [9.8, 9.8, 9.8], [5, 5, 5]], 1, 2.0,
        rows={"0": [1]}, brute=True)
    # tests/abi.rs:157-193
    add("lc_knearest_matches_packed", O10, [[0, 0, 0], [1, 0, 0], [0, 1.5, 0], [0, 0, 2],
                                            [9.7, 0.2, 0.1]], 3, None, brute=True)
    # src/lib.rs:70-84: one bin for the whole box, two images
    add("one_bin_two_images", O10, [[0.1, 0, 0], [1.0, 0, 0], [9.8, 0, 0]], 2, 10.0,
        rows={"0": [2, 1]}, d2={"0": 0.09, "1": 0.81}, brute=True)
    # src/knearest.rs:495-507, src/cell.rs:255-267
    add("brute_doc_0.64", O10, [[0.2, 0, 0], [9.4, 0, 0]], 1, None, d2={"0": 0.64},
        brute=True)
    # tests/gpu_agree.cpp:69-101 (indices compared exactly)
    add("gpu_agree_ortho_face", O10, [[0.2, 1.0, 1.0], [9.7, 1.0, 1.0]], 1, None, brute=True)
    add("gpu_agree_hex_face", HEX, [[0.2, 0.1, 1.0], [9.7, 0.1, 1.0]], 1, None, brute=True)
    add("gpu_agree_triclinic_48_k4", TRI, lcg_cell(48, TRI, 7), 4, None, brute=True)
    # tests/python/test_knearest.py:26-33 (cell as lattice rows)
    add("python_lattice_rows", [[10, 0, 0], [0, 10, 0], [0, 0, 10]], [[0, 0, 0], [1, 0, 0]],
        1, None, rows={"0": [1]}, d2={"0": 1.0})
    return out


# Box of edge 16, hint 4.0: widths 16 (exact), mesh 4 x 4 x 4 of edge 4,
# cell_min 4, max_reach 3; s = x / 16 and 16 s are exact, bin = floor(x / 4).
# A source at (4, 4, 4) sits in cell (1, 1, 1); its shell-1 cells 0..2 per
# axis need no wrap, and cell offset (dx, dy, dz) is visited at position
# P = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) (home = 13).
B16 = [16.0, 16.0, 16.0]
B32 = [32.0, 32.0, 32.0]


def tie_cases():
    out = []

    def add(name, cell, hint, k, xyz, rows, d2, derivation, plants):
        out.append(dict(name=name, cell=cell, hint=hint, k=k, xyz=xyz, rows=rows,
                        d2=d2, derivation=derivation, fails_if=plants))

    add("tie_across_cells_first_visit_kept", B16, 4.0, 2,
        [[4, 4, 4], [4, 4, 3], [4, 3, 4], [3, 4, 4]],
        {"0": [2, 3]}, {"0": [1.0, 1.0]},
        "source 0 in cell (1,1,1); 3 in (0,1,1) at P4, 2 in (1,0,1) at P10, 1 in "
        "(1,1,0) at P12, each at d2 1. Pushes 3, 2 fill the heap; 1 meets the first "
        "argmax (slot 0, d2 1) and 1 < 1 fails, so it is dropped. Stop at reach 1 "
        "(1 <= 16). finish: (1,2), (1,3).",
        ["replace_on_equal", "canonical_index"])
    add("tie_in_one_cell_decreasing_index_kept", B16, 4.0, 2,
        [[2, 2, 2], [1, 2, 2], [3, 2, 2], [2, 1, 2]],
        {"0": [2, 3]}, {"0": [1.0, 1.0]},
        "all four points in cell (0,0,0); the only occupied stencil cell of source 0 "
        "is its home, whose list runs 3, 2, 1, 0. Pushes 3, 2 fill the heap, 1 (d2 1) "
        "is not < 1 and is dropped. finish: (1,2), (1,3).",
        ["replace_on_equal", "canonical_index", "increasing_lists"])
    add("tie_eviction_takes_first_argmax", B16, 4.0, 2,
        [[4, 4, 4], [2, 4, 4], [4, 2, 4], [4, 4, 3], [6, 4, 4]],
        {"0": [3, 2]}, {"0": [1.0, 4.0]},
        "visit order for source 0: 1 (P4, d2 4) -> slot 0, 2 (P10, d2 4) -> slot 1, "
        "3 (P12, d2 1): first argmax of [4, 4] is slot 0, 1 < 4, slot 0 := 3; "
        "4 (home P13, d2 4): argmax of [1, 4] is slot 1, 4 < 4 fails. Heap {3, 2}. "
        "finish: (1,3), (4,2).",
        ["replace_on_equal", "canonical_index", "last_argmax"])
    add("tie_eviction_by_slot_not_age", B16, 4.0, 2,
        [[4, 4, 4], [1, 4, 4], [4, 4, 2], [4, 2, 4], [5, 4, 4]],
        {"0": [4, 3]}, {"0": [1.0, 4.0]},
        "visit order for source 0: 1 (P4, d2 9) -> slot 0, 3 (P10, d2 4) -> slot 1, "
        "2 (P12, d2 4): argmax of [9, 4] is slot 0, slot 0 := 2 (heap [2:4, 3:4]); "
        "4 (home P13, d2 1): first argmax of [4, 4] is slot 0, slot 0 := 4, so the "
        "newer tied entry (2) leaves and the older (3) stays. finish: (1,4), (4,3).",
        ["canonical_index", "last_argmax"])
    add("tie_across_shells_inner_shell_kept", B32, 4.0, 2,
        [[4, 4, 4], [4, 4, 5], [31.5, 4, 4], [4, 4, 8.5]],
        {"0": [1, 3]}, {"0": [1.0, 20.25]},
        "box 32, hint 4: mesh 8 x 8 x 8, cell_min 4, max_reach 5. Source 0 in cell "
        "(1,1,1). Shell 1: home list 1, 0 pushes 1 (d2 1); (0,0,1) = cell (1,1,2) "
        "pushes 3 (d2 4.5^2 = 20.25). Full, but 20.25 > (1*4)^2: walk shell 2. Point "
        "2 (cell 7) is reached only at jx = -1 (offset dx = -2) with shift -32: "
        "((31.5 - 32) - 4)^2 = 20.25, not < 20.25, dropped. Stop at reach 2 "
        "(20.25 <= 64). finish: (1,1), (20.25,3).",
        ["replace_on_equal", "canonical_index"])
    add("equal_d2_inside_row_sorted_by_index", B16, 4.0, 3,
        [[4, 4, 4], [6, 4, 4], [4, 4, 5], [3, 4, 4]],
        {"0": [2, 3, 1]}, {"0": [1.0, 1.0, 4.0]},
        "three candidates for k = 3, all kept: 3 (P4, d2 1), then the home list 3, 2, "
        "1, 0 gives 2 (d2 1) and 1 (d2 4). finish sorts by (d2, index): (1,2), (1,3), "
        "(4,1).",
        ["finish_descending_index"])
    return out


def main():
    from oracle import lcoracle as orc
    from linkcell_amd import workloads as wl

    here = Path(__file__).resolve().parent
    cs = cases()
    (here / "known_answers.json").write_text(json.dumps(
        {"source": "linkcell 0.3.3 reference tests and doc tests", "cases": cs},
        indent=1))
    ties = tie_cases()
    (here / "tie_answers.json").write_text(json.dumps(
        {"source": "hand-derived from src/knearest.rs:106-158, :423-438 (see "
                   "tests/golden/make_golden.py)", "cases": ties}, indent=1))
    c1 = wl.uniform_c1()
    nn, d2, reach, cand = orc.knearest_many(c1["xyz"], c1["cell"], c1["k"],
                                            cell_hint=c1["hint"], stats=True)
    np.savez_compressed(here / "c1_uniform_oracle.npz", nn=nn,
                        d2_bits=d2.view(np.uint64), reach=reach, cand=cand,
                        oracle=np.array([orc.load().orc_version().decode()]))
    print(f"{len(cs)} known-answer cases; C1 fixture {nn.shape}")


if __name__ == "__main__":
    main()
