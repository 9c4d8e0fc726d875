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


def main():
    from oracle import lcoracle as orc
    from linkcell_amd import workloads as wl

    here = Path(__file__).resolve().parent
    cs = cases()
    (here / "known_answers.json").write_text(json.dumps(
        {"source": "linkcell 0.3.3 reference tests and doc tests", "cases": cs},
        indent=1))
    c1 = wl.uniform_c1()
    nn, d2, reach, cand = orc.knearest_many(c1["xyz"], c1["cell"], c1["k"],
                                            cell_hint=c1["hint"], stats=True)
    np.savez_compressed(here / "c1_uniform_oracle.npz", nn=nn,
                        d2_bits=d2.view(np.uint64), reach=reach, cand=cand,
                        oracle=np.array([orc.load().orc_version().decode()]))
    print(f"{len(cs)} known-answer cases; C1 fixture {nn.shape}")


if __name__ == "__main__":
    main()
