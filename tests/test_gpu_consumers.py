"""The reference boundary on the device, from compiled consumers.

tests/consumers/gpu_agree.cpp is the C++ consumer of
linkcell::gpu::Workspace (after the reference's tests/gpu_agree.cpp:21-101:
host knearest_into against Workspace::knearest_into on device buffers, and
knearest_into_many with wait = false, then wait()). tests/consumers/
capi_device.c drives the lc_gpu_* waist in the order of the reference's
Python binding (python/src/gpu.rs:252-359: lc_gpu_alloc, lc_gpu_memcpy kind
1, lc_gpu_fill_i32 -1, lc_gpu_knearest_many wait = 0, lc_gpu_wait,
lc_gpu_memcpy kind 2) and then lc_gpu_knearest. Both are built here with
the host compilers against include/ and the in-tree library only; their
device outputs are compared bit for bit with the oracle."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gpu_util import cell12, same_bits
from oracle import lcoracle as orc
from linkcell_amd import workloads as wl

ROOT = Path(__file__).resolve().parents[1]
SRC = Path(__file__).resolve().parent / "consumers"
LIBDIR = ROOT / "linkcell_amd"


def _build(tmp, src, compiler, std):
    exe = tmp / src.stem
    cmd = [compiler, std, "-O1", "-I", str(ROOT / "include"), str(src), "-o", str(exe),
           "-L", str(LIBDIR), "-llinkcell_amd", f"-Wl,-rpath,{LIBDIR}"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return exe


@pytest.fixture(scope="module")
def consumers(tmp_path_factory):
    if not (LIBDIR / "liblinkcell_amd.so").exists():
        pytest.fail("liblinkcell_amd.so is not built")
    for cc in ("gcc", "g++"):
        if shutil.which(cc) is None:
            pytest.skip(f"{cc} not installed")
    tmp = tmp_path_factory.mktemp("consumers")
    return {"cpp": _build(tmp, SRC / "gpu_agree.cpp", "g++", "-std=c++17"),
            "c": _build(tmp, SRC / "capi_device.c", "gcc", "-std=c11")}


def test_consumers_build(consumers):
    # compiles and links against the public headers and the library alone
    assert consumers["cpp"].exists() and consumers["c"].exists()


def _write_input(path, xyz, cell, k, mask, hint, frames):
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(x) // frames
    with open(path, "wb") as f:
        np.array([n, frames, k, 0 if mask is None else 1], dtype=np.int64).tofile(f)
        np.array([hint], dtype=np.float64).tofile(f)
        np.array(cell12(cell), dtype=np.float64).tofile(f)
        x.tofile(f)
        if mask is not None:
            np.ascontiguousarray(mask, dtype=np.int32).tofile(f)
    return n


def _read_output(path, m, parts):
    raw = Path(path).read_bytes()
    out, off = [], 0
    for _ in range(parts):
        nn = np.frombuffer(raw, dtype=np.int32, count=m, offset=off)
        off += 4 * m
        d2 = np.frombuffer(raw, dtype=np.float64, count=m, offset=off)
        off += 8 * m
        out.append((nn, d2))
    assert off == len(raw)
    return out


HEX = [[10.0, 0.0, 0.0], [5.0, 8.660254037844386, 0.0], [0.0, 0.0, 10.0]]
TRI = [[10.0, 0.0, 0.0], [3.0, 9.0, 0.0], [1.0, 2.0, 8.0]]


def _cases():
    # tests/gpu_agree.cpp:71-101, then larger masked and multi-frame inputs
    c2 = wl.triclinic_c2(n=20000)
    rng = np.random.default_rng(31)
    H = np.array([[14.0, 0, 0], [3.0, 13.0, 0], [-2.0, 1.5, 12.0]])
    frames3 = (rng.random((3, 700, 3)) * 1.2 - 0.1) @ H
    return [
        ("ortho-face", [[0.2, 1.0, 1.0], [9.7, 1.0, 1.0]], [10.0, 10.0, 10.0], 1, None,
         0.0, 1),
        ("hex-prism-face", [[0.2, 0.1, 1.0], [9.7, 0.1, 1.0]], HEX, 1, None, 0.0, 1),
        ("triclinic-48-k4", wl.lcg_points_in_cell(48, TRI[0], TRI[1], TRI[2], 7), TRI, 4,
         None, 0.0, 1),
        ("c2-20000-masked-k16", c2["xyz"], c2["cell"], 16, c2["mask"], 3.0, 1),
        ("three-frames-masked-k24", frames3, H, 24, (rng.random(700) > 0.2).astype(np.int32),
         2.2, 3),
    ]


@pytest.mark.gpu
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0] if isinstance(c, tuple) else "")
def test_consumers_match_oracle(consumers, tmp_path, case):
    name, xyz, cell, k, mask, hint, frames = case
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    n = _write_input(inp, xyz, cell, k, mask, hint, frames)
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    on, od = orc.knearest_many(x.reshape(frames, n, 3) if frames > 1 else x, cell, k,
                               mask=mask, cell_hint=hint, nthreads=None)
    on, od = on.ravel(), od.ravel()
    m = frames * n * k
    for exe, parts in ((consumers["cpp"], 1), (consumers["c"], 2 if frames == 1 else 1)):
        res = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True,
                             timeout=120)
        print(res.stdout, res.stderr, flush=True)
        assert res.returncode == 0, (name, exe.name, res.stderr)
        for nn, d2 in _read_output(out, m, parts):
            assert np.array_equal(nn, on), (name, exe.name)
            assert same_bits(d2, od), (name, exe.name)
