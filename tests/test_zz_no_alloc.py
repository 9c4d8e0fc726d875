"""lc_gpu_knearest_noalloc: the search that allocates nothing (fixed device
scratch), through the C-ABI, against the existing entry point
lc_gpu_knearest_many on the same inputs, bit for bit, on consecutive calls
of decreasing size; its size limit is an error, not a fault; captured into
a graph (where an allocation would fail the capture) it replays to the same
bits."""
import ctypes as C

import numpy as np
import pytest

from gpu_util import cell12, same_bits
from linkcell_amd import _lib
from linkcell_amd import workloads as wl
from oracle import lcoracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib(gpu):
    return _lib.load()


def _device(xyz, mask):
    import torch
    dev = torch.device("cuda:0")
    x = torch.from_numpy(np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)).to(dev)
    m = None if mask is None else torch.from_numpy(np.ascontiguousarray(mask, np.int32)).to(dev)
    return x, m


def _outputs(rows, k, fill):
    import torch
    dev = torch.device("cuda:0")
    return (torch.full((rows, k), fill, dtype=torch.int32, device=dev),
            torch.full((rows, k), float(fill), dtype=torch.float64, device=dev))


def _noalloc(lib, x, n, frames, cell, k, m, hint, nn, d2, stream=None):
    import torch
    st = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    box = _lib.lc_cell.from12(cell12(cell))
    return lib.lc_gpu_knearest_noalloc(
        C.c_void_p(x.data_ptr()), n, frames, C.byref(box), k,
        C.c_void_p(m.data_ptr() if m is not None else None), float(hint),
        C.c_void_p(nn.data_ptr()), C.c_void_p(d2.data_ptr()), C.c_void_p(st))


def _many(lib, ws, x, n, frames, cell, k, m, hint, nn, d2):
    box = _lib.lc_cell.from12(cell12(cell))
    return lib.lc_gpu_knearest_many(
        ws, C.c_void_p(x.data_ptr()), n, frames, C.byref(box), k,
        C.c_void_p(m.data_ptr() if m is not None else None), float(hint),
        C.c_void_p(nn.data_ptr()), C.c_void_p(d2.data_ptr()), 1)


def _cases():
    rng = np.random.default_rng(77)
    c2 = wl.triclinic_c2(n=60000)
    two = np.concatenate([c2["xyz"], c2["xyz"][::-1].copy()])
    small = rng.random((3000, 3)) * np.array([14.0, 11.0, 9.0])
    tiny = rng.random((400, 3)) * 6.0
    return [
        # (xyz, n, frames, cell, k, mask, hint)
        (two, 60000, 2, c2["cell"], 16, c2["mask"], 3.0),
        (small, 3000, 1, [14.0, 11.0, 9.0], 8, None, 2.0),
        (tiny, 400, 1, [6.0, 6.0, 6.0], 150, None, 3.0),
    ]


def test_noalloc_matches_knearest_many_on_shrinking_calls(lib):
    import torch
    ws = lib.lc_gpu_workspace_new()
    assert ws
    try:
        for i, (xyz, n, frames, cell, k, mask, hint) in enumerate(_cases()):
            print(f"noalloc case {i}: n={n} frames={frames} k={k}", flush=True)
            x, m = _device(xyz, mask)
            a_nn, a_d2 = _outputs(n * frames, k, -7)
            b_nn, b_d2 = _outputs(n * frames, k, -9)
            torch.cuda.synchronize()
            assert _many(lib, ws, x, n, frames, cell, k, m, hint, a_nn, a_d2) == 0, \
                lib.lc_gpu_last_error()
            rc = _noalloc(lib, x, n, frames, cell, k, m, hint, b_nn, b_d2)
            assert rc == 0, lib.lc_gpu_last_error()
            torch.cuda.synchronize()
            assert torch.equal(a_nn, b_nn), i
            assert torch.equal(a_d2.view(torch.int64), b_d2.view(torch.int64)), i
            if frames == 1:
                on, od = orc.knearest_many(xyz, cell, k, mask=mask, cell_hint=hint,
                                           nthreads=None)
                assert np.array_equal(b_nn.cpu().numpy(), on), i
                assert same_bits(b_d2.cpu().numpy(), od), i
    finally:
        lib.lc_gpu_workspace_free(ws)


def test_noalloc_size_limit_is_an_error(lib):
    import torch
    assert lib.lc_gpu_noalloc_scratch_bytes() == 64 << 20
    rng = np.random.default_rng(3)
    xyz = rng.random((1200000, 3)) * 100.0
    x, _ = _device(xyz, None)
    nn, d2 = _outputs(len(xyz), 4, -7)
    torch.cuda.synchronize()
    rc = _noalloc(lib, x, len(xyz), 1, [100.0, 100.0, 100.0], 4, None, 3.0, nn, d2)
    assert rc == 1
    msg = lib.lc_gpu_last_error().decode()
    assert "bytes of scratch" in msg and str(64 << 20) in msg
    torch.cuda.synchronize()
    assert (nn == -7).all()  # nothing ran
    # argument errors keep the reference's messages
    box = _lib.lc_cell.from12(cell12([10.0, 10.0, 10.0]))
    assert lib.lc_gpu_knearest_noalloc(C.c_void_p(x.data_ptr()), 4, 1, C.byref(box), 0,
                                       None, 0.0, C.c_void_p(nn.data_ptr()), None,
                                       None) == 1
    assert lib.lc_gpu_last_error().decode() == "k must be at least 1"


def test_noalloc_replays_under_graph_capture(lib):
    import torch
    rng = np.random.default_rng(5)
    xyz = rng.random((20000, 3)) * 24.0
    x, _ = _device(xyz, None)
    nn, d2 = _outputs(20000, 12, -7)
    ref_nn, ref_d2 = _outputs(20000, 12, -9)
    torch.cuda.synchronize()
    assert _noalloc(lib, x, 20000, 1, [24.0, 24.0, 24.0], 12, None, 2.5, ref_nn, ref_d2) == 0
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, capture_error_mode="global"):
        rc = _noalloc(lib, x, 20000, 1, [24.0, 24.0, 24.0], 12, None, 2.5, nn, d2)
    assert rc == 0, lib.lc_gpu_last_error()
    for rep in range(3):
        nn.fill_(-3)
        d2.fill_(-3.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(nn, ref_nn), rep
        assert torch.equal(d2.view(torch.int64), ref_d2.view(torch.int64)), rep
        # eager work between replays
        y = (x * 2.0).sum()
        torch.cuda.synchronize()
        del y
