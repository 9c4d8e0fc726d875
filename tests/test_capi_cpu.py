"""C-ABI library checks that need no GPU: the library loads, exports every
symbol the public headers declare, and its host-side contract (validation
order, messages, thread-local error slots) matches the reference
(tests/abi.rs, src/capi.rs:233-265, src/gpu/capi.cpp)."""
import ctypes as C
import re
import threading
from pathlib import Path

import numpy as np
import pytest

from linkcell_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def header_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"\b(lc_[a-z0-9_]+)\s*\(", text):
            if not text[m.start() - 7:m.start()].strip().endswith("inline"):
                names.add(m.group(1))
    names.discard("lc_cell_ortho")  # static inline helper, not exported
    return names


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    declared = header_functions()
    assert declared, "no declarations parsed"
    assert declared == set(_lib.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_cpp_workspace_symbols_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)],
                         capture_output=True, text=True, check=True).stdout
    for sym in ("_ZN8linkcell3gpu9WorkspaceC1Ev", "_ZN8linkcell3gpu9available",
                "knearest_into_many_dcell", "_ZN8linkcell3gpu9Workspace5queueEv"):
        assert sym in out, sym


def ortho_c(lx, ly, lz):
    return _lib.lc_cell(lx, 0, 0, 0, ly, 0, 0, 0, lz, 0, 0, 0)


def last_error():
    p = _lib.load().lc_last_error()
    return None if p is None else p.decode()


def test_version_and_empty_error():
    lib = _lib.load()
    assert lib.lc_version().decode() == "0.3.3"
    box = ortho_c(10, 10, 10)
    out = (C.c_int * 1)(-1)
    dummy = (C.c_double * 3)(0, 0, 0)
    assert lib.lc_knearest(dummy, 0, C.byref(box), 1, None, 0.0, out) != 0
    assert last_error() == "no points"
    lib.lc_version()  # does not touch the slot (tests/abi.rs:289-306)
    assert last_error() == "no points"


def test_zero_k_null_and_overflow_messages():
    lib = _lib.load()
    box = ortho_c(10, 10, 10)
    out = (C.c_int * 1)(-1)
    pt = (C.c_double * 3)(0, 0, 0)
    assert lib.lc_knearest(pt, 1, C.byref(box), 0, None, 0.0, out) != 0
    assert last_error() == "k must be at least 1"
    assert lib.lc_knearest(None, 1, C.byref(box), 1, None, 0.0, out) != 0
    assert last_error() == "null pointer"
    n = (2 ** 63 - 1) // 3 + 1  # tests/abi.rs:308-320
    assert lib.lc_knearest(pt, n, C.byref(box), 1, None, 0.0, out) != 0
    assert last_error() == "n * k overflows"


def test_bad_box_and_too_fine_mesh():
    lib = _lib.load()
    out = (C.c_int * 2)(-7, -7)
    d2 = (C.c_double * 2)(0, 0)
    pts = (C.c_double * 6)(0, 0, 0, 1, 0, 0)
    bad = ortho_c(0, 10, 10)
    assert lib.lc_knearest_d2(pts, 2, C.byref(bad), 1, None, 0.0, out, d2) != 0
    assert last_error() == "singular or non-positive cell"
    big = ortho_c(1e4, 1e4, 1e4)
    assert lib.lc_knearest_d2(pts, 2, C.byref(big), 1, None, 1e-3, out, d2) != 0
    assert last_error() == "linked-cell mesh is too fine"
    assert list(out) == [-1, -1]  # write_one filled before the search failed


def test_all_masked_succeeds_without_search():
    lib = _lib.load()
    out = (C.c_int * 4)(-7, -7, -7, -7)
    d2 = (C.c_double * 4)(0, 0, 0, 0)
    pts = (C.c_double * 6)(0, 0, 0, 1, 0, 0)
    mask = (C.c_int * 2)(0, 0)
    box = ortho_c(1e4, 1e4, 1e4)
    assert lib.lc_knearest_d2(pts, 2, C.byref(box), 2, mask, 1e-3, out, d2) == 0
    assert last_error() is None
    assert list(out) == [-1] * 4
    assert np.isnan(np.array(list(d2))).all()


def test_error_slots_are_thread_local():
    lib = _lib.load()
    seen = {}

    def run(tag, n, k):
        box = ortho_c(10, 10, 10)
        out = (C.c_int * 1)(-1)
        pt = (C.c_double * 3)(0, 0, 0)
        assert lib.lc_knearest(pt, n, C.byref(box), k, None, 0.0, out) != 0
        seen[tag] = last_error()

    a = threading.Thread(target=run, args=("a", 0, 1))
    b = threading.Thread(target=run, args=("b", 1, 0))
    a.start(); b.start(); a.join(); b.join()
    assert seen == {"a": "no points", "b": "k must be at least 1"}


def _no_gpu():
    return _lib.load().lc_gpu_available() == 0


@pytest.mark.skipif(not _no_gpu(), reason="a HIP device is present")
def test_gpu_waist_fails_cleanly_without_device():
    lib = _lib.load()
    assert lib.lc_gpu_available() == 0
    assert not lib.lc_gpu_workspace_new()
    assert lib.lc_gpu_last_error().decode() == "HIP device not available"
    p = C.c_void_p()
    assert lib.lc_gpu_alloc(C.byref(p), 64) != 0
    box = ortho_c(10, 10, 10)
    assert lib.lc_gpu_knearest(None, None, 1, C.byref(box), 1, None, 0.0, None, None) != 0
    assert lib.lc_gpu_last_error().decode() == "null pointer"
    assert lib.lc_gpu_fill_i32(None, 5, 0) != 0


def test_python_surface_host_checks():
    import linkcell_amd as lc
    assert lc.__version__ == "0.3.3"
    assert isinstance(lc.gpu_available(), bool)
    with pytest.raises(ValueError, match="k must"):
        lc.knearest(np.zeros((2, 3)), np.array([10.0, 10, 10]), 0)
    with pytest.raises(ValueError, match="cell must"):
        lc.knearest(np.zeros((2, 3)), np.zeros((2, 2)), 1)
    with pytest.raises(ValueError, match="xyz must"):
        lc.knearest(np.zeros((2, 4)), np.array([10.0, 10, 10]), 1)
