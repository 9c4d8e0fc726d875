"""ctypes face of the CPU oracle (oracle/liblcoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liblcoracle.so"

MESSAGES = {
    1: "k must be at least 1",
    2: "singular or non-positive cell",
    3: "no points",
    4: "out buffer length must be n * k",
    5: "n * k overflows",
    6: "mask length must be n",
    7: "linked-cell mesh is too fine",
    8: "null pointer",
}


class OracleError(RuntimeError):
    def __init__(self, code):
        super().__init__(MESSAGES.get(code, f"oracle error {code}"))
        self.code = code


class _Cell(C.Structure):
    _fields_ = [("h", C.c_double * 9), ("hinv", C.c_double * 9),
                ("origin", C.c_double * 3), ("widths", C.c_double * 3),
                ("ortho", C.c_int)]


class _Plan(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("ncell", C.c_int64), ("cell_min", C.c_double),
                ("max_reach", C.c_int)]


_lib = None
_variants = {}


def load(path=None):
    """The oracle library; `path` loads another build of the same source
    (the tie tests' planted-mistake variants)."""
    global _lib
    if path is not None:
        path = str(path)
        if path not in _variants:
            _variants[path] = _bind(C.CDLL(path))
        return _variants[path]
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} missing; run `make -C oracle`")
    _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def _bind(lib):
    vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
    lib.orc_cell_from_vectors.argtypes = [dp, dp, dp, dp, C.POINTER(_Cell)]
    lib.orc_cell_from_vectors.restype = C.c_int
    lib.orc_fractional.argtypes = [C.POINTER(_Cell), dp, dp]
    lib.orc_cartesian.argtypes = [C.POINTER(_Cell), dp, dp]
    lib.orc_lattice_shift.argtypes = [C.POINTER(_Cell), C.c_int, C.c_int, C.c_int, dp]
    lib.orc_dist2_shifted.argtypes = [dp, dp, dp]
    lib.orc_dist2_shifted.restype = C.c_double
    lib.orc_dist2.argtypes = [C.POINTER(_Cell), dp, dp]
    lib.orc_dist2.restype = C.c_double
    lib.orc_plan_mesh.argtypes = [C.POINTER(_Cell), C.c_double, C.POINTER(_Plan)]
    lib.orc_plan_mesh.restype = C.c_int
    lib.orc_fold_bin.argtypes = [C.POINTER(_Cell), C.POINTER(_Plan), dp, dp,
                                 C.POINTER(C.c_int)]
    lib.orc_knearest_many.argtypes = [vp, sz, sz, dp, sz, vp, C.c_double, vp, vp,
                                      vp, vp, C.c_int]
    lib.orc_knearest_many.restype = C.c_int
    lib.orc_knearest_sample.argtypes = [vp, sz, dp, sz, vp, C.c_double, vp, sz, vp,
                                        vp, vp, vp, C.c_int, dp]
    lib.orc_knearest_sample.restype = C.c_int
    lib.orc_knearest_brute.argtypes = [vp, sz, dp, sz, vp, vp, vp]
    lib.orc_knearest_brute.restype = C.c_int
    lib.orc_version.restype = C.c_char_p
    return lib


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def cell12_of(cell) -> np.ndarray:
    """(3,), (3, 3) or (4, 3) cell -> 12 doubles (rows a, b, c, origin)."""
    c = np.asarray(cell, dtype=np.float64)
    if c.shape == (3,):
        return np.array([c[0], 0, 0, 0, c[1], 0, 0, 0, c[2], 0, 0, 0], dtype=np.float64)
    if c.shape == (3, 3):
        return np.concatenate([c.reshape(9), np.zeros(3)])
    if c.shape == (4, 3):
        return c.reshape(12).copy()
    if c.shape == (12,):
        return c.copy()
    raise ValueError("bad cell shape")


class Cell:
    """Cell::from_vectors restated (src/cell.rs:97-123)."""

    def __init__(self, a, b, c, origin=(0.0, 0.0, 0.0)):
        lib = load()
        self._c = _Cell()
        rc = lib.orc_cell_from_vectors(_d3(a), _d3(b), _d3(c), _d3(origin),
                                       C.byref(self._c))
        if rc:
            raise OracleError(rc)

    @classmethod
    def ortho(cls, lx, ly, lz):
        return cls([lx, 0, 0], [0, ly, 0], [0, 0, lz])

    @property
    def widths(self):
        return list(self._c.widths)

    @property
    def is_ortho(self):
        return bool(self._c.ortho)

    @property
    def hinv(self):
        return list(self._c.hinv)

    def fractional(self, r):
        out = (C.c_double * 3)()
        load().orc_fractional(C.byref(self._c), _d3(r), out)
        return list(out)

    def cartesian(self, s):
        out = (C.c_double * 3)()
        load().orc_cartesian(C.byref(self._c), _d3(s), out)
        return list(out)

    def lattice_shift(self, na, nb, nc):
        out = (C.c_double * 3)()
        load().orc_lattice_shift(C.byref(self._c), na, nb, nc, out)
        return list(out)

    def dist2(self, p, q):
        return load().orc_dist2(C.byref(self._c), _d3(p), _d3(q))

    def plan(self, hint=float("nan")):
        p = _Plan()
        rc = load().orc_plan_mesh(C.byref(self._c), float(hint), C.byref(p))
        if rc:
            raise OracleError(rc)
        return dict(nx=p.nx, ny=p.ny, nz=p.nz, ncell=p.ncell,
                    cell_min=p.cell_min, max_reach=p.max_reach)


def dist2_shifted(p, q, shift):
    return load().orc_dist2_shifted(_d3(p), _d3(q), _d3(shift))


def _threads(nthreads):
    if nthreads is None:
        try:
            return len(os.sched_getaffinity(0))
        except AttributeError:
            return os.cpu_count() or 1
    return int(nthreads)


def knearest_many(xyz, cell, k, mask=None, cell_hint=0.0, n_frames=None,
                  nthreads=1, stats=False, lib_path=None):
    """lc_knearest_many restated. xyz: (n, 3) or (F, n, 3) float64.
    Returns (nn, d2) shaped (F*n, k) [plus (reach, cand) when stats]."""
    lib = load(lib_path)
    x = np.ascontiguousarray(xyz, dtype=np.float64)
    if x.ndim == 3:
        F, n = x.shape[0], x.shape[1]
    else:
        F, n = 1, x.reshape(-1, 3).shape[0]
    if n_frames is not None:
        F = n_frames
        n = x.size // 3 // F
    c12 = cell12_of(cell)
    nn = np.full((F * n, k), -7, dtype=np.int32)
    d2 = np.full((F * n, k), -7.0, dtype=np.float64)
    reach = np.zeros(F * n, dtype=np.int32)
    cand = np.zeros(F * n, dtype=np.int32)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    rc = lib.orc_knearest_many(
        x.ctypes.data, n, F, c12.ctypes.data_as(C.POINTER(C.c_double)), k,
        m.ctypes.data if m is not None else None, float(cell_hint),
        nn.ctypes.data, d2.ctypes.data, reach.ctypes.data, cand.ctypes.data,
        _threads(nthreads))
    if rc:
        raise OracleError(rc)
    if stats:
        return nn, d2, reach, cand
    return nn, d2


def knearest_sample(xyz, cell, k, sources, mask=None, cell_hint=0.0, nthreads=None):
    """Walk only `sources` (the CPU baseline's bounded sample).
    Returns (nn, d2, reach, cand, prep_seconds, walk_seconds)."""
    lib = load()
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    c12 = cell12_of(cell)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    nn = np.full((n, k), -7, dtype=np.int32)
    d2 = np.full((n, k), -7.0, dtype=np.float64)
    reach = np.zeros(n, dtype=np.int32)
    cand = np.zeros(n, dtype=np.int32)
    times = (C.c_double * 2)()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    rc = lib.orc_knearest_sample(
        x.ctypes.data, n, c12.ctypes.data_as(C.POINTER(C.c_double)), k,
        m.ctypes.data if m is not None else None, float(cell_hint),
        src.ctypes.data, src.size, nn.ctypes.data, d2.ctypes.data,
        reach.ctypes.data, cand.ctypes.data, _threads(nthreads), times)
    if rc:
        raise OracleError(rc)
    return nn, d2, reach, cand, times[0], times[1]


def knearest_brute(xyz, cell, k, mask=None):
    """knearest_brute restated (src/knearest.rs:508-541), packed."""
    lib = load()
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0]
    c12 = cell12_of(cell)
    nn = np.full((n, k), -7, dtype=np.int32)
    d2 = np.full((n, k), -7.0, dtype=np.float64)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    rc = lib.orc_knearest_brute(x.ctypes.data, n, c12.ctypes.data_as(C.POINTER(C.c_double)),
                                k, m.ctypes.data if m is not None else None,
                                nn.ctypes.data, d2.ctypes.data)
    if rc:
        raise OracleError(rc)
    return nn, d2
