"""Periodic linked-cell k-nearest search on MI355X (gfx950).

Python surface of the drop-in library, mirroring the reference module
``linkcell`` (linkcell/__init__.py:1-11, python/src/lib.rs:318-417):

* ``knearest(xyz, cell, k, mask=None, cell_hint=None) -> (indices, dist2)``
* ``gpu_available() -> bool``
* ``__version__``

plus ``Workspace``, the Python face of ``linkcell::gpu::Workspace``.

Host (numpy / any CPU ``__dlpack__``) inputs go through the ``lc_knearest``
C API, device (torch on ROCm, DLPack device type 10, or CUDA type 2)
inputs through ``lc_gpu_knearest_many`` on device pointers. Both run the
HIP kernels; results are bit-identical to the reference CPU walk.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib

__all__ = ["__version__", "gpu_available", "knearest", "Workspace", "Error"]

# DLPack device types (dlpack.h): CPU, CUDA, CUDAHost, ROCM, ROCMHost.
_DL_CPU, _DL_CUDA, _DL_CUDA_HOST, _DL_ROCM, _DL_ROCM_HOST = 1, 2, 3, 10, 11


class Error(RuntimeError):
    """A failed search; the message is the C waist's error string."""


def _lib_handle():
    return _lib.load()


def __getattr__(name):
    # Resolved on first use so that importing the package (for example to
    # run linkcell_amd.build) never loads a stale or missing library.
    if name == "__version__":
        return _lib_handle().lc_version().decode()
    raise AttributeError(name)


def gpu_available() -> bool:
    return bool(_lib_handle().lc_gpu_available())


def _cell12(cell) -> list:
    """Cell shapes of python/src/lib.rs:235-308: (3,) ortho lengths, (3, 3)
    lattice rows, (4, 3) rows plus origin."""
    c = np.asarray(cell, dtype=np.float64)
    if c.shape == (3,):
        return [c[0], 0.0, 0.0, 0.0, c[1], 0.0, 0.0, 0.0, c[2], 0.0, 0.0, 0.0]
    if c.shape == (3, 3):
        return list(c.reshape(9)) + [0.0, 0.0, 0.0]
    if c.shape == (4, 3):
        return list(c.reshape(12))
    raise ValueError("cell must be float64 shape (3,) ortho, (3, 3) lattice rows, "
                     "or (4, 3) rows plus origin")


def _xyz_shape(shape) -> tuple:
    """(n, n_frames) for shapes (n, 3), (F, n, 3), (3n,)
    (python/src/lib.rs:213-233)."""
    shape = tuple(int(s) for s in shape)
    if len(shape) == 2 and shape[1] == 3 and shape[0] > 0:
        return shape[0], 1
    if len(shape) == 3 and shape[2] == 3 and shape[0] > 0 and shape[1] > 0:
        return shape[1], shape[0]
    if len(shape) == 1 and shape[0] > 0 and shape[0] % 3 == 0:
        return shape[0] // 3, 1
    raise ValueError("xyz must be float64 shape (n, 3), (n_frames, n, 3), or (n*3,)")


def _device_type(obj) -> int:
    dev = getattr(obj, "__dlpack_device__", None)
    if dev is None:
        return _DL_CPU
    return int(dev()[0])


def _is_torch(obj) -> bool:
    return type(obj).__module__.startswith("torch")


class Workspace:
    """One HIP stream plus grow-only device buffers
    (include/linkcell_gpu.hpp:29-78)."""

    def __init__(self):
        lib = _lib_handle()
        self._lib = lib
        self._ws = lib.lc_gpu_workspace_new()
        if not self._ws:
            msg = lib.lc_gpu_last_error()
            raise Error(msg.decode() if msg else "workspace creation failed")

    def close(self):
        if getattr(self, "_ws", None):
            self._lib.lc_gpu_workspace_free(self._ws)
            self._ws = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            msg = self._lib.lc_gpu_last_error()
            raise Error(msg.decode() if msg else "device search failed")

    @property
    def queue(self) -> int:
        return int(self._lib.lc_gpu_queue(self._ws) or 0)

    def wait(self):
        """Wait for the workspace stream; raises a failure reported there
        (a device-cell plan of a call made without waiting)."""
        self._lib.lc_gpu_wait(self._ws)
        msg = self._lib.lc_gpu_last_error()
        if msg:
            raise Error(msg.decode())

    def set_mode(self, mode: int):
        self._check(self._lib.lc_gpu_set_mode(self._ws, int(mode)))

    def set_timing(self, on: bool):
        self._check(self._lib.lc_gpu_set_timing(self._ws, 1 if on else 0))

    def last_timing(self) -> dict:
        """Milliseconds per pipeline stage of the last timed search (walk
        split into schedule set-up, first pass and retry pass), plus the
        deferred-source and retried-row counts."""
        ms = (C.c_double * 12)()
        self._check(self._lib.lc_gpu_last_timing_ex(self._ws, ms, 12))
        keys = ("fold", "scan", "scatter", "sort", "walk", "exact", "total",
                "setup", "pass1", "retry")
        out = {k: float(v) for k, v in zip(keys, ms)}
        out["deferred"] = int(ms[10])
        out["retried"] = int(ms[11])
        return out

    TIMING_KEYS = ("fold", "scan", "scatter", "sort", "walk", "exact", "total",
                   "setup", "pass1", "retry")

    def timing_history(self, n: int) -> list:
        """Stage milliseconds of up to n of the most recent timed searches,
        oldest first (the workspace keeps 64)."""
        ms = (C.c_double * (10 * max(1, n)))()
        got = C.c_int(0)
        self._check(self._lib.lc_gpu_timing_history(self._ws, ms, int(n), C.byref(got)))
        return [{k: float(ms[10 * i + j]) for j, k in enumerate(self.TIMING_KEYS)}
                for i in range(got.value)]

    def set_dense_min(self, v: int):
        self._check(self._lib.lc_gpu_set_dense_min(self._ws, int(v)))

    SCHED_CELL, SCHED_LEARNED, SCHED_GIVEN, SCHED_LEARNED_DEPTH = 0, 1, 2, 3

    def set_schedule(self, mode: int):
        """Order of the lane-per-source walk (results never change)."""
        self._check(self._lib.lc_gpu_set_schedule(self._ws, int(mode)))

    def sched_set_prediction(self, kth_ptr: int, n_cells: int):
        """Per-cell k-th-neighbour distances (device float32 array) for
        SCHED_GIVEN."""
        self._check(self._lib.lc_gpu_sched_set_prediction(
            self._ws, C.c_void_p(kth_ptr or None), int(n_cells)))

    def sched_set_tile_prediction(self, kth_ptr: int, n_tiles: int):
        """Per-tile k-th-neighbour distances (device float32 array, the
        workspace's tile order) for SCHED_GIVEN."""
        self._check(self._lib.lc_gpu_sched_set_tile_prediction(
            self._ws, C.c_void_p(kth_ptr or None), int(n_tiles)))

    def sched_set_source_prediction(self, kth_ptr: int, n_sources: int):
        """Per-source k-th-neighbour distances (device float32 array, the
        walk's slot order) for SCHED_GIVEN; 0 clears."""
        self._check(self._lib.lc_gpu_sched_set_source_prediction(
            self._ws, C.c_void_p(kth_ptr or None), int(n_sources)))

    def source_labels(self, out_ptr: int, n: int):
        """Final k-th-neighbour distance per source slot of the last search
        into a device float32 array."""
        self._check(self._lib.lc_gpu_sched_source_labels(self._ws, C.c_void_p(out_ptr),
                                                         int(n)))

    def tile_count(self) -> int:
        """Source tiles of the last search."""
        n = C.c_size_t(0)
        self._check(self._lib.lc_gpu_sched_tile_count(self._ws, C.byref(n)))
        return int(n.value)

    def tile_labels(self, out_ptr: int, n: int):
        """Largest final k-th-neighbour distance per source tile of the last
        search into a device float32 array."""
        self._check(self._lib.lc_gpu_sched_tile_labels(self._ws, C.c_void_p(out_ptr), int(n)))

    def sched_dump(self, n_cells: int, predict: bool = True) -> dict:
        """Scheduler diagnostics of the last search: the BF16 inputs of
        every tile as training collects them and (when trained, with
        predict) as inference reads them, inference's caps, and the tile
        layout (home cell, box, cell occupancy)."""
        nt = self.tile_count()
        xc = np.zeros((nt, 160), dtype=np.uint16)
        xp = np.zeros((nt, 160), dtype=np.uint16) if predict else None
        caps = np.zeros(nt, dtype=np.float32) if predict else None
        cells = np.zeros(nt, dtype=np.int32)
        box = np.zeros((nt, 8), dtype=np.float64)
        count = np.zeros(n_cells, dtype=np.int32)
        self._check(self._lib.lc_gpu_sched_dump(
            self._ws, nt, n_cells, xc.ctypes.data, xp.ctypes.data if predict else None,
            caps.ctypes.data if predict else None, cells.ctypes.data, box.ctypes.data,
            count.ctypes.data))
        return dict(x_collect=xc, x_predict=xp, caps=caps, tile_cell=cells, tile_box=box,
                    count=count)

    def last_retried(self) -> int:
        """Rows of the last timed search walked again without their cap."""
        r = C.c_int(0)
        self._check(self._lib.lc_gpu_last_retried(self._ws, C.byref(r)))
        return int(r.value)

    # learned scheduler ------------------------------------------------
    SCHED_PARAMS = 14529

    def sched_collect(self) -> int:
        """Append the source tiles of the last search (run with out_d2) to
        the predictor's training set; returns the samples added."""
        n = C.c_size_t(0)
        self._check(self._lib.lc_gpu_sched_collect(self._ws, C.byref(n)))
        return int(n.value)

    def sched_samples(self) -> int:
        n = C.c_size_t(0)
        self._check(self._lib.lc_gpu_sched_samples(self._ws, C.byref(n)))
        return int(n.value)

    def sched_clear(self):
        self._check(self._lib.lc_gpu_sched_clear(self._ws))

    def sched_train(self, epochs=4, batch=4096, lr=2e-3, seed=1, quantile=0.98):
        """Train the predictor on the collected set; returns (last epoch's
        mean loss, margin)."""
        loss = C.c_double(0.0)
        margin = C.c_double(0.0)
        self._check(self._lib.lc_gpu_sched_train(
            self._ws, int(epochs), int(batch), float(lr), int(seed) & 0xFFFFFFFF,
            float(quantile), C.byref(loss), C.byref(margin)))
        return float(loss.value), float(margin.value)

    def sched_get_params(self):
        p = np.zeros(self.SCHED_PARAMS, dtype=np.float32)
        margin = C.c_double(0.0)
        self._check(self._lib.lc_gpu_sched_get_params(
            self._ws, p.ctypes.data, p.size, C.byref(margin)))
        return p, float(margin.value)

    def sched_set_params(self, params, margin=0.0):
        p = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self._lib.lc_gpu_sched_set_params(
            self._ws, p.ctypes.data, p.size, float(margin)))

    def sched_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(x.shape[0], dtype=np.float32)
        self._check(self._lib.lc_gpu_sched_forward(self._ws, x.ctypes.data, x.shape[0],
                                                   y.ctypes.data))
        return y

    def sched_gradient(self, x, t, w):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        g = np.zeros(self.SCHED_PARAMS, dtype=np.float32)
        loss = C.c_double(0.0)
        self._check(self._lib.lc_gpu_sched_gradient(
            self._ws, x.ctypes.data, t.ctypes.data, w.ctypes.data, x.shape[0],
            g.ctypes.data, C.byref(loss)))
        return g, float(loss.value)

    def sched_adam(self, x, t, w, steps, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        loss = C.c_double(0.0)
        self._check(self._lib.lc_gpu_sched_adam(
            self._ws, x.ctypes.data, t.ctypes.data, w.ctypes.data, x.shape[0],
            int(steps), float(lr), C.byref(loss)))
        return float(loss.value)

    def set_stats(self, on: bool):
        self._check(self._lib.lc_gpu_set_stats(self._ws, 1 if on else 0))

    STAT_NAMES = ("dense_waves", "dense_tile_tests", "dense_tiles", "dense_steps",
                  "dense_merges", "dense_appends", "sparse_sources",
                  "sparse_tile_tests", "sparse_tiles", "sparse_evals",
                  "sparse_merges", "row_sources", "row_tile_tests", "row_tiles",
                  "row_merges", "row_rounds", "row_cells", "row_refines",
                  "row_useful", "row_cyc_fill", "row_cyc_drain", "row_cyc_merge",
                  "row_cyc_all", "row_flushes", "row_waves_r2", "row_rounds_r2",
                  "row_useful_r2", "row_cyc_r2", "row_dead", "row_accepts",
                  "row_dead_halves", "row_cyc_open", "row_cyc_init", "row_idle_done",
                  "row_idle_empty")

    def last_stats(self) -> dict:
        n = len(self.STAT_NAMES)
        buf = (C.c_ulonglong * n)()
        self._check(self._lib.lc_gpu_last_stats(self._ws, buf, n))
        return {k: int(v) for k, v in zip(self.STAT_NAMES, buf)}

    def knearest_ptr(self, xyz_ptr: int, n: int, n_frames: int, cell12, k: int,
                     out_nn_ptr: int, out_d2_ptr: int = 0, mask_ptr: int = 0,
                     cell_hint: float = 0.0, out_reach_ptr: int = 0,
                     wait: bool = True):
        """Raw device-pointer search (``lc_gpu_knearest_ex``)."""
        cell = _lib.lc_cell.from12(cell12)
        rc = self._lib.lc_gpu_knearest_ex(
            self._ws, C.c_void_p(xyz_ptr), n, n_frames, C.byref(cell), k,
            C.c_void_p(mask_ptr or None), float(cell_hint),
            C.c_void_p(out_nn_ptr), C.c_void_p(out_d2_ptr or None),
            C.c_void_p(out_reach_ptr or None), 1 if wait else 0)
        self._check(rc)

    def knearest_dcell_ptr(self, xyz_ptr: int, n: int, n_frames: int,
                           cell_ptr: int, cell_n: int, k: int, out_nn_ptr: int,
                           out_d2_ptr: int = 0, mask_ptr: int = 0,
                           cell_hint: float = 0.0, wait: bool = True):
        rc = self._lib.lc_gpu_knearest_many_dcell(
            self._ws, C.c_void_p(xyz_ptr), n, n_frames, C.c_void_p(cell_ptr),
            int(cell_n), k, C.c_void_p(mask_ptr or None), float(cell_hint),
            C.c_void_p(out_nn_ptr), C.c_void_p(out_d2_ptr or None),
            1 if wait else 0)
        self._check(rc)

    def set_dcell_capacity(self, cells_per_frame: int):
        """Cells per frame the device-cell path sizes its buffers for."""
        self._check(self._lib.lc_gpu_set_dcell_capacity(self._ws, int(cells_per_frame)))

    def dcell_plan(self, cell_ptr: int, cell_n: int, cell_hint: float = 0.0) -> dict:
        """The device's plan of a device cell (diagnostics)."""
        out = (C.c_double * 36)()
        self._check(self._lib.lc_gpu_dcell_plan(self._ws, C.c_void_p(cell_ptr), int(cell_n),
                                                float(cell_hint), out))
        v = list(out)
        return dict(h=v[0:9], hinv=v[9:18], origin=v[18:21], widths=v[21:24],
                    ortho=int(v[24]), cell_min=v[25], nx=int(v[26]), ny=int(v[27]),
                    nz=int(v[28]), ncell=int(v[29]), max_reach=int(v[30]), status=int(v[31]))

    @staticmethod
    def _frames(frames):
        """Build an ``lc_gpu_frame`` array from dicts with keys xyz, n,
        cell (12 values), k, and optional mask, cell_hint, out_nn, out_d2,
        out_reach; pointer values are device addresses (ints)."""
        frames = list(frames)
        arr = (_lib.lc_gpu_frame * max(1, len(frames)))()
        for i, fr in enumerate(frames):
            arr[i].xyz = fr["xyz"] or None
            arr[i].n = int(fr["n"])
            arr[i].cell = _lib.lc_cell.from12(fr["cell"])
            arr[i].k = int(fr["k"])
            arr[i].mask = fr.get("mask") or None
            arr[i].cell_hint = float(fr.get("cell_hint") or 0.0)
            arr[i].out_nn = fr["out_nn"] or None
            arr[i].out_d2 = fr.get("out_d2") or None
            arr[i].out_reach = fr.get("out_reach") or None
        return arr, len(frames)

    def knearest_batch(self, frames, wait: bool = True):
        """Uneven systems in one pipeline (``lc_gpu_knearest_batch``)."""
        arr, n = self._frames(frames)
        self._check(self._lib.lc_gpu_knearest_batch(self._ws, arr, n, 1 if wait else 0))

    def graph_capture(self, frames):
        """Capture the batch into a replayable hipGraph."""
        arr, n = self._frames(frames)
        self._check(self._lib.lc_gpu_graph_capture(self._ws, arr, n))

    def graph_launch(self, wait: bool = True):
        self._check(self._lib.lc_gpu_graph_launch(self._ws, 1 if wait else 0))

    def graph_reset(self):
        self._check(self._lib.lc_gpu_graph_reset(self._ws))

    def autotune(self, frames) -> tuple:
        """Time each walk split on the batch; keep and return the fastest
        as (dense_min, milliseconds)."""
        arr, n = self._frames(frames)
        best = C.c_int(0)
        ms = C.c_double(0.0)
        self._check(self._lib.lc_gpu_autotune(self._ws, arr, n, C.byref(best), C.byref(ms)))
        return int(best.value), float(ms.value)


_tls = threading.local()


def _workspace() -> Workspace:
    """This thread's workspace: each thread searches on its own stream and
    scratch (the reference binding creates one per call,
    python/src/gpu.rs:239; one per thread keeps the grow-only buffers)."""
    ws = getattr(_tls, "ws", None)
    if ws is None:
        ws = Workspace()
        _tls.ws = ws
    return ws


def _host_mask(mask, n):
    if mask is None:
        return None
    m = np.asarray(mask)
    if m.size != n:
        raise ValueError("mask length must be n")
    if m.dtype == np.int32:
        return np.ascontiguousarray(m.reshape(n))
    if m.dtype == np.float64:
        return np.ascontiguousarray((m.reshape(n) != 0.0).astype(np.int32))
    raise ValueError("mask must be int32 or float64")


def knearest(xyz, cell, k, mask=None, cell_hint=None):
    """k nearest periodic neighbours of every point.

    ``xyz`` is float64 ``(n, 3)``, ``(n_frames, n, 3)`` or ``(n*3,)``;
    ``cell`` is ``(3,)``, ``(3, 3)`` or ``(4, 3)`` float64 on any device;
    ``mask`` (length n, int32 or float64, nonzero keeps) drops points as
    source and candidate. Returns ``(indices, dist2)``: int32 with -1 and
    float64 with NaN in unused slots, shape ``(n, k)`` or
    ``(n_frames, n, k)``. Device inputs return device tensors.
    """
    k = int(k)
    if k <= 0:
        raise ValueError("k must be at least 1")
    hint = float(cell_hint) if cell_hint is not None else 0.0
    if _device_type(xyz) in (_DL_ROCM, _DL_CUDA):
        return _knearest_device(xyz, cell, k, mask, hint)
    if _is_torch(cell) and _device_type(cell) in (_DL_ROCM, _DL_CUDA):
        cell = cell.detach().cpu().numpy()
    x = np.asarray(xyz)
    if x.dtype != np.float64:
        raise ValueError("xyz must be float64 shape (n, 3), (n_frames, n, 3), or (n*3,)")
    n, frames = _xyz_shape(x.shape)
    x = np.ascontiguousarray(x)
    c12 = _cell12(cell)
    m = _host_mask(mask, n)
    nn = np.full(frames * n * k, -1, dtype=np.int32)
    d2 = np.full(frames * n * k, np.nan, dtype=np.float64)
    lib = _lib_handle()
    cobj = _lib.lc_cell.from12(c12)
    rc = lib.lc_knearest_many(
        x.ctypes.data, n, frames, C.byref(cobj), k,
        m.ctypes.data if m is not None else None, hint, nn.ctypes.data,
        d2.ctypes.data)
    if rc != 0:
        msg = lib.lc_last_error()
        raise Error(msg.decode() if msg else "knearest failed")
    shape = (n, k) if frames == 1 else (frames, n, k)
    return nn.reshape(shape), d2.reshape(shape)


def _knearest_device(xyz, cell, k, mask, hint):
    """Device search on the thread's workspace stream, ordered after
    torch's current stream by an event and before it by another (the
    reference binding orders on lc_gpu_queue, python/src/gpu.rs:252); the
    host does not wait."""
    import torch

    if not _is_torch(xyz):
        xyz = torch.from_dlpack(xyz)
    if xyz.dtype != torch.float64:
        raise ValueError("xyz must be float64 shape (n, 3), (n_frames, n, 3), or (n*3,)")
    n, frames = _xyz_shape(xyz.shape)
    xyz = xyz.contiguous()
    dev = xyz.device
    shape = (n, k) if frames == 1 else (frames, n, k)
    m_dev = None
    if mask is not None:
        if _is_torch(mask) and mask.is_cuda:
            # python/src/lib.rs:188-211: int32 as given, float64 nonzero keeps
            if mask.dtype == torch.int32:
                m_dev = mask.contiguous()
            elif mask.dtype == torch.float64:
                m_dev = (mask != 0).to(torch.int32).contiguous()
            else:
                raise ValueError("mask must be int32 or float64")
        else:
            m_dev = torch.from_numpy(_host_mask(mask, n)).to(dev)
        if m_dev.numel() != n:
            raise ValueError("mask length must be n")
    dcell = None
    if _is_torch(cell) and cell.is_cuda:
        if cell.dtype != torch.float64 or tuple(cell.shape) not in ((3,), (3, 3), (4, 3)):
            raise ValueError("cell must be float64 shape (3,) ortho, (3, 3) lattice rows, "
                             "or (4, 3) rows plus origin")
        dcell = cell.contiguous()
    else:
        c12 = _cell12(cell)
    nn = torch.empty(shape, dtype=torch.int32, device=dev)
    d2 = torch.empty(shape, dtype=torch.float64, device=dev)
    ws = _workspace()
    cur = torch.cuda.current_stream(dev)
    queue = torch.cuda.ExternalStream(ws.queue, device=dev)
    ready = torch.cuda.Event()
    ready.record(cur)
    queue.wait_event(ready)
    mptr = m_dev.data_ptr() if m_dev is not None else 0
    if dcell is not None:
        # waits, as the reference binding does (python/src/gpu.rs:302-322),
        # so a bad device cell raises here
        ws.knearest_dcell_ptr(xyz.data_ptr(), n, frames, dcell.data_ptr(), dcell.numel(),
                              k, nn.data_ptr(), d2.data_ptr(), mptr, hint, wait=True)
    else:
        ws.knearest_ptr(xyz.data_ptr(), n, frames, c12, k, nn.data_ptr(), d2.data_ptr(),
                        mptr, hint, wait=False)
    # the caching allocator must not hand these blocks out again before the
    # workspace stream is done with them
    for t in (xyz, nn, d2, m_dev, dcell):
        if t is not None:
            t.record_stream(queue)
    done = torch.cuda.Event()
    done.record(queue)
    cur.wait_event(done)
    return nn, d2
