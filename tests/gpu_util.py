"""Helpers for GPU tests: run the HIP path through the C-ABI on torch
device buffers and return host numpy arrays."""
from __future__ import annotations

import numpy as np


def cell12(cell):
    c = np.asarray(cell, dtype=np.float64)
    if c.shape == (3,):
        return [c[0], 0, 0, 0, c[1], 0, 0, 0, c[2], 0, 0, 0]
    if c.shape == (3, 3):
        return list(c.reshape(9)) + [0.0, 0.0, 0.0]
    return list(c.reshape(12))


def run_device(ws, xyz, cell, k, mask=None, hint=0.0, frames=1, reach=False,
               fill=-7):
    """lc_gpu_knearest_ex on device copies of the inputs. Outputs start
    filled with a sentinel so untouched slots are caught."""
    import torch

    dev = torch.device("cuda:0")
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = x.shape[0] // frames
    dx = torch.from_numpy(x).to(dev)
    nn = torch.full((frames * n, k), fill, dtype=torch.int32, device=dev)
    d2 = torch.full((frames * n, k), float(fill), dtype=torch.float64, device=dev)
    rr = torch.full((frames * n,), -1, dtype=torch.int32, device=dev) if reach else None
    dm = None
    if mask is not None:
        dm = torch.from_numpy(np.ascontiguousarray(mask, dtype=np.int32)).to(dev)
    torch.cuda.synchronize()
    ws.knearest_ptr(dx.data_ptr(), n, frames, cell12(cell), k, nn.data_ptr(),
                    d2.data_ptr(), dm.data_ptr() if dm is not None else 0,
                    float(hint), rr.data_ptr() if rr is not None else 0, True)
    out = (nn.cpu().numpy(), d2.cpu().numpy())
    if reach:
        out = out + (rr.cpu().numpy(),)
    return out


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
