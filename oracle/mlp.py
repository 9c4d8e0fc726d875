"""CPU oracle of the learned scheduler's network (numpy, FP64).

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere in the product.

The network has no counterpart in the reference; this is the restatement the
HIP kernels (linkcell_amd/csrc/lc_sched.hip) are checked against:

    z1 = X W1 + b1,  a1 = relu(z1)
    z2 = a1 W2 + b2, a2 = relu(z2)
    y  = a2 . w3 + b3
    loss = sum_i w_i (y_i - t_i)^2 / (2 sum_i w_i)

Parameters are one flat vector: W1 (160 x 64, input-major), b1 (64),
W2 (64 x 64, input-major), b2, w3, b3 -- 14,529 values.

``emulate=True`` rounds to BF16 (round to nearest even) exactly where the
device does: the inputs, W1 and W2 (every MFMA operand), relu(z1) before the
second layer and the first weight gradient, dz2 before da1 = dz2 W2^T and
the W2 gradient, dz1 before the W1 gradient. Biases, w3, a2 and all sums
stay unrounded (FP32 on the device, FP64 here). ``emulate=False`` is the
plain FP64 network, whose gradient is pinned by finite differences in
tests/test_sched_oracle.py.

Adam: m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2,
p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps)  (Kingma & Ba 2015,
Algorithm 1).
"""
from __future__ import annotations

import numpy as np

FEAT, HID = 160, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + FEAT * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAMS = OFF_B3 + 1


def bf16(x):
    """Round to BF16 (nearest even) through FP32, returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, (u >> 16) | 0x40, r).astype(np.uint32)
    return (r << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(FEAT, HID), p[OFF_B1:OFF_W2],
            p[OFF_W2:OFF_B2].reshape(HID, HID), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, w3, [b3]]).astype(np.float64)


def _r(x, emulate):
    return bf16(x) if emulate else np.asarray(x, dtype=np.float64)


def forward(p, X, emulate=True):
    """Activations of the network for inputs X (n x 160)."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xr = _r(X, emulate)
    z1 = Xr @ _r(W1, emulate) + b1
    a1 = np.maximum(z1, 0.0)
    a1r = _r(a1, emulate)
    z2 = a1r @ _r(W2, emulate) + b2
    a2 = np.maximum(z2, 0.0)
    y = a2 @ w3 + b3
    return dict(X=Xr, z1=z1, a1=a1r, z2=z2, a2=a2, y=y)


def loss_grad(p, X, t, w, emulate=True):
    """(loss, gradient) of the weighted half mean square error."""
    W1, b1, W2, b2, w3, b3 = unpack(p)
    f = forward(p, X, emulate)
    t = np.asarray(t, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    inv = 1.0 / w.sum()
    e = f["y"] - t
    loss = 0.5 * float(np.sum(w * e * e)) * inv
    dy = w * e * inv
    g_w3 = f["a2"].T @ dy
    g_b3 = dy.sum()
    dz2 = np.where(f["z2"] > 0.0, dy[:, None] * w3[None, :], 0.0)
    g_b2 = dz2.sum(axis=0)
    dz2r = _r(dz2, emulate)
    g_W2 = f["a1"].T @ dz2r
    da1 = dz2r @ _r(W2, emulate).T
    dz1 = np.where(f["z1"] > 0.0, da1, 0.0)
    g_b1 = dz1.sum(axis=0)
    g_W1 = f["X"].T @ _r(dz1, emulate)
    return loss, pack(g_W1, g_b1, g_W2, g_b2, g_w3, g_b3)


def adam(p, X, t, w, steps, lr, beta1=0.9, beta2=0.999, eps=1e-8, emulate=True):
    """`steps` full-batch Adam steps from moments at zero; returns
    (parameters, loss before the last step)."""
    p = np.asarray(p, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    loss = None
    for s in range(1, steps + 1):
        loss, g = loss_grad(p, X, t, w, emulate)
        m = beta1 * m + (1.0 - beta1) * g
        v = beta2 * v + (1.0 - beta2) * g * g
        mh = m / (1.0 - beta1 ** s)
        vh = v / (1.0 - beta2 ** s)
        p = p - lr * mh / (np.sqrt(vh) + eps)
    return p, loss


def init_params(seed):
    """He-uniform test parameters (not the library's initialiser)."""
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (FEAT, HID)) * np.sqrt(6.0 / FEAT)
    W2 = rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6.0 / HID)
    w3 = rng.uniform(-1, 1, HID) * np.sqrt(6.0 / HID)
    b1 = rng.normal(0, 0.1, HID)
    b2 = rng.normal(0, 0.1, HID)
    return pack(W1, b1, W2, b2, w3, 0.3)


def tile_features(count, mesh, home, ext_over_cell_min, pos, fill, k):
    """The 160 inputs of one source tile (float64 before BF16 rounding):
    log2(1 + occupancy) / 8 over the home cell's 5x5x5 periodic
    neighbourhood in (dz, dy, dx) order, log2(box edge / cell_min + 1/256)
    / 8 per axis, the box centre's fractional position in the home cell per
    axis, fill / 16, log2(k) / 8, then zeros. count is indexed (z, y, x)."""
    nx, ny, nz = mesh
    hx, hy, hz = home
    out = np.zeros(FEAT)
    j = 0
    for dz in range(-2, 3):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                occ = count[(hz + dz) % nz, (hy + dy) % ny, (hx + dx) % nx]
                out[j] = np.log2(1.0 + occ) / 8.0
                j += 1
    out[125:128] = np.log2(np.asarray(ext_over_cell_min, dtype=np.float64) + 1.0 / 256.0) / 8.0
    out[128:131] = pos
    out[131] = fill / 16.0
    out[132] = np.log2(k) / 8.0
    return out
