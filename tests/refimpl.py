"""Independent numpy checks for the oracle (test helpers only).

These do not call the oracle: they restate the reference tests' own
oracles (the 27-image MIC of tests/agree.rs:161-210) and the fold
(src/cell.rs:176-207) in vectorised numpy, whose elementwise float64
operations are single correctly rounded IEEE-754 ops (no FMA), so their
bits can be compared with the oracle's.
"""
from __future__ import annotations

import numpy as np


def rows_of(cell):
    c = np.asarray(cell, dtype=np.float64)
    if c.shape == (3,):
        H = np.diag(c)
        o = np.zeros(3)
    elif c.shape == (3, 3):
        H, o = c.copy(), np.zeros(3)
    elif c.shape == (4, 3):
        H, o = c[:3].copy(), c[3].copy()
    else:
        c = c.reshape(12)
        H, o = c[:9].reshape(3, 3).copy(), c[9:].copy()
    return H, o  # H rows are a, b, c


def shifts27(H):
    """All 27 shifts fa*a + fb*b + fc*c summed left to right
    (tests/agree.rs:173-175)."""
    out = []
    for na in (-1, 0, 1):
        for nb in (-1, 0, 1):
            for nc in (-1, 0, 1):
                fa, fb, fc = float(na), float(nb), float(nc)
                s = fa * H[0] + fb * H[1] + fc * H[2]
                out.append(s)
    return np.array(out)


def mic27_matrix(xyz, H):
    """d2[i, j] = min over 27 images of |q + s - p|^2, evaluated as
    (q0 + s0 - p0)^2 + (q1 + s1 - p1)^2 + (q2 + s2 - p2)^2."""
    x = np.asarray(xyz, dtype=np.float64)
    best = np.full((len(x), len(x)), np.inf)
    for s in shifts27(H):
        q = x + s  # q + shift first
        dx = q[None, :, 0] - x[:, None, 0]
        dy = q[None, :, 1] - x[:, None, 1]
        dz = q[None, :, 2] - x[:, None, 2]
        d2 = dx * dx + dy * dy
        d2 = d2 + dz * dz
        best = np.minimum(best, d2)
    return best


def brute_knn(xyz, cell, k, mask=None):
    """27-image brute force k nearest with (d2, index) order
    (tests/agree.rs:186-210). Returns packed (nn, d2) with -1 / NaN."""
    H, _ = rows_of(cell)
    x = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = len(x)
    act = np.ones(n, bool) if mask is None else np.asarray(mask) != 0
    D = mic27_matrix(x, H)
    nn = np.full((n, k), -1, dtype=np.int32)
    dd = np.full((n, k), np.nan)
    idx = np.arange(n)
    for i in range(n):
        if not act[i]:
            continue
        cand = idx[act & (idx != i)]
        order = np.lexsort((cand, D[i, cand]))
        sel = cand[order][:k]
        nn[i, :len(sel)] = sel
        dd[i, :len(sel)] = D[i, sel]
    return nn, dd


def fold(xyz, cell):
    """Cell::fractional then Cell::cartesian, vectorised in source order."""
    H, o = rows_of(cell)
    x = np.asarray(xyz, dtype=np.float64).reshape(-1, 3)
    a, b, c = H[0], H[1], H[2]
    det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) \
        + a[2] * (b[0] * c[1] - b[1] * c[0])
    bc = np.array([b[1] * c[2] - b[2] * c[1], b[2] * c[0] - b[0] * c[2],
                   b[0] * c[1] - b[1] * c[0]])
    ca = np.array([c[1] * a[2] - c[2] * a[1], c[2] * a[0] - c[0] * a[2],
                   c[0] * a[1] - c[1] * a[0]])
    ab = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                   a[0] * b[1] - a[1] * b[0]])

    def nrm(v):
        return np.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])

    w = np.array([abs(det) / nrm(bc), abs(det) / nrm(ca), abs(det) / nrm(ab)])
    scale = max(nrm(a) + nrm(b) + nrm(c), 1.0)
    tol = 1e-12 * scale
    ortho = all(abs(v) <= tol for v in (a[1], a[2], b[0], b[2], c[0], c[1]))
    if ortho:
        s = (x - o) / w
    else:
        inv = 1.0 / det
        # columns of Hinv (src/cell.rs:369-385), hinv[col][row]
        hinv = np.array([
            [(b[1] * c[2] - b[2] * c[1]) * inv, (a[2] * c[1] - a[1] * c[2]) * inv,
             (a[1] * b[2] - a[2] * b[1]) * inv],
            [(b[2] * c[0] - b[0] * c[2]) * inv, (a[0] * c[2] - a[2] * c[0]) * inv,
             (a[2] * b[0] - a[0] * b[2]) * inv],
            [(b[0] * c[1] - b[1] * c[0]) * inv, (a[1] * c[0] - a[0] * c[1]) * inv,
             (a[0] * b[1] - a[1] * b[0]) * inv]])
        d = x - o
        s = np.empty_like(d)
        for r in range(3):
            s[:, r] = hinv[0][r] * d[:, 0] + hinv[1][r] * d[:, 1] + hinv[2][r] * d[:, 2]
    s = s - np.floor(s)
    s = np.where(s >= 1.0, 0.0, s)
    f = np.empty_like(s)
    for r in range(3):
        f[:, r] = (a[r] * s[:, 0] + b[r] * s[:, 1] + c[r] * s[:, 2]) + o[r]
    return s, f, w, ortho
