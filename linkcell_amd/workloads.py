"""Seeded synthetic inputs for the benchmark configurations.

Each generator returns plain numpy arrays (host); callers move them to the
device. Configurations (sizes fixed, seeds fixed):

* ``uniform_c1``    N=4096 uniform in a 16^3 box, k=12, hint 3.0, drawn with
                    the reference test LCG (tests/agree.rs:144-147).
* ``triclinic_c2``  a=(40,0,0) b=(20,34.64,0) c=(8,6,38), N=262,144
                    uniform fractional coordinates, 5% shifted by +-1 or
                    +-2 images per axis, 5% masked, k=16.
* ``clustered_c3``  120x120x360 box, N=4,194,304: liquid slab |z-180|<60,
                    vapour elsewhere, 2000 Gaussian clusters (sigma 1.5),
                    membership in proportion 172,800 : 6,912 : 400,000
                    (the stated densities and cluster sizes, scaled to N),
                    k=16. ``clustered_c3_device`` draws further frames of the
                    same mixture on the device (training frames);
                    ``clustered_c3_small`` is the mixture at 1/16 scale.
* ``batched_c4``    256 frames, N log-uniform in [1e3, 2e5], even frames
                    orthorhombic (edges 15-80), odd frames triclinic (tilts
                    up to 0.4 of the edge), uniform or clustered, k in
                    {4, 12, 32}.
* ``small_c5i``     128 frames of N in [64, 512] in boxes of edge 4-8, k=100.
* ``fcc_c5ii``      jittered FCC, 4 x 128^3 sites at a=1.5625 (200^3 box),
                    jitter sigma 0.15, k=32.
"""
from __future__ import annotations

import math

import numpy as np

MASK64 = (1 << 64) - 1


class Lcg:
    """s = s * 6364136223846793005 + 1 (mod 2^64); u = (s >> 11) / 2^53
    (tests/agree.rs:144-147)."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def frac(self) -> float:
        self.s = (self.s * 6364136223846793005 + 1) & MASK64
        return (self.s >> 11) / float(1 << 53)


def lcg_points_box(n, lengths, seed):
    """n points (x, y, z) = (u*Lx, v*Ly, w*Lz), draws in x, y, z order
    (tests/agree.rs:61-73)."""
    g = Lcg(seed)
    out = np.empty((n, 3), dtype=np.float64)
    for i in range(n):
        out[i, 0] = g.frac() * lengths[0]
        out[i, 1] = g.frac() * lengths[1]
        out[i, 2] = g.frac() * lengths[2]
    return out


def lcg_points_in_cell(n, a, b, c, seed):
    """cell.cartesian([u, v, w]) per point (tests/agree.rs:149-159); the
    cartesian map is H s summed left to right (src/cell.rs:200-207)."""
    g = Lcg(seed)
    out = np.empty((n, 3), dtype=np.float64)
    for i in range(n):
        u, v, w = g.frac(), g.frac(), g.frac()
        for r in range(3):
            out[i, r] = a[r] * u + b[r] * v + c[r] * w + 0.0
    return out


def uniform_c1():
    xyz = lcg_points_box(4096, (16.0, 16.0, 16.0), 1)
    cell = np.array([16.0, 16.0, 16.0])
    return dict(name="uniform-16^3-N4096-k12", xyz=xyz, cell=cell, k=12,
                hint=3.0, mask=None)


C2_CELL = np.array([[40.0, 0.0, 0.0], [20.0, 34.64, 0.0], [8.0, 6.0, 38.0]])


def triclinic_c2(n=262144):
    rng = np.random.default_rng(2)
    s = rng.random((n, 3))
    out = rng.random(n) < 0.05
    shifts = rng.choice(np.array([-2, -1, 1, 2]), size=(n, 3))
    s[out] += shifts[out]
    xyz = s @ C2_CELL  # row i: s0*a + s1*b + s2*c
    mrng = np.random.default_rng(3)
    mask = (mrng.random(n) >= 0.05).astype(np.int32)
    return dict(name=f"triclinic-N{n}-k16", xyz=np.ascontiguousarray(xyz),
                cell=C2_CELL.copy(), k=16, hint=3.0, mask=mask)


C3_BOX = (120.0, 120.0, 360.0)


def clustered_c3(frame=0, n=4194304, n_clusters=2000, box=C3_BOX):
    """The slab / vapour / cluster mixture. Smaller n with a box of
    proportionally smaller volume and n_clusters keeps the same local
    densities (the training and test frames of the learned scheduler)."""
    rng = np.random.default_rng(1000 + frame)
    w = np.array([172800.0, 6912.0, 400000.0])
    counts = rng.multinomial(n, w / w.sum())
    n_slab, n_vap, n_clu = (int(c) for c in counts)
    lx, ly, lz = box
    third = lz / 3.0
    slab = np.column_stack([rng.random(n_slab) * lx, rng.random(n_slab) * ly,
                            third + rng.random(n_slab) * third])
    vz = rng.random(n_vap) * 2.0 * third
    vz = np.where(vz < third, vz, vz + third)
    vap = np.column_stack([rng.random(n_vap) * lx, rng.random(n_vap) * ly, vz])
    centres = rng.random((n_clusters, 3)) * np.array(box)
    member = rng.integers(0, n_clusters, size=n_clu)
    clu = centres[member] + rng.normal(0.0, 1.5, size=(n_clu, 3))
    xyz = np.concatenate([slab, vap, clu])
    perm = rng.permutation(n)
    xyz = np.ascontiguousarray(xyz[perm])
    bx = "x".join(f"{v:g}" for v in box)
    return dict(name=f"clustered-slab-{bx}-N{n}-k16", xyz=xyz,
                cell=np.array(box, dtype=np.float64), k=16, hint=3.0, mask=None)


def clustered_c3_device(frame, device, n=4194304, n_clusters=2000, box=C3_BOX):
    """The clustered mixture drawn on the device with torch's generator
    (seed 2000 + frame): the same distribution as clustered_c3, not the
    same points. Used for the learned scheduler's training frames, where
    32 frames of 4M points would take a minute with numpy."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(2000 + frame)
    w = np.array([172800.0, 6912.0, 400000.0])
    c = np.cumsum(w / w.sum())
    v = torch.rand(n, dtype=torch.float64, device=device, generator=g)
    comp = (v >= c[0]).to(torch.int64) + (v >= c[1]).to(torch.int64)
    lx, ly, lz = box
    third = lz / 3.0
    u = torch.rand((n, 3), dtype=torch.float64, device=device, generator=g)
    L = torch.tensor([lx, ly, lz], dtype=torch.float64, device=device)
    xyz = u * L
    slab = comp == 0
    xyz[:, 2] = torch.where(slab, third + u[:, 2] * third, xyz[:, 2])
    vz = u[:, 2] * 2.0 * third
    vz = torch.where(vz < third, vz, vz + third)
    xyz[:, 2] = torch.where(comp == 1, vz, xyz[:, 2])
    centres = torch.rand((n_clusters, 3), dtype=torch.float64, device=device, generator=g) * L
    member = torch.randint(0, n_clusters, (n,), device=device, generator=g)
    noise = torch.randn((n, 3), dtype=torch.float64, device=device, generator=g) * 1.5
    clu = comp == 2
    xyz = torch.where(clu[:, None], centres[member] + noise, xyz)
    return xyz.contiguous()


def clustered_c3_small(frame=0):
    """The clustered mixture at 1/16 of the points, the clusters and the
    volume (box 60 x 60 x 90: x and y halved, z quartered), so the slab,
    vapour and cluster densities are those of the full box."""
    return clustered_c3(frame, n=4194304 // 16, n_clusters=2000 // 16,
                        box=(60.0, 60.0, 90.0))


def _clustered_in(rng, n, H, n_clusters):
    s_c = rng.random((n_clusters, 3))
    centres = s_c @ H
    member = rng.integers(0, n_clusters, size=n)
    return centres[member] + rng.normal(0.0, 1.5, size=(n, 3))


def batched_c4(n_frames=256, seed0=4000):
    frames = []
    for f in range(n_frames):
        rng = np.random.default_rng(seed0 + f)
        n = int(math.floor(math.exp(rng.uniform(math.log(1e3), math.log(2e5)))))
        L = rng.uniform(15.0, 80.0, size=3)
        if f % 2 == 0:
            H = np.diag(L)
        else:
            t = rng.uniform(-0.4, 0.4, size=3)
            H = np.array([[L[0], 0.0, 0.0], [t[0] * L[0], L[1], 0.0],
                          [t[1] * L[0], t[2] * L[1], L[2]]])
        if rng.random() < 0.5:
            xyz = rng.random((n, 3)) @ H
        else:
            xyz = _clustered_in(rng, n, H, max(1, n // 200))
        k = int(rng.choice([4, 12, 32]))
        frames.append(dict(xyz=np.ascontiguousarray(xyz), cell=H, k=k, hint=3.0,
                           mask=None))
    return frames


def small_c5i(n_frames=128, seed0=5000):
    frames = []
    for f in range(n_frames):
        rng = np.random.default_rng(seed0 + f)
        n = int(rng.integers(64, 513))
        L = rng.uniform(4.0, 8.0, size=3)
        xyz = rng.random((n, 3)) * L
        frames.append(dict(xyz=np.ascontiguousarray(xyz), cell=L.copy(), k=100,
                           hint=3.0, mask=None))
    return frames


def fcc_c5ii(cells=128, a=1.5625, jitter=0.15, seed=5500):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(cells, dtype=np.float64)
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1)
    sites = (grid[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    xyz = sites + rng.normal(0.0, jitter, size=sites.shape)
    L = cells * a
    return dict(name=f"fcc-{4 * cells ** 3}-k32", xyz=np.ascontiguousarray(xyz),
                cell=np.array([L, L, L]), k=32, hint=3.0, mask=None)


def fcc_c5ii_device(seed, device, cells=128, a=1.5625, jitter=0.15):
    """The jittered FCC lattice of fcc_c5ii drawn on the device with torch's
    generator (same lattice, fresh jitter per seed): the learned scheduler's
    training frames for the 8M-point FCC workload."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    base = torch.tensor([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]],
                        dtype=torch.float64, device=device)
    r = torch.arange(cells, dtype=torch.float64, device=device)
    gx, gy, gz = torch.meshgrid(r, r, r, indexing="ij")
    grid = torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], dim=1)
    sites = (grid[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    noise = torch.randn(sites.shape, dtype=torch.float64, device=device, generator=g)
    return (sites + noise * jitter).contiguous()
