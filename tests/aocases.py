"""Meshes and expected values for the cavity ambient-occlusion tests
(argent-mesh cavity.rs:95-128).  Shared by the oracle tests (CPU) and the GPU
parity tests; nothing here calls the oracle or the library."""
import numpy as np

import argent_amd as ag


def hash32(x):
    """argent-mesh util.rs:244-251, on uint32 arrays."""
    x = np.asarray(x, np.uint32).copy()
    x ^= x >> np.uint32(16)
    x = (x * np.uint32(0x7FEB352D)).astype(np.uint32)
    x ^= x >> np.uint32(15)
    x = (x * np.uint32(0x846CA68B)).astype(np.uint32)
    x ^= x >> np.uint32(16)
    return x


def grid(n, lo, hi, z, flip=False):
    """An n x n quad grid over [lo, hi]^2 at height z (normals +z, or -z with
    `flip`).  Returns positions, triangles, normals."""
    xs = np.linspace(lo, hi, n + 1, dtype=np.float32)
    X, Y = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z, np.float32)], 1).astype(np.float32)
    idx = lambda i, j: j * (n + 1) + i
    tris = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris += [[a, c, b], [a, d, c]] if flip else [[a, b, c], [a, c, d]]
    nrm = np.zeros_like(pos)
    nrm[:, 2] = -1.0 if flip else 1.0
    return pos, np.asarray(tris, np.uint32), nrm


def merge(*parts):
    pos, tris, nrm, off = [], [], [], 0
    for p, t, n in parts:
        pos.append(p)
        tris.append(t + off)
        nrm.append(n)
        off += len(p)
    return np.concatenate(pos), np.concatenate(tris), np.concatenate(nrm)


def desc(pos, tris, nrm):
    d = ag.SceneDesc(pos, tris, normals=nrm)
    d.add_material(ag.Material.metal("sterling"))
    return d


def floor_and_ceiling(h=0.5):
    """A 41 x 41-vertex floor over [-5, 5]^2 at z = 0 (normals up) under a
    ceiling over [-40, 40]^2 at z = h facing down."""
    return merge(grid(40, -5.0, 5.0, 0.0), grid(2, -40.0, 40.0, h, flip=True))


def ceiling_counts(pos, n_floor, samples, distance, h):
    """Closed form for floor vertex v: its ray k leaves at height eps, and the
    ceiling (wide enough for any ray of length `distance`) blocks it iff
    (h - eps) / cos(theta) < distance with cos(theta) = sqrt(1 - u1),
    u1 = min((k + jit_v) / n, 0.999999): iff u1 < 1 - ((h - eps) / distance)^2.
    Returns (count per floor vertex, mask of vertices whose threshold is at
    least 1e-3 of a stratum away from some (k + jit) / n)."""
    lo, hi = pos.min(0), pos.max(0)
    e = (hi - lo).astype(np.float32)
    diag = np.sqrt(np.float32(e[0] * e[0] + e[1] * e[1]) + np.float32(e[2] * e[2]))
    eps = max(np.float32(2e-4) * diag, np.float32(1e-3))
    thr = 1.0 - ((h - float(eps)) / distance) ** 2
    v = np.arange(n_floor, dtype=np.uint32)
    jit = (hash32(hash32(v ^ np.uint32(0x5BD1E995))) >> np.uint32(8)).astype(np.float64) / 16777216.0
    k = np.arange(samples)[None, :]
    u1 = np.minimum((k + jit[:, None]) / samples, 0.999999)
    counts = (u1 < thr).sum(1)
    clear = np.abs(u1 - thr).min(1) > 1e-3 / samples
    return counts, clear


def groove_ridge(n=200):
    """cavity/tests.rs:28-35: a V-groove at x = -3 and a ridge at x = +3, 0.5
    deep/high and 1 wide, on a height field z = f(x) over [-5, 5]^2; normals
    from the surface slope (exact for this profile away from the kinks)."""
    xs = np.linspace(-5.0, 5.0, n + 1)

    def f(x):
        return -0.5 * np.maximum(1.0 - np.abs(x + 3.0) / 0.5, 0.0) + 0.5 * np.maximum(1.0 - np.abs(x - 3.0) / 0.5, 0.0)

    X, Y = np.meshgrid(xs, xs)
    pos = np.stack([X.ravel(), Y.ravel(), f(X.ravel())], 1).astype(np.float32)
    h = 1e-3
    dz = (f(X.ravel() + h) - f(X.ravel() - h)) / (2 * h)
    nrm = np.stack([-dz, np.zeros_like(dz), np.ones_like(dz)], 1)
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    idx = lambda i, j: j * (n + 1) + i
    tris = []
    for j in range(n):
        for i in range(n):
            tris += [[idx(i, j), idx(i + 1, j), idx(i + 1, j + 1)], [idx(i, j), idx(i + 1, j + 1), idx(i, j + 1)]]
    return pos, np.asarray(tris, np.uint32), nrm


def nearest(pos, x, y):
    return int(np.argmin((pos[:, 0] - x) ** 2 + (pos[:, 1] - y) ** 2))
