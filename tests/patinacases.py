"""Meshes for the patina-mask tests (argent-mesh cavity.rs and its tests,
cavity/tests.rs).  Shared by the oracle tests (CPU) and the GPU parity tests;
nothing here calls the oracle or the library."""
import numpy as np


def height_field(n, f):
    """cavity/tests.rs:5-26: z = f(x) over [-5, 5]^2 with n x n quads, in f32
    as the reference builds it (x = -5 + 10 i / n).  Positions, triangles."""
    i = np.arange(n + 1, dtype=np.float32)
    xs = (np.float32(-5.0) + np.float32(10.0) * i / np.float32(n)).astype(np.float32)
    X, Y = np.meshgrid(xs, xs)
    Z = f(X).astype(np.float32)
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    j, k = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (j * (n + 1) + k).ravel()
    b, c, d = a + 1, a + n + 2, a + n + 1
    tris = np.stack([np.stack([a, b, c], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3)
    return pos, tris.astype(np.uint32)


def groove_ridge_profile(x):
    """cavity/tests.rs:28-35: V-groove at x = -3, ridge at x = +3 (0.5 deep/high, 1 wide), in f32."""
    x = np.asarray(x, np.float32)
    one, half = np.float32(1.0), np.float32(0.5)
    g = -half * np.maximum(one - np.abs(x + np.float32(3.0)) / half, np.float32(0.0))
    r = half * np.maximum(one - np.abs(x - np.float32(3.0)) / half, np.float32(0.0))
    return (g + r).astype(np.float32)


def value_near(pos, vals, x, y):
    """cavity/tests.rs:37-46: the value at the vertex nearest (x, y) in the
    xy plane (first of equals)."""
    d = (pos[:, 0] - np.float32(x)) ** 2 + (pos[:, 1] - np.float32(y)) ** 2
    return vals[int(np.argmin(d))]


def cylinder(radius, length, nu, nv, inward=False):
    """A closed-loop tube about the z axis: nu segments around, nv along, with
    exact outward normals (inward, and the winding reversed, with `inward`)."""
    th = np.arange(nu) * (2.0 * np.pi / nu)
    zs = np.linspace(-length / 2, length / 2, nv + 1)
    T, Z = np.meshgrid(th, zs)
    pos = np.stack([radius * np.cos(T).ravel(), radius * np.sin(T).ravel(), Z.ravel()], 1).astype(np.float32)
    nrm = np.stack([np.cos(T).ravel(), np.sin(T).ravel(), np.zeros(T.size)], 1).astype(np.float32)
    tris = []
    for j in range(nv):
        for i in range(nu):
            a, b = j * nu + i, j * nu + (i + 1) % nu
            c, d = b + nu, a + nu
            tris += [[a, b, c], [a, c, d]]
    tris = np.asarray(tris, np.uint32)
    if inward:
        nrm = -nrm
        tris = tris[:, ::-1].copy()
    return pos, tris, nrm


def plate(n, gap):
    """Two parallel n x n grids over [-5, 5]^2, at z = 0 facing down and at
    z = gap facing up: a thin wall seen from both sides.  Positions, triangles,
    normals."""
    p0, t0 = height_field(n, lambda x: np.zeros_like(x))
    p1 = p0.copy()
    p1[:, 2] = np.float32(gap)
    n0 = np.tile(np.float32([0, 0, -1]), (len(p0), 1))
    n1 = np.tile(np.float32([0, 0, 1]), (len(p1), 1))
    pos = np.concatenate([p0, p1])
    tris = np.concatenate([t0[:, ::-1], t0 + len(p0)]).astype(np.uint32)
    return pos, tris, np.concatenate([n0, n1])
