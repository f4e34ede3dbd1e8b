"""Shared mesh-ingest cases for the oracle tests and the GPU parity tests:
seeded meshes, STL byte strings and a brute-force weld.

The brute-force weld states weld.rs:1-13 directly: vertices with
bit-identical coordinates (-0 == +0) form one class; classes whose first
positions lie within eps (f64 distance) are joined transitively; output
vertices are the joined groups in the order of their first input vertex, at
that vertex's position; triangles are remapped and those with a repeated
index dropped.  It tests all pairs, so it checks the oracle's grid search and
union-find without sharing their structure.
"""
import struct

import numpy as np

from argent_amd import synth


def weld_eps(pos, tris, rel):
    """weld.rs:40-43: rel x bbox diagonal of referenced vertices, in f32."""
    used = np.zeros(len(pos), bool)
    used[np.asarray(tris).reshape(-1)] = True
    if not used.any():
        return np.float32(0.0)
    p = np.asarray(pos, np.float32)[used]
    d = p.max(0) - p.min(0)
    diag = np.sqrt(np.float32(np.float32(d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]))
    return np.float32(np.float64(max(np.float32(rel), np.float32(0.0))) * np.float64(diag))


def brute_weld(pos, tris, rel):
    pos = np.asarray(pos, np.float32)
    tris = np.asarray(tris, np.uint32).reshape(-1, 3)
    eps = weld_eps(pos, tris, rel)
    used = np.zeros(len(pos), bool)
    used[tris.reshape(-1)] = True
    canon = (pos + np.float32(0.0)).view(np.uint32)
    first = {}
    cls = np.full(len(pos), -1, np.int64)
    reps = []
    for i in np.nonzero(used)[0]:
        k = tuple(canon[i])
        if k not in first:
            first[k] = len(reps)
            reps.append(i)
        cls[i] = first[k]
    m = len(reps)
    parent = list(range(m))

    def find(x):
        while parent[x] != x:
            x = parent[x]
        return x

    rp = pos[reps].astype(np.float64)
    e2 = np.float64(eps) * np.float64(eps)
    if eps > 0:
        for a in range(m):
            d = rp[a] - rp[:a]
            close = np.nonzero(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2] <= e2)[0]
            for b in close:
                ra, rb = find(a), find(int(b))
                if ra != rb:
                    parent[max(ra, rb)] = min(ra, rb)
    roots = [find(a) for a in range(m)]
    order = {}
    out = []
    for a in range(m):
        if roots[a] not in order:
            order[roots[a]] = len(out)
            out.append(pos[reps[a]])
    remap = np.array([order[roots[c]] if c >= 0 else -1 for c in cls], np.int64)
    t = remap[tris.astype(np.int64)]
    keep = (t[:, 0] != t[:, 1]) & (t[:, 1] != t[:, 2]) & (t[:, 0] != t[:, 2])
    return np.array(out, np.float32).reshape(-1, 3), t[keep].astype(np.uint32), {
        "exact_merges": int(used.sum()) - m, "tolerance_merges": m - len(out),
        "dropped_triangles": int((~keep).sum()), "tolerance_mm": float(eps)}


def clustered(seed, n_centres=40, per=(1, 6), rel=1e-3):
    """Points in tight clusters around random centres, with exact duplicates,
    signed zeros, chains of near points and unreferenced vertices; offsets
    span 0.2-1.6 eps so pairs fall on both sides of eps and across cell walls."""
    rng = np.random.default_rng(seed)
    centres = rng.uniform(0.0, 10.0, (n_centres, 3)).astype(np.float32)
    centres[::7, 0] = 0.0
    eps_guess = rel * 10.0 * np.sqrt(3.0)
    pts = []
    for c in centres:
        k = rng.integers(per[0], per[1] + 1)
        p = c.copy()
        for _ in range(k):
            pts.append(p.copy())
            step = rng.normal(size=3)
            step *= rng.uniform(0.2, 1.6) * eps_guess / np.linalg.norm(step)
            p = (p + step).astype(np.float32)
    pts = np.array(pts, np.float32)
    dup = rng.integers(0, len(pts), len(pts) // 3)
    pts = np.concatenate([pts, pts[dup]])
    z = pts[:, 0] == 0.0
    pts[np.nonzero(z)[0][::2], 0] = np.float32(-0.0)
    pts = pts[rng.permutation(len(pts))]
    nt = len(pts)
    tris = rng.integers(0, len(pts) - 3, (nt, 3)).astype(np.uint32)  # last 3 vertices unreferenced
    return pts, tris


def stl_bytes(pos, tris, header=b"solid argent"):
    data = bytearray(synth.to_binary_stl(pos, tris))
    data[:80] = header.ljust(80, b" ")[:80]
    return bytes(data)


def stl_with_bad(seed, n=500):
    """A soup STL where some records carry NaN or infinite coordinates."""
    rng = np.random.default_rng(seed)
    pos, tris = synth.torus(10.0, 2.0, 20, 12)
    data = bytearray(stl_bytes(pos, tris))
    nt = len(tris)
    for t in rng.choice(nt, 17, replace=False):
        off = 84 + 50 * int(t) + 12 + 4 * int(rng.integers(0, 9))
        struct.pack_into("<f", data, off, [np.nan, np.inf, -np.inf][int(rng.integers(0, 3))])
    return bytes(data)


def signed_volume(pos, tris):
    return synth.signed_volume(pos, tris)
