"""Workload generators: synth.rs:326-342 (torus volume), binary STL round
trip (io/tests.rs:14-45), and the workload sizes."""
import numpy as np

from argent_amd import synth


def mesh_volume(p, t):
    p = p.astype(np.float64)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)


def test_torus_volume():
    p, t = synth.torus_profile(256, 64, lambda u, v: (np.float32(10.0) + np.cos(v), np.sin(v)))
    v = abs(mesh_volume(p, t))
    assert abs(v - 2 * np.pi ** 2 * 10.0) / (2 * np.pi ** 2 * 10.0) < 0.01


def test_workload_sizes():
    p, t, n = synth.torus_ring(seed=1)
    assert len(t) == 131072 and len(p) == 65536
    r = np.linalg.norm(p[:, :2], axis=1)
    tube = np.sqrt((r - 9.0) ** 2 + p[:, 2] ** 2)
    assert tube.min() > 1.2 - 0.051 and tube.max() < 1.2 + 0.051
    p, t, n = synth.signet(seed=2)
    assert 1_150_000 < len(t) < 1_250_000
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)


def test_sphere_rays():
    o, d = synth.sphere_rays(20000, [1.0, 2.0, 3.0], seed=3)
    assert np.allclose(np.linalg.norm(o - [1.0, 2.0, 3.0], axis=1), 30.0, atol=1e-3)
    to_c = np.array([1.0, 2.0, 3.0]) - o
    to_c /= np.linalg.norm(to_c, axis=1, keepdims=True)
    ang = np.degrees(np.arccos(np.clip(np.sum(to_c * d, axis=1), -1, 1)))
    assert ang.max() <= 20.01 and ang.mean() > 5.0


def test_binary_stl_round_trip():
    p, t, _ = synth.torus_ring(seed=1, nu=32, nv=8)
    data = synth.to_binary_stl(p, t)
    assert data[:5] == b"solid" and len(data) == 84 + 50 * len(t)
    q, s = synth.parse_binary_stl(data)
    assert len(s) == len(t)
    assert np.array_equal(q.reshape(-1, 3, 3), p[t])


def test_hdri_shape():
    img = synth.hdri(256, 128, seed=4)
    assert img.shape == (128, 256, 3) and np.all(img > 0) and img.max() > 20 * np.median(img)


def test_brilliant_is_a_closed_convex_57_facet_solid():
    p, t = synth.brilliant(4.0, seed=3)
    a, b, c = (p[t[:, k]].astype(np.float64) for k in range(3))
    n = np.cross(b - a, c - a)
    # one plane per facet: group triangle normals
    un = n / np.linalg.norm(n, axis=1, keepdims=True)
    planes = []
    for k in range(len(un)):
        if not any(un[k] @ q[:3] > 0.99999 for q in planes):
            planes.append(np.r_[un[k], un[k] @ a[k]])
    assert len(planes) == 57
    centre = p.mean(axis=0)
    assert np.all(np.einsum("ij,ij->i", n, (a + b + c) / 3 - centre) > 0)  # outward
    # closed: every edge is shared by exactly two triangles (welded by position)
    key = {tuple(np.round(v, 5)): i for i, v in enumerate(p)}
    vid = np.array([key[tuple(np.round(v, 5))] for v in p])
    edges = {}
    for tri in vid[t]:
        for e in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            k = tuple(sorted(e))
            edges[k] = edges.get(k, 0) + 1
    assert all(v == 2 for v in edges.values())
    # convex: every vertex on or behind every facet plane
    for pl in planes:
        assert np.all(p @ pl[:3] <= pl[3] + 1e-3)
    # facets pass through the nominal girdle circle at their midlines; where
    # neighbours meet the outline lies a few per cent outside it
    girdle = 2 * np.max(np.linalg.norm(p[:, [0, 2]], axis=1))
    assert 4.0 <= girdle < 4.0 * 1.06


def test_signet_stones_do_not_overlap():
    pos, tris, nrm, mat, diam = synth.signet_stones(
        seed=2, signet_kw=dict(band_nu=64, band_nv=16, plate_nt=20, plate_np=40))
    assert len(diam) == 12 and set(np.unique(mat)) == {0, 1, 2}
    assert len(mat) == len(tris) and np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    stones = [tris[mat == m] for m in (1, 2)]
    centres = []
    for m in (1, 2):
        sel = tris[mat == m]
        for k in range(0, len(sel), 216):
            v = pos[sel[k:k + 216].ravel()]
            centres.append((v[:, [0, 2]].mean(axis=0), (v[:, 0].max() - v[:, 0].min()) / 2))
    for i in range(len(centres)):
        for j in range(i + 1, len(centres)):
            (ci, ri), (cj, rj) = centres[i], centres[j]
            assert np.linalg.norm(ci - cj) >= (ri + rj) * 0.99
