"""The oracle's crease-aware normals (normals.rs restated) pinned by the
reference's own normals tests (normals.rs:795-931) and closed forms: radial
normals on a sphere, flat faces on a creased cube with the exact split count,
one vertex per corner when faceted, inside-out shells flipped, a cavity kept
inward, inconsistent winding repaired, a broken fan on a plane merged."""
import numpy as np

import oracle as O
from argent_amd import synth
from normalcases import broken_fan, hollow, inconsistent_sphere, inside_out_sphere, welded_sphere


def _fn(p, t):
    v = p[t].astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def test_sphere_normals_are_radial():
    p, t = welded_sphere()
    q, n, s, r = O.normals(p, t, 50.0)
    assert r["split_vertices"] == 0 and r["boundary_edges"] == 0, r
    radial = q / np.linalg.norm(q, axis=1, keepdims=True)
    assert (np.einsum("ij,ij->i", radial, n) > 0.9995).all()


def test_cube_creases_keep_faces_flat():
    p, t = synth.subdivided_cuboid([2.0, 2.0, 2.0], 3)
    q, n, s, r = O.normals(p, t, 50.0)
    assert r["split_vertices"] == 8 * 2 + 12 * 2, r
    fn = _fn(q, s)
    for k in range(3):
        assert (np.einsum("ij,ij->i", fn, n[s[:, k]]) > 0.99999).all()
    assert abs(synth.signed_volume(q, s) - 8.0) < 1e-4


def test_faceted_mode_with_zero_crease():
    p, t = synth.icosphere(1.0, 1)
    q, n, s, r = O.normals(p, t, 0.0)
    assert len(q) == 3 * len(t)


def test_inside_out_shell_is_flipped():
    q, n, s, r = O.normals(*inside_out_sphere(), 50.0)
    assert r["flipped_shells"] == 1, r
    assert (np.einsum("ij,ij->i", q, n) > 0).all()
    assert synth.signed_volume(q, s) > 0


def test_hollow_cavity_is_kept_inward():
    q, n, s, r = O.normals(*hollow(), 50.0)
    assert r["flipped_shells"] == 0 and r["cavity_shells"] == 1, r
    v = 4.0 / 3.0 * np.pi * (125.0 - 27.0)
    assert abs(synth.signed_volume(q, s) - v) / v < 0.03


def test_inconsistent_faces_are_repaired():
    q, n, s, r = O.normals(*inconsistent_sphere(), 50.0)
    assert r["repaired_faces"] > 0 and r["inconsistent_edges"] == 0 and r["split_vertices"] == 0, r
    radial = q / np.linalg.norm(q, axis=1, keepdims=True)
    assert (np.einsum("ij,ij->i", radial, n) > 0.999).all()
    assert synth.signed_volume(q, s) > 0


def test_broken_fan_on_flat_surface_has_no_seam():
    q, n, s, r = O.normals(*broken_fan(), 50.0)
    assert r["split_vertices"] == 0, r
    assert (n[:, 2] > 0.99999).all()


def test_split_copies_share_positions_and_order():
    # split copies follow the originals in vertex order, at the same positions
    p, t = synth.subdivided_cuboid([1.0, 2.0, 3.0], 2)
    q, n, s, r = O.normals(p, t, 50.0)
    assert np.array_equal(q[:len(p)], p)
    extra = q[len(p):]
    assert len(extra) == r["split_vertices"]
    # every extra position is one of the corner / edge positions, repeated as often as its fans
    keys = {tuple(x) for x in p.tolist()}
    assert all(tuple(x) in keys for x in extra.tolist())
