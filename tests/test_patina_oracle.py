"""The oracle's patina mask (argent-mesh cavity.rs:52-296 restated in
oracle/cavity.cpp) pinned by the reference's own cavity tests
(cavity/tests.rs) and by what the geometry fixes: the concavity of spheres and
cylinders from either side (radius x mean curvature), zero on a plane, the
thin-wall rule, and the smoothing step's closed form on a regular grid."""
import numpy as np
import pytest

import oracle as O
from argent_amd import synth

from patinacases import cylinder, groove_ridge_profile, height_field, plate, value_near


@pytest.fixture(scope="module")
def groove():
    pos, tris = height_field(200, groove_ridge_profile)
    p, n, t, _ = O.normals(pos, tris, 60.0)  # crate::normals::compute(&mut m, 60.0)
    return p, n, t


def test_groove_dark_ridge_bright(groove):
    """cavity/tests.rs:48-61 with PatinaSettings::default()."""
    pos, nrm, tris = groove
    mask = O.patina_mask(pos, nrm, tris)
    assert mask.shape == (len(pos),)
    assert value_near(pos, mask, -3.0, 0.0) > 0.6
    assert value_near(pos, mask, 3.0, 0.0) < 0.05
    assert value_near(pos, mask, 0.0, 0.0) < 0.1
    assert ((mask >= 0.0) & (mask <= 1.0)).all()


def test_concavity_signs(groove):
    """cavity/tests.rs:63-76."""
    pos, nrm, tris = groove
    c = O.concavity(pos, nrm, tris, 0.4)
    assert value_near(pos, c, -3.0, 0.0) > 0.2
    assert value_near(pos, c, 3.0, 0.0) < -0.2
    assert abs(value_near(pos, c, 0.0, 0.0)) < 1e-3
    sp, st = synth.icosphere(4.0, 5)
    sp, sn, st, _ = O.normals(sp, st, 50.0)
    cs = O.concavity(sp, sn, st, 0.4)
    assert abs(cs.mean() + 0.1) < 0.02, cs.mean()


@pytest.mark.parametrize("inward", [False, True])
def test_sphere_and_cylinder_closed_forms(inward):
    """On a sphere of radius R every vertex reads about radius / R times the
    sign of the side (outside convex, inside concave); on a cylinder, about
    radius / (2R), the mean of its principal curvatures 1/R and 0.  The
    estimate works on cell centroids (chords, a little inside the surface),
    which reads about 10 % more convex here: the bound is 15 % on the mean and
    40 % per vertex.  A dropped factor of 2 or a flipped sign fails these."""
    R, r = 4.0, 0.4
    sign = 1.0 if inward else -1.0
    sp, st = synth.icosphere(R, 5)
    sn = sp / np.linalg.norm(sp, axis=1, keepdims=True)
    if inward:
        sn, st = -sn, st[:, ::-1].copy()
    cp, ct, cn = cylinder(R, 8.0, 512, 256, inward=inward)
    mid = np.abs(cp[:, 2]) < 2.0  # away from the open rims
    for c, want in ((O.concavity(sp, sn, st, r), sign * r / R), (O.concavity(cp, cn, ct, r)[mid], sign * r / (2 * R))):
        assert abs(c.mean() / want - 1.0) < 0.15, (c.mean(), want)
        assert (np.abs(c / want - 1.0) < 0.4).all(), (c.min(), c.max(), want)


def test_plane_is_flat_and_thin_walls_ignore_their_back():
    """A plane reads 0 exactly.  On a thin plate (faces 0.5 apart, normals
    opposed, radius 0.8, so the faces fall in different cells of size 0.4 but
    within each other's radius) each side ignores the other (cell normals
    facing away, dot < -0.5): still 0; without that rule the far face would
    read as a strong bump."""
    pos, tris = height_field(40, lambda x: np.zeros_like(x))
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    assert (O.concavity(pos, nrm, tris, 0.8) == 0.0).all()
    pp, pt, pn = plate(40, 0.5)
    c = O.concavity(pp, pn, pt, 0.8)
    assert (c == 0.0).all(), np.abs(c).max()


def test_concavity_degenerate_inputs():
    """cavity.rs:131-139 and :222-226: no radius, zero normals."""
    pos, tris = height_field(10, lambda x: 0.1 * x * x)
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    assert (O.concavity(pos, nrm, tris, 0.0) == 0.0).all()
    assert (O.concavity(pos, nrm, tris, float("nan")) == 0.0).all()
    nz = nrm.copy()
    nz[5] = 0.0
    c = O.concavity(pos, nz, tris, 2.0)
    assert c[5] == 0.0 and c[6] > 0.0


def test_smoothing_preserves_constants_and_merges_crease_copies():
    """cavity/tests.rs:101-119."""
    p, t = synth.subdivided_cuboid([2.0, 2.0, 2.0], 4)
    p, n, t, _ = O.normals(p, t, 50.0)
    v = O.smooth(p, t, np.full(len(p), 0.7, np.float32), 5, 0.5)
    assert (np.abs(v - 0.7) < 1e-6).all()
    spike = np.zeros(len(p), np.float32)
    spike[0] = 1.0
    v = O.smooth(p, t, spike, 2, 0.5)
    key = (p + np.float32(0)).view(np.uint32)
    for i in range(len(p)):
        same = (key == key[i]).all(1)
        assert (v[same] == v[i]).all()
    assert v[0] < 1.0 and (v > 0.0).sum() > 3


def test_smoothing_one_step_closed_form():
    """One step with lambda = 0.5 of a unit spike at an interior vertex of the
    diagonal-split grid (6 neighbours each): the spike keeps 1/2, each
    neighbour j gets 0.5 / deg(j) = 1/12, everything else stays 0.  Two
    copies of the spike vertex (an unwelded duplicate) average first."""
    pos, tris = height_field(6, lambda x: np.zeros_like(x))
    s = 3 * 7 + 3
    v = np.zeros(len(pos), np.float32)
    v[s] = 1.0
    out = O.smooth(pos, tris, v, 1, 0.5)
    nb = set(tris[(tris == s).any(1)].ravel()) - {s}
    assert len(nb) == 6
    assert out[s] == np.float32(0.5)
    for j in range(len(pos)):
        if j in nb:
            assert abs(out[j] - 1.0 / 12.0) < 1e-7
        elif j != s:
            assert out[j] == 0.0
    # a duplicate of the spike vertex holding 0: the class starts at 1/2
    pos2 = np.concatenate([pos, pos[s:s + 1]])
    tris2 = np.concatenate([tris, [[s, len(pos), s]]]).astype(np.uint32)
    v2 = np.concatenate([v, [0.0]]).astype(np.float32)
    out2 = O.smooth(pos2, tris2, v2, 0, 0.5)
    assert out2[s] == out2[-1] == np.float32(0.5)
    # one step: the degenerate triangle gives the class a self-edge, which
    # counts in the divisor (7 entries) but not in the sum: 0.5 - 0.25 = 0.25
    out2 = O.smooth(pos2, tris2, v2, 1, 0.5)
    assert out2[s] == out2[-1] == np.float32(0.25)
    assert all(abs(out2[j] - 1.0 / 24.0) < 1e-7 for j in nb)


def test_patina_open_plane_and_enclosure():
    """An open plane has no AO and no concavity: mask 0 everywhere
    (cavity/tests.rs:121-127 for the AO term).  A vertex under a lid reads
    AO 1 and, weighted to AO alone, mask 1."""
    pos, tris = height_field(20, lambda x: np.zeros_like(x))
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    assert (O.patina_mask(pos, nrm, tris) == 0.0).all()
    m = O.patina_mask(pos, nrm, tris, concavity_weight=0.0, smooth_iterations=0, ao=np.ones(len(pos), np.float32))
    assert (m == 1.0).all()
    assert O.patina_mask(pos, nrm, np.zeros((0, 3), np.uint32)).sum() == 0.0


def test_smooth_vertex_normals_closed_forms():
    """normals.rs:226-242, the fallback of cavity.rs:82-87: a triangulated
    cube's corners read (+-1, +-1, +-1)/sqrt(3) (each face's corner angles sum
    to 90 degrees whatever its diagonal), a flat grid reads +z exactly, an
    unreferenced vertex +z, and a fine icosphere's normals are radial."""
    p, t = synth.cuboid([2.0, 2.0, 2.0])
    n = O.smooth_vertex_normals(p, t)
    want = np.sign(p) / np.sqrt(3.0)
    assert np.abs(n - want).max() < 1e-6
    g, gt = height_field(8, lambda x: np.zeros_like(x))
    g = np.concatenate([g, [[9.0, 9.0, 9.0]]]).astype(np.float32)
    ng = O.smooth_vertex_normals(g, gt)
    assert (ng == np.float32([0, 0, 1])).all()
    s, st = synth.icosphere(3.0, 4)
    ns = O.smooth_vertex_normals(s, st)
    radial = s / np.linalg.norm(s, axis=1, keepdims=True)
    assert (np.sum(ns * radial, 1) > 0.9995).all()


def test_patina_without_normals_uses_smooth_normals():
    """patina_mask on a mesh without normals equals the mask given its smooth
    vertex normals (cavity.rs:82-87)."""
    pos, tris = height_field(60, groove_ridge_profile)
    a = O.patina_mask(pos, None, tris)
    b = O.patina_mask(pos, O.smooth_vertex_normals(pos, tris), tris)
    assert np.array_equal(a, b)
    assert value_near(pos, a, -3.0, 0.0) > 0.5
