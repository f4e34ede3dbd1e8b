"""The oracle's cavity ambient occlusion (argent-mesh cavity.rs:95-128),
pinned by the reference's own known answer (an open plane is unoccluded,
cavity/tests.rs:131-137), a closed form (a floor under a ceiling: the count is
exact per vertex from its stratification), the groove/ridge height field of
cavity/tests.rs:28-35, and brute-force occlusion of the same rays."""
import numpy as np
import pytest

import oracle as O

from aocases import ceiling_counts, desc, floor_and_ceiling, grid, groove_ridge, nearest


def test_open_plane_is_unoccluded():
    pos, tris, nrm = grid(20, -5.0, 5.0, 0.0)
    ao = O.OracleScene(desc(pos, tris, nrm)).ambient_occlusion(len(pos), 32, 1.0)
    assert np.all(ao == 0.0)


@pytest.mark.parametrize("samples,distance,h", [(48, 1.2, 0.5), (32, 2.0, 0.9), (7, 1.0, 0.25)])
def test_floor_under_ceiling_matches_closed_form(samples, distance, h):
    pos, tris, nrm = floor_and_ceiling(h)
    n_floor = 41 * 41
    ao = O.OracleScene(desc(pos, tris, nrm)).ambient_occlusion(len(pos), samples, distance)
    counts, clear = ceiling_counts(pos, n_floor, samples, distance, h)
    got = np.rint(ao[:n_floor] * samples).astype(int)
    assert clear.mean() > 0.9
    assert np.array_equal(got[clear], counts[clear])
    assert 0 < counts.mean() < samples


def test_groove_occluded_flat_and_ridge_open():
    pos, tris, nrm = groove_ridge(200)
    ao = O.OracleScene(desc(pos, tris, nrm)).ambient_occlusion(len(pos), 48, 1.2)
    assert ao[nearest(pos, 0.0, 0.0)] == 0.0
    assert ao[nearest(pos, 3.0, 0.0)] == 0.0
    assert ao[nearest(pos, -3.0, 0.0)] > 0.3
    assert np.all((ao >= 0.0) & (ao <= 1.0))


def test_no_normals_uses_smooth_normals_and_zero_normal_gives_zero():
    """cavity.rs:82-87: without the mesh's normals, smooth_vertex_normals."""
    pos, tris, nrm = floor_and_ceiling(0.5)
    a = O.OracleScene(desc(pos, tris, None)).ambient_occlusion(len(pos), 8, 1.0)
    b = O.OracleScene(desc(pos, tris, O.smooth_vertex_normals(pos, tris))).ambient_occlusion(len(pos), 8, 1.0)
    assert np.array_equal(a, b) and a.max() > 0.0
    nrm = nrm.copy()
    nrm[0] = 0.0
    ao = O.OracleScene(desc(pos, tris, nrm)).ambient_occlusion(len(pos), 8, 1.0)
    assert ao[0] == 0.0 and ao[1] > 0.0


def test_bvh_occlusion_matches_brute_force_on_the_ao_rays():
    """The oracle's any-hit BVH query against brute force over every triangle
    (O.brute_intersect), on rays built like the AO rays of the groove mesh."""
    pos, tris, nrm = groove_ridge(60)
    sc = O.OracleScene(desc(pos, tris, nrm))
    rng = np.random.default_rng(4)
    v = rng.integers(0, len(pos), 4000)
    d = nrm[v] + rng.normal(0, 0.6, (len(v), 3)).astype(np.float32)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    o = (pos[v] + nrm[v] * np.float32(2e-3)).astype(np.float32)
    tmax = np.full(len(v), 1.2, np.float32)
    occ = sc.occluded(o, d, tmax)
    bt, _, _, _ = O.brute_intersect(pos, tris, o, d, tmax)
    assert np.array_equal(occ, np.isfinite(bt) & (bt < 1.2))
    assert 0.02 < occ.mean() < 0.98
