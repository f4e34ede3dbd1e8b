"""The patina mask on the GPU (ag_patina_mask, ag_smooth_values) against the
oracle (argent-mesh cavity.rs:52-296 restated in oracle/cavity.cpp) and the
reference's own cavity tests.

Tolerances: the smoothing and the blend sum in the reference's order in f32
and match bit for bit.  The concavity sums in the same order in f64; its
Gaussian weights use the device's exp, which may differ from the host's in
the last bit of a double, so a vertex's value may move by an ulp or two of
the f32 result: the bound is 2e-6 absolute (values lie in [-4, 4]).  The AO
counts follow ag_ambient_occlusion's bound (at most 2 rays on 0.5 % of
vertices); the whole mask against the oracle's whole mask is therefore
checked on its mean and its bulk, and exactly once the GPU's AO is given to
the oracle."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth, workloads

from aocases import desc
from patinacases import groove_ridge_profile, height_field, value_near

pytestmark = pytest.mark.gpu

CAV_TOL = 2e-6


def _scene(ctx, pos, tris, nrm):
    return ag.Scene(ctx, desc(pos, tris, nrm))


@pytest.fixture(scope="module")
def groove():
    pos, tris = height_field(200, groove_ridge_profile)
    p, n, t, _ = O.normals(pos, tris, 60.0)
    return p, n, t


def test_groove_matches_reference_test_and_oracle(ctx, groove):
    """cavity/tests.rs:48-61 on the GPU, then every step against the oracle."""
    pos, nrm, tris = groove
    sc = _scene(ctx, pos, tris, nrm)
    try:
        mask, ao, cav = sc.patina_mask(parts=True)
    finally:
        sc.close()
    print(f"groove: {len(pos)} vertices, {len(tris)} triangles", flush=True)
    assert value_near(pos, mask, -3.0, 0.0) > 0.6
    assert value_near(pos, mask, 3.0, 0.0) < 0.05
    assert value_near(pos, mask, 0.0, 0.0) < 0.1
    assert ((mask >= 0.0) & (mask <= 1.0)).all()
    ref_cav = O.concavity(pos, nrm, tris, 0.4)
    assert np.abs(cav - ref_cav).max() <= CAV_TOL, np.abs(cav - ref_cav).max()
    assert (cav == ref_cav).mean() > 0.99
    # the GPU's AO through the oracle's blend and smoothing: equal to the tolerance of the concavity
    ref = O.patina_mask(pos, nrm, tris, ao=ao)
    assert np.abs(mask - ref).max() <= 1e-5, np.abs(mask - ref).max()
    # the whole oracle mask (its own AO)
    full = O.patina_mask(pos, nrm, tris)
    assert abs(float(mask.mean() - full.mean())) < 1e-3
    assert (np.abs(mask - full) < 1e-3).mean() > 0.99


@pytest.mark.parametrize("kw", [dict(), dict(ao_samples=16, ao_distance=2.5, curvature_radius=0.3, concavity_weight=0.6,
                                             smooth_iterations=5)])
def test_signet_matches_oracle(ctx, kw):
    pos, tris, nrm = synth.signet(seed=2, band_nu=256, band_nv=64, plate_nt=96, plate_np=192)
    print(f"signet: {len(pos)} vertices, settings {kw}", flush=True)
    sc = _scene(ctx, pos, tris, nrm)
    try:
        settings = sc.patina_settings(**kw)
        mask, ao, cav = sc.patina_mask(settings, parts=True)
    finally:
        sc.close()
    ref_cav = O.concavity(pos, nrm, tris, settings["curvature_radius"])
    assert np.abs(cav - ref_cav).max() <= CAV_TOL, np.abs(cav - ref_cav).max()
    oracle_kw = {k: settings[k] for k in ("ao_samples", "ao_distance", "curvature_radius", "concavity_weight",
                                          "smooth_iterations")}
    ref = O.patina_mask(pos, nrm, tris, ao=ao, **oracle_kw)
    assert np.abs(mask - ref).max() <= 1e-5, np.abs(mask - ref).max()
    full = O.patina_mask(pos, nrm, tris, **oracle_kw)
    assert mask.max() > 0.3
    assert abs(float(mask.mean() - full.mean())) < 1e-3
    assert (np.abs(mask - full) < 1e-3).mean() > 0.99


def test_settings_sized_by_the_piece(ctx):
    """shot.rs:133-141, in f32."""
    pos, tris, nrm = synth.torus_ring(seed=1, nu=128, nv=32)
    sc = _scene(ctx, pos, tris, nrm)
    try:
        s = sc.patina_settings()
    finally:
        sc.close()
    lo = np.float32([pos[:, k].min() for k in range(3)])
    hi = np.float32([pos[:, k].max() for k in range(3)])
    h = (hi - lo) * np.float32(0.5)
    r = max(np.sqrt(np.float32(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])), np.float32(1e-3))
    assert s["ao_distance"] == np.clip(r * np.float32(0.09), np.float32(0.4), np.float32(2.5))
    assert s["curvature_radius"] == np.clip(r * np.float32(0.03), np.float32(0.15), np.float32(1.0))
    assert (s["ao_samples"], s["smooth_iterations"]) == (48, 3)
    assert abs(s["concavity_weight"] - 0.35) < 1e-7


def test_sphere_concavity_on_the_gpu(ctx):
    """cavity/tests.rs:70-75: a convex sphere of radius 4 reads about -0.1."""
    sp, st = synth.icosphere(4.0, 5)
    sp, sn, st, _ = O.normals(sp, st, 50.0)
    sc = _scene(ctx, sp, st, sn)
    try:
        _, _, cav = sc.patina_mask(dict(ao_samples=0, curvature_radius=0.4), parts=True)
    finally:
        sc.close()
    assert abs(cav.mean() + 0.1) < 0.02, cav.mean()
    ref = O.concavity(sp, sn, st, 0.4)
    assert np.abs(cav - ref).max() <= CAV_TOL


def test_smoothing_matches_oracle_bit_for_bit(ctx):
    """cavity/tests.rs:101-119 on the GPU, and equal to the oracle's smoothing."""
    p, t = synth.subdivided_cuboid([2.0, 2.0, 2.0], 4)
    p, n, t, _ = O.normals(p, t, 50.0)
    sc = _scene(ctx, p, t, n)
    try:
        v = sc.smooth_values(np.full(len(p), 0.7, np.float32), 5, 0.5)
        assert (np.abs(v - 0.7) < 1e-6).all()
        rng = np.random.default_rng(5)
        x = rng.random(len(p)).astype(np.float32)
        for it in (0, 1, 2, 7):
            got = sc.smooth_values(x, it, 0.5)
            assert np.array_equal(got, O.smooth(p, t, x, it, 0.5)), it
        key = (p + np.float32(0)).view(np.uint32)
        got = sc.smooth_values(x, 2, 0.5)
        for i in range(len(p)):
            assert (got[(key == key[i]).all(1)] == got[i]).all()
    finally:
        sc.close()


def test_smoothing_self_edges_and_copies(ctx):
    """A degenerate triangle gives a class a self-edge (counted in the
    divisor, not in the sum) and an unwelded copy joins its class."""
    pos, tris = height_field(6, lambda x: np.zeros_like(x))
    s = 3 * 7 + 3
    pos2 = np.concatenate([pos, pos[s:s + 1]])
    tris2 = np.concatenate([tris, [[s, len(pos), s]]]).astype(np.uint32)
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos2), 1))
    v = np.zeros(len(pos2), np.float32)
    v[s] = 1.0
    sc = _scene(ctx, pos2, tris2, nrm)
    try:
        for it in (0, 1, 3):
            got = sc.smooth_values(v, it, 0.5)
            assert np.array_equal(got, O.smooth(pos2, tris2, v, it, 0.5)), it
        assert sc.smooth_values(v, 1, 0.5)[s] == np.float32(0.25)
    finally:
        sc.close()


def test_edge_cases(ctx):
    pos, tris = height_field(20, lambda x: np.zeros_like(x))
    nrm = np.tile(np.float32([0, 0, 1]), (len(pos), 1))
    sc = _scene(ctx, pos, tris, nrm)
    try:
        assert (sc.patina_mask() == 0.0).all()  # open plane
        with pytest.raises(ag.ArgentError):
            sc.patina_mask(dict(ao_distance=float("nan")))
    finally:
        sc.close()
    # a concavity-only mask (no AO term) equals the oracle's
    pos, tris = height_field(100, groove_ridge_profile)
    p, n, t, _ = O.normals(pos, tris, 60.0)
    sc = _scene(ctx, p, t, n)
    try:
        m = sc.patina_mask(dict(ao_samples=0, concavity_weight=1.0, smooth_iterations=2))
    finally:
        sc.close()
    ref = O.patina_mask(p, n, t, ao_samples=0, concavity_weight=1.0, smooth_iterations=2)
    assert np.abs(m - ref).max() <= 1e-5
    assert m.max() > 0.5
    # no normals: the smooth vertex normals (cavity.rs:82-87), on the device as in the oracle
    d = ag.SceneDesc(pos, tris)
    d.add_material(ag.Material.metal("sterling"))
    sc = ag.Scene(ctx, d)
    try:
        m, ao, cav = sc.patina_mask(parts=True)
    finally:
        sc.close()
    sn = O.smooth_vertex_normals(pos, tris)
    assert np.abs(cav - O.concavity(pos, sn, tris, 0.4)).max() <= CAV_TOL
    assert np.abs(m - O.patina_mask(pos, None, tris, ao=ao)).max() <= 1e-5
    assert value_near(pos, m, -3.0, 0.0) > 0.5


def test_stored_mask_is_the_cavity_shading_reads(ctx):
    """Prepared::build (shot.rs:133-143): the mask stored on the piece renders
    exactly as a scene given the same mask as its cavity array, and the
    patina layer (sterling, patina 0.8) darkens the render."""
    pos, tris, nrm = synth.signet(seed=2, band_nu=256, band_nv=64, plate_nt=96, plate_np=192)
    piece = workloads.Piece(pos, tris, nrm)
    w = h = 64
    cam, turn = piece.shot("hero", 100.0, 0.0, 1.0)
    ps = workloads.path_settings(8)
    sc = ag.Scene(ctx, piece.scene("sterling", 0.07, 0.8, "noir"))
    try:
        sc.set_turntable(turn)
        f0 = ag.Film(ctx, w, h)
        ag.render_pass(ctx, sc, cam, f0, ps, 16, 3)
        mask = sc.patina_mask(sc.patina_settings(), store=True)
        f1 = ag.Film(ctx, w, h)
        ag.render_pass(ctx, sc, cam, f1, ps, 16, 3)
        # a look created over the same piece afterwards shares the stored mask
        look = ag.Scene(ctx, piece.scene("sterling", 0.07, 0.8, "noir"), share_mesh_with=sc)
        look.set_turntable(turn)
        f3 = ag.Film(ctx, w, h)
        ag.render_pass(ctx, look, cam, f3, ps, 16, 3)
        look.close()
    finally:
        sc.close()
    assert mask.max() > 0.3
    piece.cavity = mask
    sc2 = ag.Scene(ctx, piece.scene("sterling", 0.07, 0.8, "noir"))
    try:
        sc2.set_turntable(turn)
        f2 = ag.Film(ctx, w, h)
        ag.render_pass(ctx, sc2, cam, f2, ps, 16, 3)
    finally:
        sc2.close()
    c0, c1, c2, c3 = (f.color() for f in (f0, f1, f2, f3))
    assert np.array_equal(c1, c2) and np.array_equal(c1, c3)
    assert c1.sum() < c0.sum()
