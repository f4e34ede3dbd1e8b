"""The HIP path's light transport against closed forms (the same cases as
test_transport_oracle.py, which pins the oracle by them), and against the
oracle at matched samples.  The GPU renders with the reference's MIS only; the
closed forms fix NEE to panels and to the importance-sampled map, the MIS
weights, the multiple-scattering conductor, Fresnel, Beer-Lambert, Russian
roulette and the indirect clamp independently of either restatement."""
import numpy as np
import pytest

import argent_amd as ag
from argent_amd import synth

pytestmark = pytest.mark.gpu

O = pytest.importorskip("oracle")

import transportcases as T  # noqa: E402
from test_transport_oracle import FURNACE_EXPECT, FURNACE_MATS, SLABS  # noqa: E402


def gpu_render(ctx, desc, cam, w, h, spp, seed=1, ps=None, xyz=False):
    sc = ag.Scene(ctx, desc)
    film = ag.Film(ctx, w, h)
    try:
        ag.render_pass(ctx, sc, cam, film, ps or ag.PathSettings(), spp, seed)
        return film.xyz() if xyz else film.sums()[0] / spp
    finally:
        film.close()
        sc.close()


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-20))


@pytest.mark.parametrize("alloy", T.GOLDPROBE_ALLOYS)
def test_goldprobe_gpu(ctx, alloy):
    desc = T.goldprobe_desc(alloy)
    cam = T.front_camera(32, 32)
    img = gpu_render(ctx, desc, cam, 32, 32, 512)
    assert np.all(np.abs(img[16, 16] - np.array(T.F0_DOC[alloy])) < 2e-3), (alloy, img[16, 16])
    ref = T.render_oracle(desc, cam, 32, 32, 512)
    assert rel_rmse(img, ref) < 0.01


@pytest.mark.parametrize("env", ["gradient", "hdri", "hdri+panel"])
@pytest.mark.parametrize("mat", sorted(FURNACE_MATS))
def test_uniform_furnace_gpu(ctx, env, mat):
    want, tol = FURNACE_EXPECT[mat]
    desc = T.furnace_desc(FURNACE_MATS[mat](), env)
    cam = T.front_camera(64, 64)
    img = gpu_render(ctx, desc, cam, 64, 64, 256)
    r = T.block_ratio(img)
    assert np.all(np.abs(r - want) < tol), (env, mat, r)
    assert rel_rmse(img, T.render_oracle(desc, cam, 64, 64, 256)) < 0.01


def test_plane_under_panel_gpu(ctx):
    w = h = 48
    cam = T.plane_camera(w, h)
    exp = T.plane_expected(cam, w, h)
    m = T.interior(exp)
    desc = T.plane_under_panel_desc()
    img = gpu_render(ctx, desc, cam, w, h, 1024)
    y = img.mean(-1)
    assert abs(y[m].sum() / exp[m].sum() - 1) < 0.004
    assert np.sqrt(np.mean((y[m] - exp[m]) ** 2)) / np.sqrt(np.mean(exp[m] ** 2)) < 0.008
    assert rel_rmse(img, T.render_oracle(desc, cam, w, h, 1024)) < 0.01


@pytest.mark.parametrize("eta,sigma,rr,tol", SLABS)
def test_absorbing_slab_gpu(ctx, eta, sigma, rr, tol):
    w = h = 48
    cam = T.front_camera(w, h)
    exp = T.slab_expected(cam, w, h, eta, sigma)
    m = ~np.isnan(exp)
    desc = T.slab_desc(eta, sigma)
    ps = ag.PathSettings(rr_depth=rr)
    img = gpu_render(ctx, desc, cam, w, h, 256, ps=ps)
    bg = img[:2, :2].reshape(-1, 3).mean()
    r = img.mean(-1)[m] / bg / exp[m]
    assert abs(r.mean() - 1) < tol, r.mean()
    if eta == 1.0:
        assert np.all(np.abs(r - 1) < 0.005), (r.min(), r.max())
    assert rel_rmse(img, T.render_oracle(desc, cam, w, h, 256, ps=ps)) < 0.01


def test_slab_spectral_absorption_gpu(ctx):
    """The film's XYZ readout (AG_FILM_XYZ) of the spectral slab against 1 nm quadrature."""
    lam = np.arange(360, 831)
    sigma = np.where(lam < 550, 0.5, 0.0).astype(np.float32)
    m = T.slab_material(1.0, 0.0)
    for i in range(len(lam)):
        m.spec_b[i] = float(sigma[i])
    p, t = synth.cuboid(T.SLAB)
    desc = ag.SceneDesc(p, t)
    desc.add_material(m)
    desc.set_gradient_env([1, 1, 1], [1, 1, 1], [1, 1, 1], 1.0)
    w = h = 32
    xyz = gpu_render(ctx, desc, T.front_camera(w, h), w, h, 1024, xyz=True)
    slab = xyz[12:20, 12:20].reshape(-1, 3).mean(0)
    bg = xyz[:2, :2].reshape(-1, 3).mean(0)
    want = T.slab_spectral_expected(sigma)
    assert np.all(np.abs(slab / bg - want) < 0.004), (slab / bg, want)


@pytest.mark.parametrize("mat", ["grey", "gold-r0.3"])
def test_hotspot_map_gpu_matches_each_strategy(ctx, mat):
    """Under the hot-spot map with the clamp off, the GPU's MIS estimate agrees
    with the oracle's light-sampling-only and BSDF-sampling-only estimates."""
    m = ag.Material.diffuse([0.5, 0.5, 0.5]) if mat == "grey" else ag.Material.metal("gold14k", 0.3, None)
    desc = T.sphere_desc(m)
    desc.set_gradient_env([0, 0, 0], [0, 0, 0], [0, 0, 0], 1.0)
    desc.set_env_map(synth.hdri(**T.HDRI_SPOTS), 0.0, 30.0)
    cam = ag.look_at([0, 0, 40], [0, 0, 0], [0, 1, 0], 100.0, 0.0, 1.0)
    ps = ag.PathSettings(clamp_indirect=1e30)
    g = gpu_render(ctx, desc, cam, 32, 32, 2048, seed=5, ps=ps)[12:20, 12:20].reshape(-1, 3).mean(0)
    for s in (O.STRATEGY_NEE, O.STRATEGY_BSDF):
        r = T.render_oracle(desc, cam, 32, 32, 2048, seed=3, strategy=s, ps=ps, rows=(12, 20))[12:20, 12:20]
        r = r.reshape(-1, 3).mean(0)
        assert np.all(np.abs(g / r - 1) < 0.025), (s, g, r)


def test_clamp_in_flat_furnace_gpu(ctx):
    cam = T.front_camera(64, 64)
    cases = [(ag.Material.diffuse([0.5] * 3), 100.0, 0.24), (ag.Material.diffuse([0.5] * 3), 40.0, 0.5),
             (T.unit_fresnel_conductor(0.3), 100.0, 0.24), (T.unit_fresnel_conductor(0.02), 100.0, 1.0)]
    for mat, k, want in cases:
        r = T.block_ratio(gpu_render(ctx, T.flat_furnace_desc(mat, k), cam, 64, 64, 256))
        assert np.all(np.abs(r - want) < 0.004 + 0.005 * want), (k, want, r)


def test_russian_roulette_from_depth_zero_gpu(ctx):
    img = gpu_render(ctx, T.furnace_desc(ag.Material.diffuse([0.5] * 3), "gradient"), T.front_camera(64, 64), 64,
                     64, 256, ps=ag.PathSettings(rr_depth=0))
    assert np.all(np.abs(T.block_ratio(img) - 0.5) < 0.015)
