"""BVH build on the device (ag_scene_create_ex with AG_SCENE_GPU_BVH): the
tree differs from the host's binned SAH, closest-hit results must not.  Hits
against the oracle (ids, t, u, v bit-exact, with hitcheck.py's one permitted
grazing artefact), any-hit against the closest-hit miss set, the full
16.8 M-ray batch on the signet, small and degenerate meshes (one triangle,
at most one leaf, coincident centroids, Morton ties), and a rendered image
equal to the host-tree render."""
import ctypes

import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth, workloads
from hitcheck import assert_same_closest

pytestmark = pytest.mark.gpu


def _scene(ctx, pos, tris, gpu):
    desc = ag.SceneDesc(pos, tris)
    desc.add_material(ag.Material.metal("sterling"))
    return desc, ag.Scene(ctx, desc, gpu_bvh=gpu)


def _check(ctx, pos, tris, o, d, tag):
    desc, sc = _scene(ctx, pos, tris, True)
    st = sc.stats()
    gt, gu, gv, gp = sc.intersect(o, d)
    osc = O.OracleScene(desc)
    rt, ru, rv, rp = osc.intersect(o, d, mode=O.MODE_ROBUST, threads=16)
    assert_same_closest(pos, tris, o, d, (gt, gp), (rt, rp))
    same = gp == rp
    hit = (rp != 0xFFFFFFFF) & same
    assert np.array_equal(gt[hit], rt[hit]) and np.array_equal(gu[hit], ru[hit]), tag
    assert np.array_equal(gv[hit], rv[hit]), tag
    occ = sc.occluded(o, d)
    assert np.array_equal(occ.astype(bool)[same], (rp != 0xFFFFFFFF)[same]), tag
    sc.close()
    return st


def _rays(pos, n, seed):
    c = (pos.min(0) + pos.max(0)) * 0.5
    r = float(np.linalg.norm(pos.max(0) - pos.min(0))) + 1.0
    return synth.sphere_rays(n, c, radius=r, seed=seed)


def test_small_and_degenerate_meshes(ctx):
    rng = np.random.default_rng(7)
    cases = []
    # one triangle, three, five (one leaf, then two)
    for k in (1, 3, 5):
        p = rng.uniform(-1, 1, (3 * k, 3)).astype(np.float32)
        cases.append((f"soup{k}", p, np.arange(3 * k, dtype=np.uint32).reshape(-1, 3)))
    # 200 triangles sharing one centroid (every Morton code equal)
    a = rng.normal(size=(200, 3)).astype(np.float32)
    p = np.concatenate([a, -a, np.zeros_like(a)]).astype(np.float32)
    t = np.stack([np.arange(200), 200 + np.arange(200), 400 + np.arange(200)], 1).astype(np.uint32)
    cases.append(("same-centroid", p, t))
    # a flat grid (zero extent on one axis) and the cube
    g = np.stack(np.meshgrid(np.arange(40), np.arange(30), indexing="ij"), -1).reshape(-1, 2).astype(np.float32)
    gp = np.concatenate([g, np.zeros((len(g), 1), np.float32)], 1)
    q = np.array([[i * 30 + j, (i + 1) * 30 + j, i * 30 + j + 1] for i in range(39) for j in range(29)], np.uint32)
    cases.append(("flat", gp, q))
    cases.append(("cube", *synth.cuboid([2.0, 3.0, 4.0])))
    for tag, p, t in cases:
        o, d = _rays(p, 20000, 5)
        _check(ctx, p, t, o, d, tag)


def test_torus_ring(ctx):
    pos, tris, _ = synth.torus_ring()
    o, d = _rays(pos, 1 << 18, 9)
    st = _check(ctx, pos, tris, o, d, "torus")
    assert st["nodes"] > 1000


def test_signet_full_batch_and_tree(ctx):
    pos, tris, nrm = synth.signet(seed=2)
    c = (pos.min(0) + pos.max(0)) * 0.5
    n = 1 << 24
    o, d = synth.sphere_rays(n, c, seed=3)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    sc = ag.Scene(ctx, desc, gpu_bvh=True)
    print("gpu tree", sc.stats())
    rays = ag.DeviceRays(ctx, o, d)
    ms = min(ag.intersect_device(ctx, sc, rays) for _ in range(3))
    gt, gu, gv, gp = rays.hits_host()
    print(f"closest hit on the GPU-built tree: {n / (ms * 1e-3) / 1e9:.2f} Grays/s")
    osc = O.OracleScene(desc)
    rt, ru, rv, rp = osc.intersect(o, d, mode=O.MODE_ROBUST, threads=16)
    assert_same_closest(pos, tris, o, d, (gt, gp), (rt, rp))
    same = gp == rp
    hit = (rp != 0xFFFFFFFF) & same
    assert np.array_equal(gt[hit], rt[hit]) and np.array_equal(gu[hit], ru[hit]) and np.array_equal(gv[hit], rv[hit])
    ag.occluded_device(ctx, sc, rays)
    occ = np.empty(n, np.uint8)
    ag.check(ag.A.lib().ag_memcpy_to_host(ctx.h, occ.ctypes.data_as(ctypes.c_void_p), rays.out[3], n))
    assert np.array_equal(occ.astype(bool)[same], (rp != 0xFFFFFFFF)[same])
    rays.close()
    sc.close()


def test_render_matches_host_tree(ctx):
    wl = workloads.torus_still()
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])
    cam, turn = wl["piece"].shot("hero", 100.0, 0.0, 1.0)
    desc.d.turntable_deg = turn
    imgs = []
    for gpu in (False, True):
        sc = ag.Scene(ctx, desc, gpu_bvh=gpu)
        film = ag.Film(ctx, 128, 128)
        ag.render_pass(ctx, sc, cam, film, workloads.path_settings(8), 16, 3)
        imgs.append(film.color())
        film.close()
        sc.close()
    # same closest hits, same shading: equal up to grazing-ray artefacts
    diff = np.abs(imgs[0] - imgs[1])
    assert np.mean(diff > 1e-6) < 1e-3, float(np.mean(diff > 1e-6))


def test_gems_scene_render_matches_host_tree(ctx):
    # 12 dielectric stones set in the plate (separate shells, refraction and
    # Beer-Lambert inside), the HDRI and the softboxes, depth 24
    wl = workloads.signet_gems_still(hdri_size=(256, 128))
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"], gems=wl["gems"])
    desc.set_env_map(wl["hdri"])
    cam, turn = wl["piece"].shot("hero", 100.0, 0.0, 1.0)
    desc.d.turntable_deg = turn
    imgs = []
    for gpu in (False, True):
        sc = ag.Scene(ctx, desc, gpu_bvh=gpu)
        film = ag.Film(ctx, 96, 96)
        ag.render_pass(ctx, sc, cam, film, workloads.path_settings(24), 8, 5)
        imgs.append(film.color())
        film.close()
        sc.close()
    diff = np.abs(imgs[0] - imgs[1])
    assert np.mean(diff > 1e-6) < 2e-3, float(np.mean(diff > 1e-6))
