"""Scenes built straight from a device mesh (ag_scene_create_from_mesh):
the whole path a new STL takes in serving -- binary STL, weld and crease
normals on the device, scene and BVH on the device -- must give the hits of
the oracle's chain (io.rs load, normals.rs, bvh.rs) and the same image as a
scene built from the same arrays on the host."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import synth, workloads
from hitcheck import assert_same_closest

pytestmark = pytest.mark.gpu


def _ingest(ctx, pos, tris):
    data = synth.to_binary_stl(pos, tris)
    m, _ = ag.load_stl(ctx, data)
    m.normals(50.0)
    return m, data


def test_signet_from_stl_hits_match_oracle(ctx):
    pos, tris, _ = synth.signet()
    m, data = _ingest(ctx, pos, tris)
    mp, mt = m.read()
    mn = m.read_normals()
    desc = ag.SceneDesc(mp, mt, normals=mn)
    desc.add_material(ag.Material.metal("sterling"))
    sc = ag.Scene.from_mesh(ctx, m, desc, gpu_bvh=True)
    m.close()  # the scene keeps its own copy
    c = (mp.min(0) + mp.max(0)) * 0.5
    o, d = synth.sphere_rays(1 << 20, c, seed=5)
    gt, gu, gv, gp = sc.intersect(o, d)
    # the oracle's chain from the same bytes
    wp, wt, _ = O.load_stl(data)
    wq, wn, ws, _ = O.normals(wp, wt, 50.0)
    assert np.array_equal(wq.view(np.uint32), mp.view(np.uint32)) and np.array_equal(ws, mt)
    od = ag.SceneDesc(wq, ws, normals=wn)
    od.add_material(ag.Material.metal("sterling"))
    rt, ru, rv, rp = O.OracleScene(od).intersect(o, d, mode=O.MODE_ROBUST, threads=16)
    assert_same_closest(wq, ws, o, d, (gt, gp), (rt, rp))
    hit = (rp != 0xFFFFFFFF) & (gp == rp)
    assert np.array_equal(gt[hit], rt[hit]) and np.array_equal(gu[hit], ru[hit]) and np.array_equal(gv[hit], rv[hit])
    sc.close()


@pytest.mark.parametrize("gpu_bvh", [True, False])
def test_render_equals_host_array_scene(ctx, gpu_bvh):
    pos, tris, _ = synth.torus_ring()
    m, _ = _ingest(ctx, pos, tris)
    mp, mt = m.read()
    mn = m.read_normals()
    piece = workloads.Piece(mp, mt, mn)
    desc = piece.scene("gold14k", 0.07, 0.0, "noir")
    cam, turn = piece.shot("hero", 100.0, 0.0, 1.0)
    desc.d.turntable_deg = turn
    imgs = []
    for sc in (ag.Scene.from_mesh(ctx, m, desc, gpu_bvh=gpu_bvh), ag.Scene(ctx, desc, gpu_bvh=gpu_bvh)):
        film = ag.Film(ctx, 96, 96)
        ag.render_pass(ctx, sc, cam, film, workloads.path_settings(8), 8, 11)
        imgs.append(film.color())
        film.close()
        sc.close()
    m.close()
    assert np.array_equal(imgs[0], imgs[1])


def test_out_of_range_index_rejected(ctx):
    m = ag.Mesh.from_arrays(ctx, np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))
    desc = ag.SceneDesc(np.zeros((3, 3), np.float32), np.array([[0, 1, 2]], np.uint32))
    desc.add_material(ag.Material.metal("sterling"))
    with pytest.raises(ag.ArgentError):
        ag.Scene.from_mesh(ctx, m, desc)
    m.close()


def test_frame_sample_equals_host_framing(ctx):
    """ag_mesh_frame_sample gives Piece's bounds and hull exactly, and the
    framed camera is the one framed from the positions on the host."""
    pos, tris, _ = synth.signet(seed=2, band_nu=128, band_nv=32, plate_nt=48, plate_np=96)
    m = ag.Mesh.from_arrays(ctx, pos, tris)
    hull, lo, hi = m.frame_sample()
    host = workloads.Piece(pos, tris)
    assert np.array_equal(lo, pos.min(0)) and np.array_equal(hi, pos.max(0))
    assert np.array_equal(hull, host.hull)
    dev = workloads.Piece.from_frame(lo, hi, hull)
    assert dev.radius == host.radius and dev.floor_y == host.floor_y
    for aspect in (1.0, 540 / 960):
        ca, ta = host.shot("hero", 100.0, 0.0, aspect)
        cb, tb = dev.shot("hero", 100.0, 0.0, aspect)
        assert ta == tb and bytes(ca) == bytes(cb)
    for stride in (1, 7, len(pos) + 3):
        h, lo2, hi2 = m.frame_sample(stride)
        assert np.array_equal(h, pos[::stride]) and np.array_equal(lo2, lo) and np.array_equal(hi2, hi)
    with pytest.raises(ag.ArgentError):
        m.frame_sample(0)
    m.close()
