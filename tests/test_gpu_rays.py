"""The 16,777,216-ray batch against the ~1.2 M-triangle signet (origins on a
30 mm sphere, directions toward the centre jittered up to 20 degrees):
closest-hit ids, t, u, v bit-exact against the oracle at full size, from
rays resident in HBM; any-hit agrees with the closest-hit miss set.
Rays whose reported hit is an f32 artefact outside its triangle (grazing)
are the one permitted difference, each verified in f64 (hitcheck.py)."""
import ctypes

import numpy as np
import pytest

import argent_amd as ag
from argent_amd import synth
from hitcheck import assert_same_closest

pytestmark = pytest.mark.gpu
O = pytest.importorskip("oracle")


def test_microbenchmark_rays_full_size(ctx):
    pos, tris, nrm = synth.signet(seed=2)
    c = (pos.min(0) + pos.max(0)) * 0.5
    n = 1 << 24
    o, d = synth.sphere_rays(n, c, seed=3)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    sc = ag.Scene(ctx, desc)
    rays = ag.DeviceRays(ctx, o, d)
    ms = [ag.intersect_device(ctx, sc, rays) for _ in range(3)]
    gt, gu, gv, gp = rays.hits_host()
    print(f"closest hit: {n / (min(ms) * 1e-3) / 1e9:.2f} Grays/s ({min(ms):.2f} ms)")
    osc = O.OracleScene(desc)
    rt, ru, rv, rp = osc.intersect(o, d, mode=O.MODE_ROBUST, threads=16)
    nbad = assert_same_closest(pos, tris, o, d, (gt, gp), (rt, rp))
    same = gp == rp
    hit = (rp != 0xFFFFFFFF) & same
    print("hit fraction", hit.mean(), "artefact rays", nbad)
    assert np.array_equal(gt[hit], rt[hit]) and np.array_equal(gu[hit], ru[hit]) and np.array_equal(gv[hit], rv[hit])
    ms_occ = ag.occluded_device(ctx, sc, rays)
    occ = np.empty(n, np.uint8)
    ag.check(ag.A.lib().ag_memcpy_to_host(ctx.h, occ.ctypes.data_as(ctypes.c_void_p), rays.out[3], n))
    assert np.array_equal(occ.astype(bool)[same], (rp != 0xFFFFFFFF)[same])
    print(f"any hit: {n / (ms_occ * 1e-3) / 1e9:.2f} Grays/s")
    rays.close()
    sc.close()
