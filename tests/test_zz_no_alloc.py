"""ag_intersect_device_noalloc (the allocation-free closest-hit query) against
ag_intersect_device through the C ABI: ids, t, u and v bit for bit, on two
calls in a row, the second over fewer rays (the scratch array keeps the first
call's cursors and stacks, so the second call proves they are reset)."""
import ctypes as C

import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import synth

pytestmark = pytest.mark.gpu


def _query(ctx, sc, rays, n, noalloc):
    hits = A.HitsSoA(rays.hits.t, rays.hits.u, rays.hits.v, rays.hits.prim)
    if noalloc:
        A.check(A.lib().ag_intersect_device_noalloc(ctx.h, sc.h, C.byref(rays.rays), n, C.byref(hits)))
        A.check(A.lib().ag_ctx_synchronize(ctx.h))
    else:
        A.check(A.lib().ag_intersect_device(ctx.h, sc.h, C.byref(rays.rays), n, C.byref(hits)))
    out = rays.hits_host()
    return tuple(a[:n].copy() for a in out)


def _poison(ctx, rays):
    """Fill the hit buffers with a pattern no query writes."""
    bad = np.full(rays.n, 0x7FC00001, np.uint32)
    for p in rays.out:
        A.check(A.lib().ag_memcpy_to_device(ctx.h, p, bad.ctypes.data_as(C.c_void_p), bad.nbytes))


def test_noalloc_equals_device_query_twice(ctx):
    pos, tris, nrm = synth.torus_ring(seed=1, nu=256, nv=64)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    sc = ag.Scene(ctx, desc)
    c = (pos.min(0) + pos.max(0)) * 0.5
    n1 = 300_000
    o, d = synth.sphere_rays(n1, c, radius=30.0, seed=11)
    rays = ag.DeviceRays(ctx, o, d)
    try:
        for n in (n1, 70_001):
            print(f"noalloc query: {n} rays", flush=True)
            _poison(ctx, rays)
            want = _query(ctx, sc, rays, n, noalloc=False)
            _poison(ctx, rays)
            got = _query(ctx, sc, rays, n, noalloc=True)
            assert (want[3] != 0xFFFFFFFF).mean() > 0.2  # the batch hits the piece
            for a, b in zip(want, got):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    finally:
        rays.close()
        sc.close()


def test_noalloc_rejects_missing_buffers(ctx):
    pos, tris = synth.cuboid((2, 2, 2))
    desc = ag.SceneDesc(pos, tris)
    desc.add_material(ag.Material.diffuse([0.5] * 3))
    sc = ag.Scene(ctx, desc)
    try:
        rays = A.RaysSoA()
        hits = A.HitsSoA()
        rc = A.lib().ag_intersect_device_noalloc(ctx.h, sc.h, C.byref(rays), 10, C.byref(hits))
        assert rc == 1 and b"missing" in A.lib().ag_last_error()
    finally:
        sc.close()
