"""The continuous-batching server: a request's image is bit-identical whether
it is served alone or in a mix of requests on other scenes; admission by HBM
budget; rejection of requests that can never fit; progressive previews."""
import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import synth, workloads
from argent_amd.serve import Server

pytestmark = pytest.mark.gpu


def _pieces():
    a = workloads.Piece(*synth.torus_ring(seed=1, nu=128, nv=32))
    b = workloads.Piece(*synth.signet(seed=2, band_nu=128, band_nv=32, plate_nt=48, plate_np=96))
    return a, b


def _serve(ctx, reqs, budget=8 << 30, step_samples=1 << 14):
    srv = Server(ctx, budget, step_samples, workloads.path_settings(8))
    ids = [srv.submit(**r) for r in reqs]
    for _ in range(1000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    out = [(srv.status(i), srv.read(i)) for i in ids]
    srv.close()
    return out


def test_alone_equals_mix(ctx):
    pa, pb = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    sb = ag.Scene(ctx, pb.scene("sterling", 0.07, 0.8, "noir"))
    ca, ta = pa.shot("hero", 100.0, 0.0, 1.0)
    cb, tb = pb.shot("front", 100.0, 16.0, 48 / 40, intersect=sb.intersect)
    gold = ag.Material.metal("gold14k", 0.07, 0.8)
    reqs = [dict(scene=sa, camera=ca, width=40, height=40, spp=12, seed=3, turntable_deg=ta),
            dict(scene=sb, camera=cb, width=48, height=40, spp=20, seed=4, turntable_deg=tb, material0=gold),
            dict(scene=sa, camera=ca, width=32, height=32, spp=7, seed=5, turntable_deg=ta + 40.0, material0=gold)]
    mixed = _serve(ctx, reqs)
    for k, r in enumerate(reqs):
        st, img = mixed[k]
        assert st["state"] == A.REQ_DONE and st["samples"] == r["spp"]
        assert 0 <= st["t_submit"] <= st["t_admit"] <= st["t_first_preview"] <= st["t_done"]
        alone = _serve(ctx, [r])[0][1]
        assert np.array_equal(alone, img), k
    sa.close()
    sb.close()


def test_budget_admission_and_rejection(ctx):
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    film_bytes = 32 * 32 * 9 * 4 * 2
    srv = Server(ctx, budget_bytes=film_bytes + 10, step_samples=1 << 12, ps=workloads.path_settings(6))
    big = srv.submit(sa, cam, 64, 64, 4, 1, turn)
    a1 = srv.submit(sa, cam, 32, 32, 4, 1, turn)
    a2 = srv.submit(sa, cam, 32, 32, 4, 2, turn)
    assert srv.status(big)["state"] == A.REQ_REJECTED
    active, queued = srv.step()
    assert (active, queued) in [(1, 1), (0, 1)]  # only one film fits at a time
    assert srv.status(a2)["state"] == A.REQ_QUEUED
    while srv.status(a1)["state"] != A.REQ_DONE:
        srv.step()
    prev = srv.read(a1)
    assert np.all(np.isfinite(prev)) and prev.mean() > 0
    srv.release(a1)
    for _ in range(100):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    assert srv.status(a2)["state"] == A.REQ_DONE
    srv.close()
    sa.close()
