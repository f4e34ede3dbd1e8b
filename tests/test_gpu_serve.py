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


def _serve(ctx, reqs, budget=64 << 30, step_samples=1 << 14):
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


def _one_film_budget(ctx, sa, cam, turn):
    """The budget that admits exactly one 32 x 32 request on `sa`: what a server
    charges with it admitted (its film, the scene look and piece, the wave state)."""
    probe = Server(ctx, budget_bytes=1 << 50, step_samples=1 << 12, ps=workloads.path_settings(6))
    # a first request sizes the wave state for the scene's material keys
    first = probe.submit(sa, cam, 8, 8, 1, 1, turn)
    probe.step()
    probe.release(first)
    idle, _ = probe.budget()
    i = probe.submit(sa, cam, 32, 32, 1, 1, turn)
    j = probe.submit(sa, cam, 32, 32, 1, 2, turn)
    probe.step()
    used2, _ = probe.budget()
    assert probe.status(i)["state"] == A.REQ_DONE and probe.status(j)["state"] == A.REQ_DONE
    film_bytes = 32 * 32 * 9 * 4 * 2
    # two films, and the look and piece they share charged once
    assert used2 - idle == 2 * film_bytes + sa.stats()["bytes"]
    probe.release(j)
    used = probe.budget()[0]
    assert used - idle == film_bytes + sa.stats()["bytes"]
    probe.release(i)
    assert probe.budget()[0] == idle
    probe.close()
    return used


def test_budget_admission_and_rejection(ctx):
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    srv = Server(ctx, budget_bytes=_one_film_budget(ctx, sa, cam, turn) + 10, step_samples=1 << 12,
                 ps=workloads.path_settings(6))
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


def test_failed_request_does_not_stop_the_others(ctx):
    """A request whose film cannot exist on the device (65535 x 65535: 309 GB)
    fails alone with a reason; the requests around it finish and equal their
    renders alone (queue.rs:514-557 isolates each job the same way)."""
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    ok1 = dict(scene=sa, camera=cam, width=32, height=32, spp=6, seed=3, turntable_deg=turn)
    ok2 = dict(scene=sa, camera=cam, width=24, height=40, spp=5, seed=4, turntable_deg=turn)
    srv = Server(ctx, 1 << 50, 1 << 14, workloads.path_settings(8))
    i1 = srv.submit(**ok1)
    bad = srv.submit(sa, cam, 65535, 65535, 4, 1, turn)
    i2 = srv.submit(**ok2)
    for _ in range(200):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    assert srv.status(bad)["state"] == A.REQ_FAILED
    assert "free memory" in srv.error(bad)
    with pytest.raises(ag.ArgentError):
        srv.read(bad)
    imgs = []
    for i in (i1, i2):
        assert srv.status(i)["state"] == A.REQ_DONE and srv.error(i) == ""
        imgs.append(srv.read(i))
    srv.release(bad)
    srv.close()
    assert np.array_equal(imgs[0], _serve(ctx, [ok1])[0][1])
    assert np.array_equal(imgs[1], _serve(ctx, [ok2])[0][1])
    sa.close()


def test_render_failure_fails_only_its_request(ctx):
    """A request whose chunks fail to render (forced from the host: every pass
    with a view at seed 77 fails before launching) fails alone once retried by
    itself; the requests that shared its steps finish and equal their renders
    alone."""
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    ok1 = dict(scene=sa, camera=cam, width=32, height=32, spp=6, seed=3, turntable_deg=turn)
    ok2 = dict(scene=sa, camera=cam, width=24, height=40, spp=5, seed=4, turntable_deg=turn)
    ctx.debug_fail_seed(77)
    try:
        srv = Server(ctx, 64 << 30, 1 << 14, workloads.path_settings(8))
        i1 = srv.submit(**ok1)
        bad = srv.submit(sa, cam, 16, 16, 4, 77, turn)
        i2 = srv.submit(**ok2)
        for _ in range(200):
            a, q = srv.step()
            if a == 0 and q == 0:
                break
        assert srv.status(bad)["state"] == A.REQ_FAILED
        assert "injected" in srv.error(bad)
        imgs = []
        for i in (i1, i2):
            assert srv.status(i)["state"] == A.REQ_DONE and srv.error(i) == ""
            imgs.append(srv.read(i))
        used, _ = srv.budget()
        srv.release(bad)
        assert srv.budget()[0] == used  # the failed request's film was freed when it failed
        srv.close()
    finally:
        ctx.debug_fail_seed(None)
    assert np.array_equal(imgs[0], _serve(ctx, [ok1])[0][1])
    assert np.array_equal(imgs[1], _serve(ctx, [ok2])[0][1])
    sa.close()


def test_live_pass_failure_does_not_fail_the_request_beside_it(ctx):
    """A step that carries an attached live view's pass and one request: when
    the live pass is what fails (seed 1, the live view's), the request is
    retried alone and carries on; its image equals its render alone."""
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    req = dict(scene=sa, camera=cam, width=32, height=32, spp=6, seed=5, turntable_deg=turn)
    srv = Server(ctx, 64 << 30, 1 << 14, workloads.path_settings(8))
    rid = srv.submit(**req)
    lv = ag.Live(ctx)
    lv.attach(srv)
    lv.set(sa, cam, 32, 32, max_spp=8, now_ms=0.0)
    ctx.debug_fail_seed(1)
    try:
        with pytest.raises(ag.ArgentError):
            lv.step(now_ms=0.0)
    finally:
        ctx.debug_fail_seed(None)
    st = srv.status(rid)
    assert st["state"] == A.REQ_ACTIVE and st["samples"] > 0, st
    lv.attach(None)
    for _ in range(200):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    assert srv.status(rid)["state"] == A.REQ_DONE
    img = srv.read(rid)
    srv.close()
    lv.close()
    assert np.array_equal(img, _serve(ctx, [req])[0][1])
    sa.close()


def test_arrivals_every_step_do_not_starve_work_in_progress(ctx):
    """A new request before every step: at most 4 preview-only steps run in a
    row, so a request in progress keeps gaining samples."""
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    srv = Server(ctx, 1 << 40, 32 * 32 * 4, workloads.path_settings(4))
    long_id = srv.submit(sa, cam, 32, 32, 64, 1, turn)
    srv.step()
    assert srv.status(long_id)["t_first_preview"] >= 0
    seen = [srv.status(long_id)["samples"]]
    for k in range(15):
        srv.submit(sa, cam, 16, 16, 2, 100 + k, turn)
        srv.step()
        seen.append(srv.status(long_id)["samples"])
    grew = [i for i in range(1, len(seen)) if seen[i] > seen[i - 1]]
    assert len(grew) >= 3, seen  # progress at least every fifth step
    assert max(np.diff([0] + grew)) <= 5, seen
    srv.close()
    sa.close()


def test_new_request_previews_before_long_one_finishes(ctx):
    """With a step budget smaller than the work in flight, a request that
    arrives while a long one renders gets its first preview at the next step,
    a step that renders nothing else, and, having less work left, finishes first; both still match their
    renders alone."""
    pa, _ = _pieces()
    sa = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    cam, turn = pa.shot("hero", 100.0, 0.0, 1.0)
    step = 48 * 48 * 4
    srv = Server(ctx, 64 << 30, step, workloads.path_settings(8))  # the depth _serve renders with
    long_id = srv.submit(sa, cam, 48, 48, 64, 1, turn)
    for _ in range(3):
        srv.step()
    assert srv.status(long_id)["state"] == A.REQ_ACTIVE
    short_id = srv.submit(sa, cam, 24, 24, 8, 2, turn)
    before = srv.status(long_id)["samples"]
    srv.step()
    st_s = srv.status(short_id)
    assert st_s["t_first_preview"] >= 0, "no preview after one step"
    # that step carried only the request awaiting its first preview
    assert srv.status(long_id)["samples"] == before and st_s["samples"] == 2
    for _ in range(1000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    st_l, st_s = srv.status(long_id), srv.status(short_id)
    assert st_l["state"] == A.REQ_DONE and st_s["state"] == A.REQ_DONE
    assert st_s["t_done"] < st_l["t_done"]
    img_l, img_s = srv.read(long_id), srv.read(short_id)
    srv.close()
    assert np.array_equal(img_l, _serve(ctx, [dict(scene=sa, camera=cam, width=48, height=48, spp=64, seed=1,
                                                   turntable_deg=turn)], step_samples=step)[0][1])
    assert np.array_equal(img_s, _serve(ctx, [dict(scene=sa, camera=cam, width=24, height=24, spp=8, seed=2,
                                                   turntable_deg=turn)], step_samples=step)[0][1])
    sa.close()


def test_scenes_sharing_a_mesh_render_as_their_own(ctx):
    """ag_scene_create_shared: a second look over the same device mesh and BVH
    renders exactly as a scene built from scratch; the mesh outlives the scene
    it was built with."""
    pa, _ = _pieces()
    base = ag.Scene(ctx, pa.scene("sterling", 0.07, 0.8, "white"))
    d_noir = pa.scene("gold14k", 0.07, 0.8, "noir")
    shared = ag.Scene(ctx, d_noir, share_mesh_with=base)
    own = ag.Scene(ctx, pa.scene("gold14k", 0.07, 0.8, "noir"))
    base.close()  # the shared scene keeps the mesh alive
    cam, turn = pa.shot("hero", 100.0, 16.0, 1.0)
    imgs = []
    for sc in (shared, own):
        film = ag.Film(ctx, 40, 40)
        sc.desc.d.turntable_deg = turn
        ag.render_views(ctx, sc, [dict(camera=cam, film=film, count=4, seed=9, turntable_deg=turn)],
                        workloads.path_settings(8))
        imgs.append(film.sums())
        film.close()
    assert np.array_equal(imgs[0], imgs[1])
    shared.close()
    own.close()
