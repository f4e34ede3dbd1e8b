"""The live view's render worker (live.rs:637-830) on the GPU, driven with an
injected clock: first look at half resolution and 2 spp, 1-spp passes while
in motion, pass sizing toward 70 ms, publish pacing, convergence at max_spp,
the settle test, generation restarts, look-only re-development.  Frames are
checked against the same pipeline run by hand (render at seed 1, à-trous,
develop) and against the oracle's render, à-trous and develop of the same
passes; relative_change against the oracle."""
import numpy as np
import pytest

import argent_amd as ag
import oracle as O
from argent_amd import workloads

pytestmark = pytest.mark.gpu

W, H = 96, 128


@pytest.fixture(scope="module")
def piece_scene(ctx):
    wl = workloads.torus_still()
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])
    cam, turn = wl["piece"].shot("hero", 100.0, 0.0, W / H)
    desc.d.turntable_deg = turn
    sc = ag.Scene(ctx, desc)
    yield wl, sc, cam
    sc.close()


def _hand_frame(ctx, sc, cam, w, h, passes, tonemap="agx_punchy", ev=0.0):
    """The same pipeline by hand: passes of the given sample counts at seed 1
    (a film sums per pass, render.rs:107-111, so the split matters), then
    à-trous and develop."""
    film = ag.Film(ctx, w, h)
    for n in passes:
        ag.render_pass(ctx, sc, cam, film, ag.PathSettings(), n, 1)
    out = film.develop(tonemap, exposure_ev=ev, denoise=True)
    film.close()
    return out


def _oracle_frame(wl, cam, turn, w, h, passes, tonemap="agx_punchy"):
    """The live frame's pipeline on the oracle: the same passes at seed 1
    summed pass by pass (render.rs:107-111), the film's means
    (render.rs:35-64), à-trous with the albedo and normal guides, develop."""
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])
    desc.d.turntable_deg = turn
    osc = O.OracleScene(desc)
    sums = np.zeros((3, h, w, 3), np.float32)
    first = 0
    for n in passes:
        osc.render_pass(cam, w, h, sums, first, ag.PathSettings(), n, 1)
        first += n
    osc.close()
    k = np.float32(1.0) / np.float32(first)
    color, albedo, nv = sums[0] * k, sums[1] * k, sums[2] * k
    ln = np.sqrt(nv[..., 0] * nv[..., 0] + nv[..., 1] * nv[..., 1] + nv[..., 2] * nv[..., 2])[..., None]
    normal = np.where(ln > 1e-6, nv / np.maximum(ln, np.float32(1e-30)), np.float32(0.0)).astype(np.float32)
    return O.develop(O.atrous(color, albedo, normal), tonemap)


def _byte_gap(a, b):
    d = np.abs(a.astype(np.int32) - b.astype(np.int32))
    return float(d.mean()), float(np.mean(d > 2)), int(d.max())


def test_live_frames_match_oracle_pipeline(ctx, piece_scene):
    """The live view's first look and its final frame against the oracle's
    render, à-trous and develop of the same passes.  The device and the
    oracle agree on the films to float rounding (about 1e-3 relative RMSE at
    a few spp); after denoising and 8-bit development that leaves the bytes
    equal or, where a value lands next to a rounding boundary, one step
    apart: mean gap at most 0.05 step, no value more than 1 step apart
    (measured: 0.002 step, 1)."""
    wl, sc, cam = piece_scene
    _, turn = wl["piece"].shot("hero", 100.0, 0.0, W / H)
    lv = ag.Live(ctx)
    lv.set(sc, cam, W, H, max_spp=16, now_ms=0.0)
    f, img = lv.step(now_ms=0.0)
    assert f["published"] and f["half"]
    gap = _byte_gap(img, _oracle_frame(wl, cam, turn, W // 2, H // 2, [2]))
    print("first look vs oracle: mean step gap, share > 2 steps, max", gap)
    assert gap[0] <= 0.05 and gap[2] <= 1, gap
    t, passes = 1000.0, []
    while not f["converged"]:
        f, img = lv.step(now_ms=t)
        passes.append(f["pass_spp"])
        t += 2000.0
    assert f["spp"] == 16
    gap = _byte_gap(img, _oracle_frame(wl, cam, turn, W, H, passes))
    print("final frame vs oracle: mean step gap, share > 2 steps, max", gap, "passes", passes)
    assert gap[0] <= 0.05 and gap[2] <= 1, gap
    lv.close()


def test_relative_change_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    for h, w in [(7, 5), (540, 960), (33, 1)]:
        a = rng.uniform(0, 4, (h, w, 3)).astype(np.float32)
        b = (a + rng.normal(0, 0.05, a.shape)).astype(np.float32)
        got, want = ag.relative_change(ctx, a, b), O.relative_change(a, b)
        assert abs(got - want) <= 1e-12 * max(want, 1e-30), (h, w, got, want)
    assert ag.relative_change(ctx, a, np.zeros_like(a)) == 0.0


def test_first_look_then_refinement(ctx, piece_scene):
    wl, sc, cam = piece_scene
    lv = ag.Live(ctx)
    lv.set(sc, cam, W, H, max_spp=16, now_ms=0.0)
    f, img = lv.step(now_ms=0.0)
    # live.rs:744-756: half resolution, 2 spp, published at once
    assert f["published"] and f["half"] and (f["width"], f["height"]) == (W // 2, H // 2)
    assert f["pass_spp"] == 2 and f["denoised"] and not f["converged"]
    assert np.array_equal(img, _hand_frame(ctx, sc, cam, W // 2, H // 2, [2]))
    passes = []
    # in motion (< 350 ms since the change): 1-spp passes, the half look stands in below 4 spp
    for k in range(3):
        f, img = lv.step(now_ms=10.0 + k)
        passes.append(f["pass_spp"])
        assert f["pass_spp"] == 1 and f["spp"] == k + 1 and not f["published"], f
    # out of motion: pass sized toward 70 ms from the running ms/spp, frame due
    prev = f["ms_per_spp"]
    f, img = lv.step(now_ms=1000.0)
    passes.append(f["pass_spp"])
    want = min(max(int(np.floor(70.0 / prev)), 1), 32, 16 - 3)
    assert f["pass_spp"] == want, (f, prev)
    assert f["published"] and not f["half"] and (f["width"], f["height"]) == (W, H)
    # run to max_spp: the last frame is final and equals the hand-run pipeline
    t = 2000.0
    while not f["converged"]:
        f, img = lv.step(now_ms=t)
        passes.append(f["pass_spp"])
        t += 2000.0
    assert f["spp"] == 16 and f["published"] and not f["settled"]
    assert np.array_equal(img, _hand_frame(ctx, sc, cam, W, H, passes))
    f, _ = lv.step(now_ms=t)
    assert f["idle"] and not f["published"]
    lv.close()


def test_refine_interval_paces_frames(ctx, piece_scene):
    wl, sc, cam = piece_scene
    lv = ag.Live(ctx)
    lv.set(sc, cam, W, H, max_spp=4096, now_ms=0.0)
    lv.step(now_ms=0.0)
    for k in range(4):
        lv.step(now_ms=1.0 + k)
    f, _ = lv.step(now_ms=1000.0)
    assert f["published"]
    # live.rs:421-424: the next refinement frame is due 166 ms x sqrt(spp / 32)
    # (at least 166 ms, at most 1.5 s) after the last one
    f, _ = lv.step(now_ms=1000.0 + 100.0)
    assert not f["published"] and f["pass_spp"] >= 1
    f, _ = lv.step(now_ms=1000.0 + 1600.0)
    assert f["published"]
    lv.close()


def test_settles_before_max_spp(ctx, piece_scene):
    wl, sc, cam = piece_scene
    lv = ag.Live(ctx)
    lv.set(sc, cam, 48, 64, max_spp=65536, now_ms=0.0)
    t, f, seen = 0.0, None, []
    for _ in range(400):
        f, _ = lv.step(now_ms=t)
        t += 5000.0
        if f["relative_change"] >= 0:
            seen.append((f["spp"], f["relative_change"]))
        if f["converged"]:
            break
    assert f["converged"] and f["settled"] and f["spp"] < 65536
    assert seen and all(s >= 64 for s, _ in seen)       # live.rs:462: from 64 spp on
    assert seen[-1][1] < 0.005 and all(r >= 0.005 for _, r in seen[:-1])
    f, _ = lv.step(now_ms=t)
    assert f["idle"]
    lv.close()


def test_generation_restart_and_burst_pacing(ctx, piece_scene):
    wl, sc, cam = piece_scene
    lv = ag.Live(ctx)
    lv.set(sc, cam, W, H, max_spp=64, now_ms=0.0)
    f0, _ = lv.step(now_ms=0.0)
    for k in range(6):
        lv.step(now_ms=1000.0 + 500.0 * k)
    # a lone change (> 350 ms after the last one): the first look goes out at once
    lv.set(sc, cam, W, H, max_spp=64, now_ms=5000.0)
    f, _ = lv.step(now_ms=5000.0)
    assert f["generation"] == f0["generation"] + 1 and f["half"] and f["published"] and f["spp"] == 0
    # a change hard on its heels is a burst: its first look waits for the 83 ms gap
    lv.set(sc, cam, W, H, max_spp=64, now_ms=5010.0)
    f, _ = lv.step(now_ms=5010.0)
    assert f["pass_spp"] == 2 and not f["published"]
    f, _ = lv.step(now_ms=5020.0)
    assert f["pass_spp"] == 1 and not f["published"]
    f, img = lv.step(now_ms=5200.0)
    assert f["published"] and f["half"] and f["generation"] == f0["generation"] + 2
    lv.close()


def test_look_only_change_redevelops(ctx, piece_scene):
    wl, sc, cam = piece_scene
    lv = ag.Live(ctx)
    lv.set(sc, cam, W, H, max_spp=8, now_ms=0.0)
    t, passes = 0.0, []
    while True:
        f, img = lv.step(now_ms=t)
        if not f["half"]:
            passes.append(f["pass_spp"])
        t += 2000.0
        if f["converged"]:
            break
    lv.set_look("khronos", exposure_ev=0.5, now_ms=t)
    f2, img2 = lv.step(now_ms=t)
    assert f2["redeveloped"] and f2["published"] and f2["pass_spp"] == 0 and f2["spp"] == f["spp"]
    assert np.array_equal(img2, _hand_frame(ctx, sc, cam, W, H, passes, "khronos", 0.5))
    lv.close()


def test_live_session_on_the_server_shares_steps_with_finals(ctx, piece_scene):
    """A session attached to a continuous-batching server: its passes run in
    the server's steps beside the queued requests' chunks (the reference
    pauses the live view instead, live.rs:344-352).  Its frames equal the same
    passes run by hand, the requests progress while it refines, and each
    request's image equals its render alone."""
    from argent_amd import _abi as A
    from argent_amd.serve import Server
    wl, sc, cam = piece_scene
    ps = workloads.path_settings(8)
    reqs = [dict(scene=sc, camera=cam, width=40, height=48, spp=24, seed=3),
            dict(scene=sc, camera=cam, width=32, height=32, spp=40, seed=4,
                 material0=ag.Material.metal("rose14k", 0.07, 0.8))]
    srv = Server(ctx, 64 << 30, 1 << 16, ps)
    ids = [srv.submit(**r) for r in reqs]
    lv = ag.Live(ctx)
    lv.attach(srv)
    lv.set(sc, cam, W, H, max_spp=16, now_ms=0.0)
    f, img = lv.step(now_ms=0.0)
    assert f["published"] and f["half"]
    assert np.array_equal(img, _hand_frame(ctx, sc, cam, W // 2, H // 2, [2]))
    t, passes, progressed = 1000.0, [], False
    while not f["converged"]:
        f, img = lv.step(now_ms=t)
        passes.append(f["pass_spp"])
        progressed |= any(srv.status(i)["samples"] > 0 for i in ids)
        t += 2000.0
    assert progressed, "the requests sat out the live session's steps"
    assert f["spp"] == 16 and np.array_equal(img, _hand_frame(ctx, sc, cam, W, H, passes))
    lv.attach(None)
    for _ in range(1000):
        a, q = srv.step()
        if a == 0 and q == 0:
            break
    imgs = [srv.read(i) for i in ids]
    assert all(srv.status(i)["state"] == A.REQ_DONE for i in ids)
    srv.close()
    for r, img_mixed in zip(reqs, imgs):
        alone = Server(ctx, 64 << 30, 1 << 16, ps)
        i = alone.submit(**r)
        while alone.step() != (0, 0):
            pass
        assert np.array_equal(alone.read(i), img_mixed)
        alone.close()
    lv.close()
