"""The still at its full shape, and the batched and served workloads at their
real structure, reduced in resolution and samples:

  * the signet still and the gems still (HDRI and softboxes), 1080 x 1920,
    against the oracle on a band of rows;

  * the turntable batch: the full ~1.2 M-triangle signet, 36 turntable steps
    of 10 degrees x 4 alloys = 144 views with spp log-uniform (seeded), all in
    one schedule (ag_render_views), against 144 plain render_pass calls (one
    per view, the per-frame loop of jobs.rs:689-710) bit for bit, and two
    views against the oracle;
  * a serving trace: seeded Poisson arrivals over meshes from the torus and
    signet generators, mixed sizes, spp, alloys and studios, replayed through
    ag_server with steps between arrivals; every request bit-identical to the
    same request served alone, one checked against the oracle.
"""
import numpy as np
import pytest

import argent_amd as ag
from argent_amd import _abi as A
from argent_amd import synth, workloads
from argent_amd.serve import ALLOYS, Server, make_trace

pytestmark = pytest.mark.gpu

O = pytest.importorskip("oracle")


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-20))


def _oracle_render(desc, cam, turn, w, h, spp, seed, ps):
    desc.d.turntable_deg = turn
    osc = O.OracleScene(desc)
    sums = np.zeros((3, h, w, 3), np.float32)
    osc.render_pass(cam, w, h, sums, 0, ps, spp, seed, threads=16)
    osc.close()
    return sums


def test_signet_still_full_shape_matches_oracle_rows(ctx):
    """The headline still at its full shape -- the 1,198,332-triangle signet
    with its patina mask prepared on the device, 1080 x 1920, depth 12,
    sterling in the white rig, the bench's camera and seed -- rendered whole
    on the GPU and against the oracle on a band of 60 rows through the ring
    (8 spp keeps the oracle's share to seconds)."""
    wl = workloads.signet_still()
    piece = wl["piece"]
    w, h, seed = wl["width"], wl["height"], wl["seed"]
    ps = workloads.path_settings(wl["depth"])
    sc0 = ag.Scene(ctx, piece.scene(wl["metal"], 0.07, 0.8, wl["studio"]))
    piece.prepare_patina(sc0)
    cam, turn = piece.shot("hero", 100.0, wl["f_stop"], w / h, intersect=sc0.intersect)
    sc0.close()
    desc = piece.scene(wl["metal"], 0.07, 0.8, wl["studio"])
    assert len(piece.triangles) == 1198332 and piece.cavity is not None and piece.cavity.max() > 0.0
    desc.d.turntable_deg = turn
    sc = ag.Scene(ctx, desc)
    film = ag.Film(ctx, w, h)
    ag.render_pass(ctx, sc, cam, film, ps, 8, seed)
    gpu = film.sums()
    film.close()
    sc.close()
    r0, r1 = 930, 990
    osc = O.OracleScene(desc)
    ref = np.zeros((3, h, w, 3), np.float32)
    osc.render_pass(cam, w, h, ref, 0, ps, 8, seed, threads=16, rows=(r0, r1))
    osc.close()
    # the band crosses the ring: its guides see the metal
    assert np.mean(np.abs(ref[1][r0:r1] - ref[1][r0:r1].mean())) > 0.0
    e = rel_rmse(gpu[0][r0:r1], ref[0][r0:r1])
    ea = rel_rmse(gpu[1][r0:r1], ref[1][r0:r1])
    print(f"full-shape signet still, rows {r0}-{r1}: relative RMSE against the oracle {e:.2e} (albedo {ea:.2e})")
    assert e < 0.01 and ea < 0.01


def test_gems_still_full_shape_matches_oracle_rows(ctx):
    """The gems still at its full shape -- the signet with its 12 diamond and
    sapphire brilliants, the 2048 x 1024 HDRI importance-sampled with MIS
    against the white rig's softboxes, 1080 x 1920, depth 24 -- rendered
    whole on the GPU and against the oracle on a band of 60 rows through the
    head plate and its stones (8 spp)."""
    wl = workloads.signet_gems_still()
    piece = wl["piece"]
    w, h, seed = wl["width"], wl["height"], wl["seed"]
    ps = workloads.path_settings(wl["depth"])
    desc = piece.scene(wl["metal"], 0.07, 0.8, wl["studio"], gems=wl["gems"])
    desc.set_env_map(wl["hdri"])
    sc0 = ag.Scene(ctx, desc)
    cam, turn = piece.shot("hero", 100.0, wl["f_stop"], w / h, intersect=sc0.intersect)
    sc0.close()
    desc.d.turntable_deg = turn
    sc = ag.Scene(ctx, desc)
    film = ag.Film(ctx, w, h)
    ag.render_pass(ctx, sc, cam, film, ps, 8, seed)
    gpu = film.sums()
    film.close()
    sc.close()
    r0, r1 = 930, 990
    osc = O.OracleScene(desc)
    ref = np.zeros((3, h, w, 3), np.float32)
    osc.render_pass(cam, w, h, ref, 0, ps, 8, seed, threads=16, rows=(r0, r1))
    osc.close()
    e = rel_rmse(gpu[0][r0:r1], ref[0][r0:r1])
    ea = rel_rmse(gpu[1][r0:r1], ref[1][r0:r1])
    print(f"full-shape gems still, rows {r0}-{r1}: relative RMSE against the oracle {e:.2e} (albedo {ea:.2e})")
    assert e < 0.01 and ea < 0.01


def test_chunked_queues_equal_exact_appends(ctx):
    """A shading launch over at least 2^26 hits reserves its queue slots in
    chunks per wavefront and deals its hits in runs (dead slots in each
    chunk's tail); smaller launches append exactly.  The headline still at
    40 spp (82.9 M paths) rendered as one wave of 2^27 path slots (chunked at
    depths 0 and 1) and as waves of 2^24 (never chunked): films bit-identical,
    and the emitted ray counts equal."""
    wl = workloads.signet_still()
    piece = wl["piece"]
    w, h = wl["width"], wl["height"]
    spp = 40
    assert w * h * spp >= 1 << 26
    ps = workloads.path_settings(wl["depth"])
    sc = ag.Scene(ctx, piece.scene(wl["metal"], 0.07, 0.8, wl["studio"]))
    cam, turn = piece.shot("hero", 100.0, 0.0, w / h)
    sc.set_turntable(turn)
    keep = ctx.tuning
    films, rays = [], []
    try:
        for ppw in (1 << 27, 1 << 24):
            ctx.set_tuning(paths_per_wave=ppw, sort_by_material=1, ordered_emission=0)
            f = ag.Film(ctx, w, h)
            ag.render_pass(ctx, sc, cam, f, ps, spp, 7)
            films.append(f.sums())
            f.close()
            st = ctx.depth_stats(wl["depth"])
            rays.append((st["extension_rays"], st["shadow_rays"]))
    finally:
        ctx.set_tuning(**keep)
    sc.close()
    print(f"extension rays per depth {rays[0][0][:4]}, shadow rays {rays[0][1][:4]}")
    assert np.array_equal(films[0], films[1])
    assert rays[0] == rays[1]


def test_turntable_144_views_batched_equal_plain_passes(ctx):
    pos, tris, nrm = synth.signet(seed=2)
    piece = workloads.Piece(pos, tris, nrm)
    w, h = 54, 96
    ps = workloads.path_settings(12)
    alloys = ["sterling", "gold14k", "rose14k", "bronze"]
    mats = [ag.Material.metal(a, 0.07, 0.8) for a in alloys]
    base = ag.Scene(ctx, piece.scene("sterling", 0.07, 0.8, "white"))
    cam, turn0 = piece.shot("hero", 100.0, 0.0, w / h)
    rng = np.random.default_rng(44)
    views = []
    for step in range(36):
        for m in range(4):
            spp = max(1, int(round(np.exp(rng.uniform(np.log(64), np.log(512))) / 16)))
            views.append(dict(camera=cam, turntable_deg=turn0 + 10.0 * step, material0=m, count=spp,
                              seed=1000 + len(views), film=ag.Film(ctx, w, h)))
    print(f"turntable: {len(views)} views, {sum(v['count'] for v in views) * w * h / 1e6:.1f} M samples", flush=True)
    ag.render_views(ctx, base, views, ps, mats)
    batched = [v["film"].sums() for v in views]
    # the plain path: one look per alloy over the same device mesh, one pass per view
    looks = [ag.Scene(ctx, piece.scene(a, 0.07, 0.8, "white"), share_mesh_with=base) for a in alloys]
    for k, v in enumerate(views):
        sc = looks[v["material0"]]
        sc.set_turntable(v["turntable_deg"])
        f = ag.Film(ctx, w, h)
        ag.render_pass(ctx, sc, cam, f, ps, v["count"], v["seed"])
        assert np.array_equal(f.sums(), batched[k]), k
        f.close()
    for k in (0, 77):
        v = views[k]
        ref = _oracle_render(piece.scene(alloys[v["material0"]], 0.07, 0.8, "white"), cam, v["turntable_deg"], w, h,
                             v["count"], v["seed"], ps)
        e = rel_rmse(batched[k][0], ref[0])
        print(f"view {k}: relative RMSE against the oracle {e:.2e}")
        assert e < 0.01
    for v in views:
        v["film"].close()
    for sc in looks + [base]:
        sc.close()


SMALL = {(512, 512): (32, 32), (1080, 1920): (54, 96), (2000, 2000): (40, 40)}


def _trace_requests():
    """A seeded trace of the serving workload's shape at 1/16 to 1/50 of its
    resolutions and spp in [4, 48]."""
    trace = make_trace(6.0, 4.0, seed=5, spp_lo=4, spp_hi=48)
    for r in trace:
        r["width"], r["height"] = SMALL[(r["width"], r["height"])]
    return trace


def _pieces():
    rng = np.random.default_rng(6)
    out = []
    for k in range(4):
        if k % 2 == 0:
            nu = int(rng.integers(160, 400))
            out.append(workloads.Piece(*synth.torus_ring(seed=int(rng.integers(1 << 30)), nu=nu, nv=max(32, nu // 4))))
        else:
            s = float(rng.uniform(0.2, 0.4))
            out.append(workloads.Piece(*synth.signet(seed=int(rng.integers(1 << 30)), band_nu=int(1024 * s),
                                                     band_nv=int(256 * s), plate_nt=int(412 * s),
                                                     plate_np=int(820 * s))))
    return out


def test_serving_trace_each_request_equals_alone(ctx):
    trace = _trace_requests()
    pieces = _pieces()
    ps = workloads.path_settings(12)
    scenes, bases, cams = {}, {}, {}
    mats = {a: ag.Material.metal(a, 0.07, 0.8) for a in ALLOYS}
    reqs = []
    for r in trace:
        key = (r["mesh"] % len(pieces), r["studio"])
        if key not in scenes:
            p = pieces[key[0]]
            sc = ag.Scene(ctx, p.scene("sterling", 0.07, 0.8, key[1]), share_mesh_with=bases.get(key[0]))
            bases.setdefault(key[0], sc)
            scenes[key] = sc
        p, sc = pieces[key[0]], scenes[key]
        ck = (key[0], r["width"], r["height"], r["studio"])
        if ck not in cams:
            cams[ck] = p.shot("hero", 100.0, 0.0 if r["studio"] == "white" else 16.0, r["width"] / r["height"],
                              intersect=sc.intersect)
        cam, turn = cams[ck]
        reqs.append((r, dict(scene=sc, camera=cam, width=r["width"], height=r["height"], spp=r["spp"],
                             seed=r["seed"], turntable_deg=turn, material0=mats[r["alloy"]])))
    print(f"trace: {len(reqs)} requests, {len(scenes)} scene looks over {len(pieces)} meshes", flush=True)
    step_samples = 1 << 15
    # replay on a simulated clock: one step every 50 ms of trace time
    srv = Server(ctx, 64 << 30, step_samples, ps)
    ids, t, k = [], 0.0, 0
    images = {}
    for _ in range(100000):
        while k < len(reqs) and reqs[k][0]["t"] <= t:
            ids.append(srv.submit(**reqs[k][1]))
            k += 1
        a, q = srv.step()
        t += 0.05
        for j, i in enumerate(ids):
            if i is not None and srv.status(i)["state"] == A.REQ_DONE:
                images[j] = srv.read(i)
                srv.release(i)
                ids[j] = None
        if a == 0 and q == 0 and k == len(reqs):
            break
    srv.close()
    assert len(images) == len(reqs)
    for j, (r, q) in enumerate(reqs):
        alone = Server(ctx, 64 << 30, step_samples, ps)
        i = alone.submit(**q)
        while alone.step() != (0, 0):
            pass
        assert alone.status(i)["state"] == A.REQ_DONE
        assert np.array_equal(alone.read(i), images[j]), j
        alone.close()
    # one request against the oracle (same look, alloy and turntable)
    j = min(range(len(reqs)), key=lambda j: reqs[j][1]["width"] * reqs[j][1]["height"] * reqs[j][1]["spp"])
    r, q = reqs[j]
    p = pieces[r["mesh"] % len(pieces)]
    ref = _oracle_render(p.scene(r["alloy"], 0.07, 0.8, r["studio"]), q["camera"], q["turntable_deg"], q["width"],
                         q["height"], q["spp"], q["seed"], ps)
    e = rel_rmse(images[j], ref[0] / q["spp"])
    print(f"request {j} ({r['alloy']}, {r['studio']}, {q['width']}x{q['height']}, {q['spp']} spp): relative RMSE "
          f"against the oracle {e:.2e}; {len(reqs)} requests checked alone")
    assert e < 0.01
    base_ids = {id(sc) for sc in bases.values()}
    for sc in scenes.values():  # looks first, then the scenes that own the meshes
        if id(sc) not in base_ids:
            sc.close()
    for sc in bases.values():
        sc.close()
