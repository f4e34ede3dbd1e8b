"""Benchmark: Msamples/s of the signet still (1080 x 1920, 256 spp, depth 12,
sterling, white studio, ~1.2 M triangles) on one MI355X, with the oracle
(the C++ restatement of the reference CPU renderer) timed on the host cores
on a bounded sample of the same image.

One step = one full 256-spp render_pass of the still.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

HBM traffic of the traversal kernel: run the bench twice under
`rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` with `--copy-calibration
MiB` (a plain streaming copy of known size in the same process, whose counter
readings scale both counters to bytes), then

    python bench.py --reduce-traffic FETCH_DIR WRITE_DIR

writes profiles/k_extend_signet_traffic.json, which later runs report as
roofline.traffic.
"""
import glob
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
EXTEND_BYTES_PER_RAY = 44  # read ray 28 B + write hit 16 B (compulsory SoA traffic per extension ray)
SHADOW_BYTES_PER_RAY = 32  # read ray 28 B + write 4 B
RAYGEN_BYTES_PER_SAMPLE = 112
FINALIZE_BYTES_PER_SAMPLE = 72
SEGMENT_BYTES = 534  # extend 44 + shade 402 + shadow 32 + resolve 56


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """Threads for the CPU legs: every CPU this process may run on, at most 16
    (the host share that goes with one GPU on the measurement machine)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    return min(16, avail)


def host_info():
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = None
    return {"nproc": os.cpu_count(), "affinity": avail, "threads_used": host_threads(), "cpu_model": _cpu_model(),
            "build": "g++ -O3 -march=x86-64-v3 -ffp-contract=off (oracle/Makefile)"}


def cpu_baseline(piece_desc, cam, w, h, spp, ps, seed, budget_s):
    """The oracle on the host's cores over evenly spaced rows of the same
    image (same per-pixel seeds), until about budget_s of work.  Each call
    renders one row with the pixels dealt to every thread."""
    import oracle as O
    threads = host_threads()
    osc = O.OracleScene(piece_desc)
    sums = np.zeros((3, h, w, 3), np.float32)
    blocks = 32
    stride = h // blocks
    rows_per_block = 1
    done_samples = segs = 0
    t_total = 0.0
    order = [b * stride + stride // 2 for b in range(blocks)]
    row_ptr = {r: r for r in order}
    while t_total < budget_s:
        for r0 in order:
            r = row_ptr[r0]
            if r >= r0 + stride:
                continue
            t0 = time.perf_counter()
            st = osc.render_pass(cam, w, h, sums, 0, ps, spp, seed, threads=threads, rows=(r, r + rows_per_block))
            t_total += time.perf_counter() - t0
            done_samples += st[2]
            segs += st[0]
            row_ptr[r0] = r + rows_per_block
            if t_total >= budget_s:
                break
        if all(row_ptr[r0] >= r0 + stride for r0 in order):
            break
    rows = sum(row_ptr[r0] - r0 for r0 in order)
    msps = done_samples / t_total / 1e6
    return {
        "value": round(msps, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
        "sample": f"{rows} of {h} rows ({rows * w} pixels x {spp} spp = {done_samples / 1e6:.1f} M samples) "
                  f"spread over the image, {t_total:.1f} s",
        "per_core": round(msps / threads, 4),
        "segments_per_sample": round(segs / max(done_samples, 1), 4),
        "cpu_model": _cpu_model(),
        "host": host_info(),
    }


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


WORKLOAD_NAMES = {"signet": "signet still", "torus": "torus ring still", "gems": "signet with 12 brilliants still"}
EXTEND_KERNEL = "ag::k_extend("
FUSED_KERNEL = "ag::k_shadow_extend("
COPY_KERNEL = "neg_kernel"


def copy_calibration(mib):
    """A plain streaming kernel of known traffic (read mib MiB, write mib MiB):
    torch's elementwise negation of a float32 buffer, run twice."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(2):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    assert float(b[0]) == -1.0
    del a, b
    torch.cuda.empty_cache()


def _pmc_rows(d, counter):
    import csv
    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    rows = []
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] == counter:
                    rows.append((int(r.get("Dispatch_Id", 0) or 0), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    return [(k, v) for _, k, v in rows]


def _shipped_still_waves():
    from argent_amd import autotune
    t, src = autotune.load("still")
    return (t or {}).get("paths_per_wave", 1 << 26), src


def reduce_traffic(fetch_dir, write_dir, mib, spp):
    """HBM bytes per k_extend dispatch from two counter runs.  FETCH_SIZE and
    WRITE_SIZE read KiB per dispatch; each is scaled by known/counted bytes of
    the calibration copy in the same run."""
    out = {}
    for name, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_rows(d, name)
        cal = [v for k, v in rows if COPY_KERNEL in k]
        # the stage-timed pass comes last and alone runs every depth's
        # k_extend as a launch of its own: its dispatches follow the last
        # fused shadow/extend launch of the timed passes
        fused = [i for i, (k, _) in enumerate(rows) if k.startswith(FUSED_KERNEL)]
        ext = [v for k, v in rows[(fused[-1] + 1 if fused else 0):] if k.startswith(EXTEND_KERNEL)]
        if not cal or not ext:
            raise SystemExit(f"{name}: calibration dispatches {len(cal)}, k_extend dispatches {len(ext)}")
        known_kib = mib * 1024.0
        scale = known_kib / (sum(cal) / len(cal))
        out[name] = {"calibration_kib_read": round(sum(cal) / len(cal), 1), "calibration_kib_known": known_kib,
                     "scale": round(scale, 4), "k_extend_dispatches": len(ext),
                     "k_extend_kib_raw_mean": round(sum(ext) / len(ext), 1),
                     "k_extend_bytes_mean": round(sum(ext) / len(ext) * scale * 1024.0)}
    total = out["FETCH_SIZE"]["k_extend_bytes_mean"] + out["WRITE_SIZE"]["k_extend_bytes_mean"]
    ppw, _ = _shipped_still_waves()
    res = {"kernel": "k_extend",
           "workload": f"signet still 1080x1920 {spp} spp depth 12 with the bench's launch shape "
                       f"({ppw} path slots per wave): the stage-timed pass's k_extend dispatches",
           "bytes_per_dispatch": int(total), "fetch": out["FETCH_SIZE"], "write": out["WRITE_SIZE"]}
    path = os.path.join(HERE, "profiles", "k_extend_signet_traffic.json")
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
        fh.write("\n")
    print(json.dumps(res))


def turntable(args):
    """144 views (36 turntable steps of 10 degrees x 4 alloys) of the signet
    (patina mask stored), --tt-size (default 1080 x 1920), spp log-uniform in
    [64, 512] (seeded) times --spp-scale: rendered as one batched schedule
    (ag_render_views: every view's samples share waves) and as one
    render_pass per view; equal samples, and every view's film must come out
    bit-identical both ways, so the two schedules reach equal error and their
    times compare directly."""
    import argent_amd as ag
    from argent_amd import workloads
    wl = workloads.signet_still()
    piece = wl["piece"]
    w, h = args.tt_size if args.tt_size else (wl["width"], wl["height"])
    depth = wl["depth"]
    ps = workloads.path_settings(depth)
    ctx = ag.Context(0)
    scene = ag.Scene(ctx, piece.scene("sterling", 0.07, 0.8, "white"))
    piece.prepare_patina(scene)
    alloys = ["sterling", "gold14k", "rose14k", "bronze"]
    mats = [ag.Material.metal(a, 0.07, 0.8) for a in alloys]
    rng = np.random.default_rng(44)
    cam, turn0 = piece.shot("hero", 100.0, 0.0, w / h)
    views = []
    for step in range(36):
        for m in range(4):
            spp = int(round(np.exp(rng.uniform(np.log(64), np.log(512))) * args.spp_scale))
            views.append(dict(camera=cam, turntable_deg=turn0 + 10.0 * step, material0=m, count=max(spp, 1),
                              seed=1000 + len(views)))
    total = sum(v["count"] for v in views) * w * h
    log(f"turntable: 144 views, {total / 1e9:.2f} G samples")
    films = [ag.Film(ctx, w, h) for _ in views]
    for v, f in zip(views, films):
        v["film"] = f
    ctx.synchronize()
    t0 = time.perf_counter()
    ag.render_views(ctx, scene, views, ps, materials=mats)
    ctx.synchronize()
    t_batch = time.perf_counter() - t0
    batch = [f.sums() for f in films]
    for f in films:
        f.clear()
    t0 = time.perf_counter()
    for v in views:
        ag.render_views(ctx, scene, [v], ps, materials=mats)
    ctx.synchronize()
    t_each = time.perf_counter() - t0
    same = all(np.array_equal(b, f.sums()) for b, f in zip(batch, films))
    res = {"metric": "Msamples/s, turntable batch (144 views, one schedule) vs one pass per view, 1 GPU",
           "value": round(total / t_batch / 1e6, 2), "unit": "Msamples/s", "n_gpus": 1,
           "per_view_msamples_per_s": round(total / t_each / 1e6, 2), "batch_s": round(t_batch, 2),
           "per_view_s": round(t_each, 2), "bit_identical": bool(same), "dtype": "f32", "data": "synthetic",
           "speedup": round(t_each / t_batch, 3),
           "config": {"workload": f"signet turntable 36 x 10 deg x 4 alloys at {w}x{h}, spp log-uniform "
                                  f"64..512 x {args.spp_scale:g}, depth {depth}, white studio",
                      "views": 144, "samples": int(total), "width": w, "height": h,
                      "paths_per_wave": ctx.tuning["paths_per_wave"]}}
    for f in films:
        f.close()
    scene.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def ray_batch(args):
    """The 2^24-ray batch against the signet (origins uniform on a 30 mm
    sphere, directions toward the centre jittered up to 20 degrees), rays and
    hits resident in HBM: closest hit and any hit, and the oracle's closest
    hit on the host's cores over a bounded subset."""
    import argent_amd as ag
    from argent_amd import synth
    pos, tris, nrm = synth.signet(seed=2)
    c = (pos.min(0) + pos.max(0)) * 0.5
    n = 1 << 24
    o, d = synth.sphere_rays(n, c, seed=3)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    ctx = ag.Context(0)
    sc = ag.Scene(ctx, desc)
    rays = ag.DeviceRays(ctx, o, d)
    for _ in range(args.warmup):
        ag.intersect_device(ctx, sc, rays)
    ms = [ag.intersect_device(ctx, sc, rays) for _ in range(max(args.steps, 1))]
    ms_any = [ag.occluded_device(ctx, sc, rays) for _ in range(max(args.steps, 1))]
    t_ms = sum(ms) / len(ms)
    achieved = n * 44 / (t_ms * 1e-3) / 1e9
    res = {"metric": "Grays/s closest hit, 2^24 rays vs the 1.2 M-triangle signet, rays resident in HBM, 1 GPU",
           "value": round(n / (t_ms * 1e-3) / 1e9, 3), "unit": "Grays/s", "n_gpus": 1, "steps": len(ms),
           "warmup": args.warmup, "ms_per_step": round(t_ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "any_hit_grays_per_s": round(n / (sum(ms_any) / len(ms_any) * 1e-3) / 1e9, 3),
           "config": {"workload": "2^24 rays, origins on a 30 mm sphere, toward the centre +- 20 degrees",
                      "rays": n, "triangles": int(len(tris))},
           "roofline": {"kernel": "k_rays_intersect", "bound": "hbm", "achieved": round(achieved, 1),
                        "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 5),
                        "traffic": None, "bytes_per_unit": 44, "unit_of_work": "ray (28 B in, 16 B hit out)"}}
    rays.close()
    sc.close()
    if not args.no_cpu:
        import oracle as O
        osc = O.OracleScene(desc)
        threads = host_threads()
        # The whole batch, passed over again until the CPU has worked about
        # --cpu-seconds (one pass takes well under a second on 16 cores).
        done, passes, t0 = 0, 0, time.perf_counter()
        while True:
            osc.intersect(o, d, mode=O.MODE_ROBUST, threads=threads)
            done += n
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= args.cpu_seconds or passes >= 64:
                break
        res["cpu_baseline"] = {"value": round(done / dt / 1e9, 4), "unit": "Grays/s", "cores": threads,
                               "kind": "port", "sample": f"the whole 2^24-ray batch, {passes} passes, {dt:.1f} s",
                               "cpu_model": _cpu_model()}
        res["gpu_over_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
        # The secondary "BVH-equivalent" figure: what the reference's own 4-wide tree
        # (bvh.rs, restated by the oracle) touches per ray -- Node4 visits x 128 B plus
        # triangle tests x 48 B -- at the GPU's ray rate.  Served from L2/MALL, not HBM.
        m = 1 << 16
        nv, nt = osc.intersect_stats(o[:m], d[:m])
        per_ray = (nv * 128 + nt * 48) / m
        res["bvh_equivalent"] = {"bytes_per_ray": round(per_ray, 1), "node4_visits_per_ray": round(nv / m, 3),
                                 "tri_tests_per_ray": round(nt / m, 3),
                                 "achieved_GBs": round(per_ray * res["value"], 1),
                                 "sample": "the first 2^16 rays on the reference's 4-wide tree"}
    ctx.close()
    print(json.dumps(res), flush=True)


def cavity_ao(args):
    """Scene prep: cavity ambient occlusion of the signet (600,804 vertices x
    48 rays, argent-mesh cavity.rs:95-128) with the reach the reference sizes
    by the piece (shot.rs:132-137: 0.09 x the half-diagonal, clamped to
    [0.4, 2.5] mm), on the GPU and by the oracle on the host's cores."""
    import argent_amd as ag
    from argent_amd import synth
    pos, tris, nrm = synth.signet(seed=2)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    radius = max(float(np.linalg.norm((pos.max(0) - pos.min(0)) * 0.5)), 1e-3)
    dist = float(np.clip(radius * 0.09, 0.4, 2.5))
    samples = 48
    ctx = ag.Context(0)
    sc = ag.Scene(ctx, desc)
    for _ in range(args.warmup):
        sc.ambient_occlusion(samples, dist)
    ms = []
    for _ in range(max(args.steps, 1)):
        sc.ambient_occlusion(samples, dist)
        ms.append(ctx.last_query_ms())
    t_ms = sum(ms) / len(ms)
    rays = len(pos) * samples
    # the whole patina mask (cavity.rs:52-80) with the settings sized by the piece
    from argent_amd import workloads
    pset = sc.patina_settings()
    sc.patina_mask(pset)
    pm_wall, pm_dev = [], []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        mask = sc.patina_mask(pset)
        pm_wall.append((time.perf_counter() - t0) * 1e3)
        pm_dev.append(ctx.last_query_ms())
    res = {"metric": "Mrays/s, cavity ambient occlusion (48 rays per vertex) of the 1.2 M-triangle signet, 1 GPU",
           "value": round(rays / (t_ms * 1e-3) / 1e6, 1), "unit": "Mrays/s", "n_gpus": 1, "steps": len(ms),
           "warmup": args.warmup, "ms_per_step": round(t_ms, 3), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"signet AO, {len(pos)} vertices x {samples} rays, reach {dist:.3f} mm",
                      "vertices": int(len(pos)), "rays": int(rays), "distance_mm": round(dist, 4)},
           "patina_mask": {"settings": {k: (round(v, 5) if isinstance(v, float) else v) for k, v in pset.items()},
                           "device_ms": round(float(np.median(pm_dev)), 2),
                           "wall_ms": round(float(np.median(pm_wall)), 2), "mean": round(float(mask.mean()), 4),
                           "timing": "device_ms: HIP events around AO, concavity, blend and smoothing; wall_ms: the "
                                     "call with the mask copied to host memory"}}
    sc.close()
    if not args.no_cpu:
        import oracle as O
        threads = host_threads()
        osc = O.OracleScene(desc)
        passes, t0 = 0, time.perf_counter()
        while True:  # the whole mesh, again until about --cpu-seconds of work
            osc.ambient_occlusion(len(pos), samples, dist, threads=threads)
            passes += 1
            dt = time.perf_counter() - t0
            if dt >= args.cpu_seconds or passes >= 64:
                break
        res["cpu_baseline"] = {"value": round(passes * rays / dt / 1e6, 2), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "sample": f"the whole mesh, {passes} passes, {dt:.1f} s",
                               "cpu_model": _cpu_model()}
        res["gpu_over_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
        t0 = time.perf_counter()
        ao = osc.ambient_occlusion(len(pos), pset["ao_samples"], pset["ao_distance"], threads=threads)
        t1 = time.perf_counter()
        ref = O.patina_mask(pos, nrm, tris, ao=ao, **{k: pset[k] for k in ("curvature_radius", "concavity_weight",
                                                                            "smooth_iterations")})
        t2 = time.perf_counter()
        res["patina_mask"]["cpu_baseline"] = {
            "value": round((t2 - t0) * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": f"the whole mask once: AO {(t1 - t0) * 1e3:.0f} ms on {threads} threads, concavity + blend + "
                      f"smoothing {(t2 - t1) * 1e3:.0f} ms on one", "cpu_model": _cpu_model()}
        res["patina_mask"]["max_abs_diff_vs_oracle"] = float(np.abs(ref - mask).max())
    ctx.close()
    print(json.dumps(res), flush=True)


def preview_develop(args):
    """The live preview's development (live.rs:429-496) of a 1080x1920 signet
    film (16 spp): à-trous with the albedo and normal guides, then develop
    (AgX Punchy, dithered 8-bit), on the GPU from the device film to host
    bytes, and by the oracle on the host (one thread) from the same film."""
    import argent_amd as ag
    from argent_amd import workloads
    wl = workloads.signet_still()
    w, h = wl["width"], wl["height"]
    ctx = ag.Context(0)
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])
    cam, turn = wl["piece"].shot("hero", 100.0, wl["f_stop"], w / h)
    desc.d.turntable_deg = turn
    sc = ag.Scene(ctx, desc)
    wl["piece"].prepare_patina(sc)
    film = ag.Film(ctx, w, h)
    ag.render_pass(ctx, sc, cam, film, workloads.path_settings(wl["depth"]), 16, wl["seed"])
    for _ in range(args.warmup):
        film.develop("agx_punchy", denoise=True)
    t_dn, t_dev = [], []
    for _ in range(max(args.steps, 1)):
        t0 = time.perf_counter()
        film.develop("agx_punchy", denoise=True)
        t_dn.append((time.perf_counter() - t0) * 1e3)
        t0 = time.perf_counter()
        film.develop("agx_punchy")
        t_dev.append((time.perf_counter() - t0) * 1e3)
    ms_dn, ms_dev = sorted(t_dn)[len(t_dn) // 2], sorted(t_dev)[len(t_dev) // 2]
    mp = w * h / 1e6
    res = {"metric": "ms per preview frame, à-trous + develop of a 1080x1920 film to 8-bit, 1 GPU",
           "value": round(ms_dn, 2), "unit": "ms", "n_gpus": 1, "steps": len(t_dn), "warmup": args.warmup,
           "ms_per_step": round(ms_dn, 2), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic", "develop_only_ms": round(ms_dev, 2),
           "mpix_per_s": round(mp / (ms_dn * 1e-3), 1),
           "config": {"workload": "signet still 1080x1920 at 16 spp, à-trous (5 passes, albedo + normal guides), "
                                  "AgX Punchy, TPDF dither, 8-bit; film on the device to bytes on the host",
                      "width": w, "height": h},
           "timing": "wall clock per call, median, including the 6.2 MB copy to the host"}
    if not args.no_cpu:
        import oracle as O
        color, albedo, normal = film.color(), film.albedo(), film.normal()
        t0 = time.perf_counter()
        O.develop(O.atrous(color, albedo, normal), "agx_punchy")
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                               "sample": "the same frame once", "cpu_model": _cpu_model()}
        res["cpu_over_gpu"] = round(dt * 1e3 / ms_dn, 1)
    film.close()
    sc.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def mesh_ingest(args):
    """Mesh ingest of a binary STL (io.rs:87-126: parse io/stl.rs:34-62, drop
    invalid triangles, weld weld.rs:38-62, crease normals normals.rs:47-220
    at 50 degrees) for the 1.2 M-triangle signet
    written as an STL triangle soup (3.6 M vertices, 59.9 MB), the file's
    bytes already in HBM; the oracle (one thread) loads the same bytes on the
    host.  A step is one ag_mesh_load_stl call, allocations and the few
    host reads of counts included."""
    import argent_amd as ag
    from argent_amd import synth
    pos, tris, _ = synth.signet()
    data = synth.to_binary_stl(pos, tris)
    nt = len(tris)
    ctx = ag.Context(0)
    buf = ag.DeviceBytes(ctx, data)
    for _ in range(args.warmup):
        m, r = ag.load_stl(ctx, len(data), device_ptr=buf.ptr)
        m.close()
    steps = max(args.steps, 1)
    ctx.synchronize()
    parse_ms, weld_ms, normals_ms = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        m, r = ag.load_stl(ctx, len(data), device_ptr=buf.ptr)
        t1 = time.perf_counter()
        nr = m.normals(50.0)
        normals_ms.append((time.perf_counter() - t1) * 1e3)
        parse_ms.append(r["parse_ms"])
        weld_ms.append(r["weld_ms"])
        nv_out = m.counts[0]
        m.close()
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / steps
    w = r["weld"]
    # algorithmic bytes of the whole chain: the file read once (50 B per
    # record), the welded mesh written once (12 B per vertex and triangle,
    # 12 B per vertex normal)
    alg = len(data) + 24 * nv_out + 12 * (nt - w["dropped_triangles"])
    res = {"metric": "binary STL ingest (parse + weld + crease normals) throughput, 1 GPU",
           "value": round(nt / dt / 1e6, 2),
           "unit": "Mtriangles/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
           "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32/u32", "data": "synthetic",
           "config": {"workload": "signet ring (1,198,332 triangles) as a binary STL soup, 59.9 MB in HBM; "
                                  "weld tolerance 1e-6 x bbox diagonal, unit scale 1, crease angle 50 degrees",
                      "triangles": nt, "stl_bytes": len(data)},
           "report": {"output_vertices": w["output_vertices"], "exact_merges": w["exact_merges"],
                      "tolerance_merges": w["tolerance_merges"], "dropped_triangles": w["dropped_triangles"],
                      "output_vertices_after_normals": nv_out},
           "stage_ms": {"parse": round(sorted(parse_ms)[steps // 2], 3), "weld": round(sorted(weld_ms)[steps // 2], 3),
                        "normals": round(sorted(normals_ms)[steps // 2], 3)},
           "normals_report": nr,
           "hbm_alg_gbps": round(alg / dt / 1e9, 1),
           "timing": "wall clock over the timed steps; stage_ms are device event pairs (median) around each stage, "
                     "host waits for counts included"}
    if not args.no_cpu:
        import oracle as O
        n, t1 = 0, time.perf_counter()
        while True:
            wp, wt, _ = O.load_stl(data)
            O.normals(wp, wt, 50.0)
            n += 1
            el = time.perf_counter() - t1
            if el >= args.cpu_seconds:
                break
        res["cpu_baseline"] = {"value": round(n * nt / el / 1e6, 3), "unit": "Mtriangles/s", "cores": 1,
                               "kind": "port", "sample": f"{n} whole loads of the same file", "cpu_model": _cpu_model()}
        res["gpu_over_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
    buf.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def live_view(args):
    """The live view's worker (live.rs:637-830) on the signet at the preview
    size (540x960, params.rs:346-351), max_spp 256, denoised frames: each of
    --steps generations starts with a change (ag_live_set) and steps on the
    wall clock until its final frame; reported are the medians of time to the
    first look (half resolution, 2 spp, developed 8-bit on the host), to the
    first full-resolution frame and to the final frame.  The oracle's first
    look (render on the host cores, à-trous and develop on one) is the CPU
    baseline."""
    import argent_amd as ag
    from argent_amd import workloads
    wl = workloads.signet_still()
    w, h, max_spp = 540, 960, 256
    ctx = ag.Context(0)
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])
    cam, turn = wl["piece"].shot("hero", 100.0, wl["f_stop"], w / h)
    desc.d.turntable_deg = turn
    sc = ag.Scene(ctx, desc)
    wl["piece"].prepare_patina(sc)
    desc = wl["piece"].scene(wl["metal"], 0.07, 0.8, wl["studio"])  # with the mask, for the oracle
    desc.d.turntable_deg = turn
    lv = ag.Live(ctx)

    clock = {"offset": 0.0}

    def generation():
        # each generation is a lone change: the session clock jumps 10 s ahead
        # (a change within 350 ms of the previous one is a burst, whose first
        # look is paced, live.rs:709, :750)
        clock["offset"] += 1e4
        t0 = time.monotonic() * 1e3
        lv.set(sc, cam, w, h, max_spp=max_spp, now_ms=t0 + clock["offset"])
        g = {"first_look_ms": None, "first_full_ms": None, "final_ms": None, "frames": 0, "passes": 0}
        while True:
            f, _ = lv.step(now_ms=time.monotonic() * 1e3 + clock["offset"])
            now = time.monotonic() * 1e3 - t0
            g["passes"] += 1 if f["pass_spp"] else 0
            if f["published"]:
                g["frames"] += 1
                if f["half"] and g["first_look_ms"] is None:
                    g["first_look_ms"] = now
                if not f["half"] and g["first_full_ms"] is None:
                    g["first_full_ms"] = now
            if f["converged"]:
                g.update(final_ms=now, spp=f["spp"], settled=bool(f["settled"]), last_change=f["relative_change"],
                         ms_per_spp=f["ms_per_spp"])
                return g
            if f["idle"]:
                raise RuntimeError("live view went idle before its final frame")

    for _ in range(args.warmup):
        generation()
    gens = [generation() for _ in range(max(args.steps, 1))]
    med = lambda k: round(float(np.median([g[k] for g in gens])), 2)
    res = {"metric": "live view time to first look (half resolution, 2 spp, denoised, 8-bit), 1 GPU",
           "value": med("first_look_ms"), "unit": "ms", "n_gpus": 1, "steps": len(gens), "warmup": args.warmup,
           "ms_per_step": med("final_ms"), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "signet ring, sterling, white rig, preview 540x960, max_spp 256, seed 1, "
                                  "PathSettings::default(); a generation = one change stepped to its final frame",
                      "width": w, "height": h, "max_spp": max_spp},
           "first_full_ms": med("first_full_ms"), "final_ms": med("final_ms"), "frames": med("frames"),
           "passes": med("passes"), "final_spp": med("spp"), "settled": [g["settled"] for g in gens],
           "ms_per_spp": med("ms_per_spp"),
           "timing": "wall clock from ag_live_set to the end of the step that published the frame, host bytes in hand"}
    if not args.no_cpu:
        import oracle as O
        threads = host_threads()
        osc = O.OracleScene(desc)
        hw, hh = w // 2, h // 2
        sums = np.zeros((3, hh, hw, 3), np.float32)
        t0 = time.perf_counter()
        osc.render_pass(cam, hw, hh, sums, 0, workloads.path_settings(32), 2, 1, threads=threads)
        t1 = time.perf_counter()
        k = np.float32(1.0 / 2.0)
        O.develop(O.atrous(sums[0] * k, sums[1] * k, sums[2] * k), "agx_punchy")
        t2 = time.perf_counter()
        res["cpu_baseline"] = {"value": round((t2 - t0) * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
                               "sample": f"the first look once: render {hw}x{hh} at 2 spp on {threads} threads "
                                         f"({(t1 - t0) * 1e3:.0f} ms), à-trous + develop on one "
                                         f"({(t2 - t1) * 1e3:.0f} ms)", "cpu_model": _cpu_model()}
        res["cpu_over_gpu"] = round(res["cpu_baseline"]["value"] / res["value"], 1)
    lv.close()
    sc.close()
    ctx.close()
    print(json.dumps(res), flush=True)


def bvh_build(args):
    """Scene preparation's BVH for the 1.2 M-triangle signet: the host binned
    SAH (ag_scene_create) against the device build (ag_scene_create_ex with
    AG_SCENE_GPU_BVH), each timed as the whole scene creation from host
    arrays (uploads and tables included), and the 2^24-ray closest-hit batch
    on each tree.  The reference builds its 4-wide binned SAH on the host
    (bvh.rs:117-146) inside a ~3 s prepare (API.md:19)."""
    import argent_amd as ag
    from argent_amd import synth
    pos, tris, nrm = synth.signet(seed=2)
    desc = ag.SceneDesc(pos, tris, normals=nrm)
    desc.add_material(ag.Material.metal("sterling"))
    ctx = ag.Context(0)
    c = (pos.min(0) + pos.max(0)) * 0.5
    o, d = synth.sphere_rays(1 << 24, c, seed=3)
    rays = ag.DeviceRays(ctx, o, d)
    res = {}
    for name, gpu in (("host_sah", False), ("gpu", True)):
        reps = max(args.steps, 1) if gpu else 1
        for _ in range(args.warmup if gpu else 0):
            ag.Scene(ctx, desc, gpu_bvh=gpu).close()
        times = []
        for _ in range(reps):
            t0 = time.perf_counter()
            sc = ag.Scene(ctx, desc, gpu_bvh=gpu)
            times.append((time.perf_counter() - t0) * 1e3)
            st = sc.stats()
            if len(times) < reps:
                sc.close()
        ms = min(ag.intersect_device(ctx, sc, rays) for _ in range(3))
        res[name] = {"scene_ms": round(float(np.median(times)), 1), "nodes": st["nodes"], "depth": st["depth"],
                     "bytes": st["bytes"], "rays_grays_per_s": round((1 << 24) / (ms * 1e-3) / 1e9, 2)}
        sc.close()
    rays.close()
    ctx.close()
    g, h = res["gpu"], res["host_sah"]
    out = {"metric": "scene creation with its BVH, 1.2 M-triangle signet, 1 GPU", "value": g["scene_ms"],
           "unit": "ms", "n_gpus": 1, "steps": max(args.steps, 1), "warmup": args.warmup,
           "ms_per_step": g["scene_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "signet ring (1,198,332 triangles), positions/triangles/normals from host arrays; "
                                  "rays: 2^24 closest-hit batch", "triangles": int(len(tris))},
           "gpu_bvh": g, "host_sah": h, "build_speedup": round(h["scene_ms"] / g["scene_ms"], 1),
           "traversal_ratio": round(g["rays_grays_per_s"] / h["rays_grays_per_s"], 3),
           "cpu_baseline": {"value": h["scene_ms"], "unit": "ms", "cores": host_threads(),
                            "kind": "port", "sample": "the host binned-SAH build of this scene once",
                            "cpu_model": _cpu_model()}}
    print(json.dumps(out), flush=True)


def new_mesh(args):
    """Serving a new STL: from the file's bytes in host memory to the first
    developed preview (the live view's half-resolution first look at
    540x960) -- upload, parse and weld (ag_mesh_load_stl), crease normals
    (ag_mesh_normals), framing from the device bounds and hull (ag_mesh_frame_sample), scene
    with its BVH built on the device (ag_scene_create_from_mesh), then the
    first live step.  The reference spends ~3 s in prepare for a 1.5 M-triangle
    mesh (API.md:19) before its first pass; the oracle's chain (load, normals,
    host BVH, first look) is timed as the CPU baseline."""
    import argent_amd as ag
    from argent_amd import synth, workloads
    pos, tris, _ = synth.signet()
    data = synth.to_binary_stl(pos, tris)
    w, h = 540, 960
    ctx = ag.Context(0)
    clock = {"offset": 0.0}

    def once():
        t0 = time.perf_counter()
        m, _ = ag.load_stl(ctx, data)
        t1 = time.perf_counter()
        m.normals(50.0)
        t2 = time.perf_counter()
        # framing needs the bounds and a 4000-point hull: taken on the device,
        # the scene's arrays stay there
        hull, lo, hi = m.frame_sample()
        piece = workloads.Piece.from_frame(lo, hi, hull)
        desc = piece.scene("sterling", 0.07, 0.8, "white")
        cam, turn = piece.shot("hero", 100.0, 0.0, w / h)
        desc.d.turntable_deg = turn
        t3 = time.perf_counter()
        sc = ag.Scene.from_mesh(ctx, m, desc, gpu_bvh=True)
        t4 = time.perf_counter()
        sc.patina_mask(sc.patina_settings(), store=True)  # Prepared::build's patina step (shot.rs:133-143)
        t4p = time.perf_counter()
        lv = ag.Live(ctx)
        clock["offset"] += 1e4
        lv.set(sc, cam, w, h, max_spp=256, now_ms=time.monotonic() * 1e3 + clock["offset"])
        f, img = lv.step(now_ms=time.monotonic() * 1e3 + clock["offset"])
        t5 = time.perf_counter()
        assert f["published"] and f["half"]
        lv.close()
        sc.close()
        m.close()
        ms = lambda a, b: (b - a) * 1e3
        return {"total_ms": ms(t0, t5), "load_ms": ms(t0, t1), "normals_ms": ms(t1, t2), "framing_ms": ms(t2, t3),
                "scene_ms": ms(t3, t4), "patina_ms": ms(t4, t4p), "first_look_ms": ms(t4p, t5)}

    for _ in range(args.warmup):
        once()
    runs = [once() for _ in range(max(args.steps, 1))]
    med = {k: round(float(np.median([r[k] for r in runs])), 2) for k in runs[0]}
    res = {"metric": "new STL to first preview (upload, ingest, normals, scene + BVH, patina mask, first look), 1 GPU",
           "value": med["total_ms"], "unit": "ms", "n_gpus": 1, "steps": len(runs), "warmup": args.warmup,
           "ms_per_step": med["total_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "signet ring as a binary STL (1,198,332 triangles, 59.9 MB in host memory), "
                                  "sterling, white rig, preview 540x960, first look at 270x480 x 2 spp",
                      "stl_bytes": len(data)},
           "stages_ms": med}
    if not args.no_cpu:
        import oracle as O
        threads = host_threads()
        t0 = time.perf_counter()
        wp, wt, _ = O.load_stl(data)
        wq, wn, ws, _ = O.normals(wp, wt, 50.0)
        piece = workloads.Piece(wq, ws, wn)
        desc = piece.scene("sterling", 0.07, 0.8, "white")
        cam, turn = piece.shot("hero", 100.0, 0.0, w / h)
        desc.d.turntable_deg = turn
        t1 = time.perf_counter()
        osc = O.OracleScene(desc)
        t2 = time.perf_counter()
        ps_ = piece.patina_settings()
        ao = osc.ambient_occlusion(len(wq), ps_["ao_samples"], ps_["ao_distance"], threads=threads)
        piece.cavity = O.patina_mask(wq, wn, ws, ao=ao, **{k: ps_[k] for k in ("curvature_radius", "concavity_weight",
                                                                            "smooth_iterations")})
        t2p = time.perf_counter()
        # the oracle's scene takes its cavity at creation: the render scene
        # with the mask is built again, outside the timed chain
        osc.close()
        desc = piece.scene("sterling", 0.07, 0.8, "white")
        desc.d.turntable_deg = turn
        osc = O.OracleScene(desc)
        t2q = time.perf_counter()
        hw, hh = w // 2, h // 2
        sums = np.zeros((3, hh, hw, 3), np.float32)
        osc.render_pass(cam, hw, hh, sums, 0, workloads.path_settings(32), 2, 1, threads=threads)
        k = np.float32(0.5)
        O.develop(O.atrous(sums[0] * k, sums[1] * k, sums[2] * k), "agx_punchy")
        t3 = time.perf_counter()
        total = (t3 - t0) - (t2q - t2p)
        res["cpu_baseline"] = {"value": round(total * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
                               "sample": f"the chain once: load + normals {(t1 - t0) * 1e3:.0f} ms (1 thread), "
                                         f"BVH {(t2 - t1) * 1e3:.0f} ms, patina mask {(t2p - t2) * 1e3:.0f} ms "
                                         f"(AO on {threads} threads), first look {(t3 - t2q) * 1e3:.0f} ms "
                                         f"({threads} threads render, 1 thread post)", "cpu_model": _cpu_model()}
        res["cpu_over_gpu"] = round(res["cpu_baseline"]["value"] / res["value"], 1)
    ctx.close()
    print(json.dumps(res), flush=True)


def graph_launch(args):
    """Host launch cost: render passes of the signet (patina mask stored) with
    every kernel launched on its own against the same pass replayed from a
    captured HIP graph (tuning launch_graph), wall clock per synchronous pass,
    median of --steps x 10 passes after the capture.  Shapes: the live view's
    first look (270x480, 2 spp, depth 32), a served preview step (135x240,
    1 spp, depth 12), a 64x64 pass at 1 spp, and the 256-spp still (one wave
    of 2^29 paths, shipped tuning)."""
    import argent_amd as ag
    from argent_amd import autotune, workloads
    wl = workloads.signet_still()
    piece = wl["piece"]
    ctx = ag.Context(0)
    sc = ag.Scene(ctx, piece.scene(wl["metal"], 0.07, 0.8, wl["studio"]))
    piece.prepare_patina(sc)
    library = ctx.tuning
    shapes = [("first look 270x480 2spp depth 32", 270, 480, 2, ag.PathSettings(), 10),
              ("preview step 135x240 1spp depth 12", 135, 240, 1, workloads.path_settings(12), 10),
              ("64x64 1spp depth 12", 64, 64, 1, workloads.path_settings(12), 10),
              ("still 1080x1920 256spp depth 12", 1080, 1920, 256, workloads.path_settings(12), 1)]
    if args.graph_shape is not None:
        shapes = [shapes[args.graph_shape]]
    out = []
    for name, w, h, spp, ps, reps in shapes:
        ctx.set_tuning(**library)
        if spp == 256:
            chosen, _ = autotune.load("still")
            autotune.apply(ctx, chosen)
        cam, turn = piece.shot("hero", 100.0, 0.0, w / h)
        sc.set_turntable(turn)
        film = ag.Film(ctx, w, h)
        row = {"shape": name}
        for on in (0, 1):
            ctx.set_tuning(launch_graph=on)
            for k in range(2):  # warm-up; the first graph pass captures
                ag.render_pass(ctx, sc, cam, film, ps, spp, k)
            ms = []
            for k in range(max(args.steps, 1) * reps):
                t0 = time.perf_counter()
                ag.render_pass(ctx, sc, cam, film, ps, spp, 10 + k)
                ms.append((time.perf_counter() - t0) * 1e3)
            row["graph_ms" if on else "eager_ms"] = round(float(np.median(ms)), 3)
        row["eager_over_graph"] = round(row["eager_ms"] / row["graph_ms"], 3)
        row["launches_per_pass"] = 2 * ps.max_depth + 5
        log(f"{name}: {row}")
        out.append(row)
        film.close()
    caps, reps_ = ctx.graph_stats()
    sc.close()
    ctx.close()
    res = {"metric": "ms per render pass, kernels launched one by one vs one captured HIP graph, 1 GPU",
           "value": out[0]["graph_ms"], "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 2,
           "ms_per_step": out[0]["graph_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "signet ring (patina mask stored), sterling, white rig; shapes as listed"},
           "shapes": out, "graph_captures": caps, "graph_replays": reps_,
           "timing": "wall clock around ag_render_pass (synchronous: it waits for the pass's ray tallies)"}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the still's launch shape first (and persist it) instead of the stored choice")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--spp", type=int, default=None, help="samples per pixel (default: the workload's)")
    ap.add_argument("--workload", choices=["signet", "torus", "gems"], default="signet",
                    help="signet: the still above (default); torus: the 131 K-triangle ring, 512x512, 64 spp, "
                         "noir, the reference's CPU-runnable case; gems: the signet with 12 brilliants, an HDRI "
                         "and the white rig's softboxes, 512 spp, depth 24")
    ap.add_argument("--rays", action="store_true", help="the 2^24-ray closest/any-hit batch instead of a still")
    ap.add_argument("--ao", action="store_true", help="cavity ambient occlusion of the signet (scene prep) instead")
    ap.add_argument("--develop", action="store_true", help="preview development (à-trous + tone map) instead")
    ap.add_argument("--ingest", action="store_true", help="binary STL ingest (parse + weld) of the signet instead")
    ap.add_argument("--live", action="store_true", help="the live view's first look / refinement loop instead")
    ap.add_argument("--bvh", action="store_true", help="scene creation with the device BVH build vs the host SAH")
    ap.add_argument("--new-mesh", action="store_true", help="a new STL's bytes to its first preview (serving)")
    ap.add_argument("--graphs", action="store_true", help="render passes launched one by one vs a captured HIP graph")
    ap.add_argument("--graph-shape", type=int, default=None, choices=range(4),
                    help="--graphs: only this shape (0 first look, 1 preview step, 2 64x64, 3 still)")
    ap.add_argument("--serve-seconds", type=float, default=120.0,
                    help="length of the served trace (Poisson arrivals at --serve-rate); 0 skips serving")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-step-log2", type=int, default=26, help="serving: log2 of the samples per server step")
    ap.add_argument("--serve-wall", type=float, default=135.0, help="stop serving after this many seconds")
    ap.add_argument("--plain-seconds", type=float, default=25.0,
                    help="render time of the serial plain leg on the same trace (0 skips it)")
    ap.add_argument("--turntable", action="store_true", help="the 144-view turntable batch instead of the still")
    ap.add_argument("--spp-scale", type=float, default=1.0, help="turntable: scale every view's spp")
    ap.add_argument("--tt-size", type=int, nargs=2, metavar=("W", "H"), default=None,
                    help="turntable: view size (default 1080 1920)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="run a streaming copy of this size first (counter calibration)")
    ap.add_argument("--reduce-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="turn two rocprofv3 --pmc runs into profiles/k_extend_signet_traffic.json")
    args = ap.parse_args()
    if args.reduce_traffic:
        reduce_traffic(*args.reduce_traffic, args.copy_calibration or 1024, args.spp or 256)
        return
    if args.turntable:
        turntable(args)
        return
    if args.rays:
        ray_batch(args)
        return
    if args.ao:
        cavity_ao(args)
        return
    if args.develop:
        preview_develop(args)
        return
    if args.ingest:
        mesh_ingest(args)
        return
    if args.live:
        live_view(args)
        return
    if args.bvh:
        bvh_build(args)
        return
    if args.new_mesh:
        new_mesh(args)
        return
    if args.graphs:
        graph_launch(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the render core runs on a single MI355X; --gpus {args.gpus} is not "
              f"supported; a still shards by image tiles or sample ranges, a server by requests", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)

    import argent_amd as ag
    from argent_amd import workloads

    t0 = time.time()
    wl = {"signet": workloads.signet_still, "torus": workloads.torus_still,
          "gems": workloads.signet_gems_still}[args.workload]()
    piece = wl["piece"]
    w, h, depth, seed = wl["width"], wl["height"], wl["depth"], wl["seed"]
    spp = args.spp or wl["spp"]
    desc = piece.scene(wl["metal"], 0.07, 0.8, wl["studio"], gems=wl.get("gems", ()))
    if "hdri" in wl:
        desc.set_env_map(wl["hdri"])
    ps = workloads.path_settings(depth)
    ctx = ag.Context(0)
    scene = ag.Scene(ctx, desc)
    # Prepared::build's patina step (shot.rs:133-143), then the same look
    # described with the mask for the oracle
    t_p = time.perf_counter()
    piece.prepare_patina(scene)
    prep = {"patina_ms": round((time.perf_counter() - t_p) * 1e3, 1), "patina_device_ms": round(ctx.last_query_ms(), 1),
            "mask_mean": round(float(piece.cavity.mean()), 4)}
    desc = piece.scene(wl["metal"], 0.07, 0.8, wl["studio"], gems=wl.get("gems", ()))
    if "hdri" in wl:
        desc.set_env_map(wl["hdri"])
    cam, turn = piece.shot("hero", 100.0, wl["f_stop"], w / h, intersect=scene.intersect)  # autofocus at f/16
    scene.set_turntable(turn)
    desc.d.turntable_deg = turn
    from argent_amd import autotune
    tune_info = None
    library_tuning = ctx.tuning
    if args.autotune:
        log(f"autotune: trials at {spp} spp ...")
        tune_info = autotune.tune_render(ctx, scene, cam, w, h, spp, ps, repeats=1)
        autotune.save("still", tune_info)
        tuning_source = f"autotuned in this run (render passes at {spp} spp)"
        log(f"autotune: {tune_info}")
    else:
        chosen, src = autotune.load("still")
        autotune.apply(ctx, chosen)
        tuning_source = src.split(":", 1)[0] + " autotuner choice" if src else "library defaults"
    stats = scene.stats()
    log(f"scene: {len(piece.triangles)} triangles, BVH {stats['nodes']} nodes depth {stats['depth']}, "
        f"{stats['bytes'] / 2**20:.0f} MiB on device, built in {time.time() - t0:.1f} s")
    film = ag.Film(ctx, w, h)
    # the timed steps run as in production (shadow rays of one depth and
    # extension rays of the next in one launch); one more pass afterwards,
    # with per-stage event timing (separate launches), gives the stage split
    # and the k_extend launch time of the roofline
    ctx.enable_timing(False)
    for i in range(args.warmup):
        film.clear()
        ag.render_pass(ctx, scene, cam, film, ps, spp, seed)
        log(f"warmup {i}: {ctx.stats()['ms_total']:.0f} ms")
    ctx.synchronize()
    totals = dict(extension_rays=0, shadow_rays=0, samples=0)
    t_start = time.perf_counter()
    for i in range(args.steps):
        film.clear()
        ag.render_pass(ctx, scene, cam, film, ps, spp, seed)
        s = ctx.stats()
        for k in totals:
            totals[k] += s[k]
        log(f"step {i}: {s['ms_total']:.0f} ms, {s['extension_rays'] / s['samples']:.3f} segments/sample")
    ctx.synchronize()
    elapsed = time.perf_counter() - t_start
    samples = w * h * spp * args.steps
    msps = samples / elapsed / 1e6
    img = film.color()
    assert np.all(np.isfinite(img)) and img.mean() > 0, "render produced no light"

    ctx.enable_timing(True)
    film.clear()
    ag.render_pass(ctx, scene, cam, film, ps, spp, seed)
    staged = ctx.stats()
    ds = ctx.depth_stats(depth)  # the stage-timed pass, per bounce depth
    ctx.enable_timing(False)
    log(f"stage-timed pass: {staged['ms_total']:.0f} ms")
    ext_ms = staged["ms_extend"] / max(staged["extend_launches"], 1)
    ext_bytes = EXTEND_BYTES_PER_RAY * staged["extension_rays"] / max(staged["extend_launches"], 1)
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    dbar = totals["extension_rays"] / totals["samples"]
    pipe_bytes = samples * (RAYGEN_BYTES_PER_SAMPLE + FINALIZE_BYTES_PER_SAMPLE + dbar * SEGMENT_BYTES)
    rays = totals["extension_rays"] + totals["shadow_rays"]
    traffic = None  # measured for the signet still only
    tf = os.path.join(HERE, "profiles", "k_extend_signet_traffic.json")
    if args.workload == "signet" and os.path.exists(tf):
        traffic = json.load(open(tf)).get("bytes_per_dispatch")
    result = {
        "metric": ("Msamples/s at 1080x1920 256spp vs CPU; served p50/p99 latency, 1 GPU" if args.workload == "signet"
                   else "Msamples/s, 1 GPU"),
        "value": round(msps, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(msps / 17.7, 3) if args.workload == "signet" and spp == 256 else None,
        "vs_baseline_note": "published CPU still rate: 1920x1080 256 spp in ~30 s on 32 cores = 17.7 Msamples/s "
                            "(README.md:4-6)",
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{WORKLOAD_NAMES[args.workload]} {w}x{h} {spp}spp depth {depth}, {wl['metal']}, "
                               f"{wl['studio']} studio{', HDRI + ' + ' and '.join(wl['gems']) if 'gems' in wl else ''}, "
                               f"{len(piece.triangles)} triangles (seeded procedural)",
                   "width": w, "height": h, "spp": spp, "max_depth": depth, "triangles": int(len(piece.triangles))},
        "rays_per_s": round(rays / elapsed / 1e6, 1),
        "rays_per_s_unit": "Mrays/s (extension + shadow)",
        "segments_per_sample": round(dbar, 4),
        "kernel_ms_per_step": {k[3:]: round(staged[k], 1) for k in
                               ("ms_extend", "ms_shade", "ms_shadow", "ms_raygen", "ms_accum")},
        "kernel_ms_note": "from one extra untimed pass with per-stage event timing (separate extend and shadow "
                          "launches, whose pass took " + f"{staged['ms_total']:.0f} ms)",
        "roofline": {
            "kernel": "k_extend (closest-hit traversal + panels + floor, material bucketing)",
            "bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 5), "traffic": traffic,
            "bytes_per_unit": EXTEND_BYTES_PER_RAY, "unit_of_work": "extension ray",
            "launch_ms": round(ext_ms, 4),
        },
        "per_depth": {"extension_rays": ds["extension_rays"], "shadow_rays": ds["shadow_rays"],
                      "ms_extend": [round(x, 2) for x in ds["ms_extend"]],
                      "ms_shade": [round(x, 2) for x in ds["ms_shade"]],
                      "ms_shadow": [round(x, 2) for x in ds["ms_shadow"]]},
        "tuning": {"values": ctx.tuning, "source": tuning_source, "autotune": tune_info},
        "prep": prep,
        "pipeline_hbm": {"bytes_alg_per_step": int(pipe_bytes / args.steps),
                         "achieved_GBs": round(pipe_bytes / elapsed / 1e9, 2),
                         "frac": round(pipe_bytes / elapsed / 1e9 / PEAK_HBM_GBS, 5)},
    }
    film.close()
    scene.close()
    # the server runs with the library's launch shape (a still's tuning may
    # hold a wave far larger than serving's memory budget allows)
    ctx.set_tuning(**library_tuning)
    if args.serve_seconds > 0 and args.workload == "signet":
        from argent_amd import serve
        trace = serve.make_trace(args.serve_seconds, args.serve_rate)
        log(f"serving: {len(trace)} requests over {args.serve_seconds:.0f} s, building meshes ...")
        pieces = serve.build_pieces(8)
        log("serving meshes (triangles):", [len(p.triangles) for p in pieces])
        res = serve.run_trace(ctx, trace, pieces, budget_bytes=96 << 30, step_samples=1 << args.serve_step_log2,
                              depth=12, max_wall=args.serve_wall, log=log)
        res["step_samples"] = 1 << args.serve_step_log2
        res["trace"] = (f"Poisson {args.serve_rate:g}/s for {args.serve_seconds:g} s; resolution uniform over "
                        f"512x512, 1080x1920, 2000x2000; spp log-uniform 16..1024; 8 meshes "
                        f"{min(len(p.triangles) for p in pieces)}..{max(len(p.triangles) for p in pieces)} "
                        f"triangles; 6 alloys x 5 studios; depth 12; stopped after {args.serve_wall:g} s")
        if args.plain_seconds > 0:
            plain = serve.serial_plain(ctx, trace, pieces, depth=12, budget_s=args.plain_seconds, log=log)
            res["serial_plain"] = plain
            res["served_over_plain"] = round(res["served_msamples_per_s"] / max(plain["msamples_per_s"], 1e-9), 3)
        result["serving"] = res
        log("serving:", json.dumps(res))
    if not args.no_cpu:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline(desc, cam, w, h, spp, ps, seed, args.cpu_seconds)
        result["gpu_over_cpu"] = round(msps / result["cpu_baseline"]["value"], 1)
    ctx.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
