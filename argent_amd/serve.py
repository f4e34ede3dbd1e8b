"""Continuous-batching render server (the studio's live worker and job queue,
crates/argent/src/serve/live.rs:637-823 and queue.rs:475-691, for many
clients at once) and a replay driver for a seeded serving trace.

    python -m argent_amd.serve --seconds 20 --rate 4 [--out result.json]

The scheduler itself is native (ag_server_* in include/argent.h): requests
are admitted FIFO while their films fit the HBM budget, and every step
renders the next sample chunk of every active request in one shared
wavefront schedule, so requests arriving while others render join at the
next step.  This module only submits requests at their arrival times,
drives steps and records latency.
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import Context, Material, PathSettings, Scene, _f32, _ptr, check
from . import _abi as A


class Server:
    def __init__(self, ctx, budget_bytes=64 << 30, step_samples=1 << 24, ps=None):
        ps = ps or PathSettings(12)
        self.ctx = ctx
        h = C.c_void_p()
        check(A.lib().ag_server_create(ctx.h, budget_bytes, step_samples, C.byref(ps.c()), C.byref(h)))
        self.h = h
        self.dims = {}

    def close(self):
        if self.h:
            A.lib().ag_server_destroy(self.h)
            self.h = None

    def submit(self, scene, camera, width, height, spp, seed, turntable_deg=0.0, material0=None, preview_spp=2):
        r = A.Request()
        r.scene = scene.h
        r.camera = camera
        r.turntable_deg = turntable_deg
        if material0 is not None:
            r.has_material0 = 1
            r.material0 = material0
        r.width, r.height, r.spp, r.seed, r.preview_spp = width, height, spp, seed, preview_spp
        i = C.c_uint64()
        check(A.lib().ag_server_submit(self.h, C.byref(r), C.byref(i)))
        self.dims[i.value] = (height, width)
        return i.value

    def step(self):
        a, q = C.c_uint32(), C.c_uint32()
        check(A.lib().ag_server_step(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def status(self, i):
        s = A.RequestStatus()
        check(A.lib().ag_server_status(self.h, i, C.byref(s)))
        return {k: getattr(s, k) for k, _ in A.RequestStatus._fields_}

    def read(self, i, which=A.FILM_COLOR):
        h, w = self.dims[i]
        out = np.empty((h, w, 3), np.float32)
        check(A.lib().ag_server_read(self.h, i, which, _ptr(out)))
        return out

    def release(self, i):
        check(A.lib().ag_server_release(self.h, i))

    def error(self, i):
        """Why request i failed ('' unless its state is REQ_FAILED)."""
        return A.lib().ag_server_error(self.h, i).decode()

    def budget(self):
        """(bytes charged now: films, scenes, wave state; the budget)."""
        u, b = C.c_size_t(), C.c_size_t()
        check(A.lib().ag_server_budget(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value


ALLOYS = ["sterling", "gold14k", "rose14k", "bronze", "gold18k", "platinum"]
STUDIOS = ["noir", "white", "velvet", "ember", "window"]
RESOLUTIONS = [(512, 512), (1080, 1920), (2000, 2000)]


def make_trace(seconds, rate, seed=5, spp_lo=16, spp_hi=1024):
    """Poisson arrivals; per request: resolution uniform over three sizes,
    spp log-uniform in [spp_lo, spp_hi], alloy uniform over 6, studio over 5."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= seconds:
            break
        w, h = RESOLUTIONS[rng.integers(len(RESOLUTIONS))]
        spp = int(round(np.exp(rng.uniform(np.log(spp_lo), np.log(spp_hi)))))
        out.append(dict(t=t, mesh=int(rng.integers(1 << 30)), width=w, height=h, spp=spp,
                        alloy=ALLOYS[rng.integers(len(ALLOYS))], studio=STUDIOS[rng.integers(len(STUDIOS))],
                        seed=int(rng.integers(1 << 31))))
    return out


def build_pieces(n_meshes=8, seed=6):
    """Seeded meshes of 50k-2M triangles from the torus and signet generators."""
    from . import synth, workloads
    rng = np.random.default_rng(seed)
    pieces = []
    for k in range(n_meshes):
        if k % 2 == 0:
            nu = int(rng.integers(160, 1400))
            nv = max(32, nu // 4)
            pos, tris, nrm = synth.torus_ring(seed=int(rng.integers(1 << 30)), nu=nu, nv=nv)
        else:
            scale = float(rng.uniform(0.3, 1.2))
            pos, tris, nrm = synth.signet(seed=int(rng.integers(1 << 30)), band_nu=int(1024 * scale),
                                          band_nv=int(256 * scale), plate_nt=int(412 * scale),
                                          plate_np=int(820 * scale))
        pieces.append(workloads.Piece(pos, tris, nrm))
    return pieces


def run_trace(ctx, trace, pieces, budget_bytes=96 << 30, step_samples=1 << 25, depth=12, max_wall=None,
              log=lambda *a: None):
    """Replay `trace` in real time against one server; returns metrics.

    Each request is released as soon as it is done (its film's bytes go back
    to the budget).  Latencies run from the arrival time; a request still
    unfinished when the replay stops counts in the time-to-final percentiles
    as censored at that moment (a lower bound), and `final_censored` says how
    many did."""
    from . import workloads
    ps = workloads.path_settings(depth)
    scenes = {}
    bases = {}
    # one device scene per (mesh, studio), all scenes of a mesh sharing its
    # BVH; alloys are per-request slot-0 overrides
    for r in trace:
        key = (r["mesh"] % len(pieces), r["studio"])
        if key not in scenes:
            p = pieces[key[0]]
            d = p.scene("sterling", 0.07, 0.8, key[1])
            sc = Scene(ctx, d, share_mesh_with=bases.get(key[0]))
            if key[0] not in bases:
                p.prepare_patina(sc)  # piece preparation, before any look of it
            bases.setdefault(key[0], sc)
            scenes[key] = (sc, p, d)
    mats = {a: Material.metal(a, 0.07, 0.8) for a in ALLOYS}
    cams = {}
    srv = Server(ctx, budget_bytes, step_samples, ps)
    pending = list(trace)
    ids = []
    finished = {}  # id -> status when it left the server
    t0 = time.perf_counter()
    steps = 0
    active = queued = 0
    while True:
        now = time.perf_counter() - t0
        while pending and pending[0]["t"] <= now:
            r = pending.pop(0)
            key = (r["mesh"] % len(pieces), r["studio"])
            sc, p, _ = scenes[key]
            ck = (key[0], r["width"], r["height"], r["studio"])
            if ck not in cams:
                f_stop = 0.0 if r["studio"] == "white" else 16.0
                cams[ck] = p.shot("hero", 100.0, f_stop, r["width"] / r["height"], intersect=sc.intersect)
            cam, turn = cams[ck]
            i = srv.submit(sc, cam, r["width"], r["height"], r["spp"], r["seed"], turn, mats[r["alloy"]])
            # submissions happen between steps: keep how long the request waited since its arrival
            ids.append((i, r, ((time.perf_counter() - t0) - r["t"]) * 1e3))
        active, queued = srv.step()
        steps += 1
        for i, _, _ in ids:
            if i in finished:
                continue
            st = srv.status(i)
            if st["state"] in (A.REQ_DONE, A.REQ_FAILED, A.REQ_REJECTED):
                if st["state"] == A.REQ_FAILED:
                    log(f"request {i} failed: {srv.error(i)}")
                finished[i] = st
                srv.release(i)
        if active == 0:
            if queued == 0 and not pending:
                break
            # nothing on the device: wait for the next arrival (or, should the
            # head of the queue not fit, yield instead of spinning)
            nxt = pending[0]["t"] - (time.perf_counter() - t0) if pending else 0.001
            time.sleep(max(0.0005, min(0.05, nxt)))
        if max_wall is not None and time.perf_counter() - t0 > max_wall:
            log("wall limit reached")
            break
    wall = time.perf_counter() - t0
    backlog_active, backlog_queued = active, queued
    ttfp, ttf, done_samples, served, failed, censored = [], [], 0, 0, 0, 0
    offered = sum(r["width"] * r["height"] * r["spp"] for _, r, _ in ids) + sum(
        r["width"] * r["height"] * r["spp"] for r in pending)
    for i, r, lag in ids:  # latencies from the arrival time
        st = finished.get(i) or srv.status(i)
        done_samples += st["samples"] * r["width"] * r["height"]
        if st["t_first_preview"] >= 0:
            ttfp.append(st["t_first_preview"] - st["t_submit"] + lag)
        if st["state"] == A.REQ_DONE:
            served += 1
            ttf.append(st["t_done"] - st["t_submit"] + lag)
        elif st["state"] in (A.REQ_FAILED, A.REQ_REJECTED):
            failed += 1
        else:
            censored += 1
            ttf.append(wall * 1e3 - r["t"] * 1e3)
    srv.close()
    for sc, _, _ in scenes.values():
        if sc.base is not None:
            sc.close()
    for sc in bases.values():
        sc.close()
    pct = lambda v, q: float(np.percentile(v, q)) if v else None
    return {
        "requests": len(trace), "submitted": len(ids), "completed": served, "failed": failed, "steps": steps,
        "wall_s": round(wall, 2), "latency_from": "arrival",
        "ttfp_ms_p50": pct(ttfp, 50), "ttfp_ms_p99": pct(ttfp, 99), "previewed": len(ttfp),
        "final_ms_p50": pct(ttf, 50), "final_ms_p99": pct(ttf, 99), "final_censored": censored,
        "served_msamples_per_s": round(done_samples / wall / 1e6, 2),
        "backlog_active": backlog_active, "backlog_queued": backlog_queued,
        "not_submitted": len(pending),
        "offered_msamples_per_s": round(offered / max(trace[-1]["t"], 1e-9) / 1e6, 2) if trace else 0.0,
    }


def serial_plain(ctx, trace, pieces, depth=12, budget_s=30.0, log=lambda *a: None):
    """The plain per-request path on the same trace: each request, in arrival
    order, rendered alone by one render_pass of its full spp over a look whose
    slot 0 is its alloy (the reference's queue runs one job at a time,
    queue.rs:475-558).  Runs requests until budget_s of render time is spent;
    returns their throughput and the time to final a serial server would give
    them (each starts at its arrival or when the previous one ends)."""
    from . import Film, render_pass, workloads
    ps = workloads.path_settings(depth)
    bases, looks, cams = {}, {}, {}
    done, total_s, samples = [], 0.0, 0
    for r in trace:
        if total_s >= budget_s:
            break
        mi = r["mesh"] % len(pieces)
        p = pieces[mi]
        key = (mi, r["studio"], r["alloy"])
        if key not in looks:
            sc = Scene(ctx, p.scene(r["alloy"], 0.07, 0.8, r["studio"]), share_mesh_with=bases.get(mi))
            if mi not in bases:
                p.prepare_patina(sc)
            bases.setdefault(mi, sc)
            looks[key] = sc
        sc = looks[key]
        ck = (mi, r["width"], r["height"], r["studio"])
        if ck not in cams:
            f_stop = 0.0 if r["studio"] == "white" else 16.0
            cams[ck] = p.shot("hero", 100.0, f_stop, r["width"] / r["height"], intersect=sc.intersect)
        cam, turn = cams[ck]
        sc.set_turntable(turn)
        film = Film(ctx, r["width"], r["height"])
        ctx.synchronize()
        t0 = time.perf_counter()
        render_pass(ctx, sc, cam, film, ps, r["spp"], r["seed"])
        dt = time.perf_counter() - t0
        film.close()
        total_s += dt
        samples += r["width"] * r["height"] * r["spp"]
        done.append((r["t"], dt))
    end, lat = 0.0, []
    for t, dt in done:
        end = max(end, t) + dt
        lat.append((end - t) * 1e3)
    base_ids = {id(sc) for sc in bases.values()}
    for sc in looks.values():
        if id(sc) not in base_ids:
            sc.close()
    for sc in bases.values():
        sc.close()
    pct = lambda v, q: float(np.percentile(v, q)) if v else None
    log(f"serial plain: {len(done)} requests, {samples / 1e9:.2f} G samples in {total_s:.1f} s")
    return {"requests": len(done), "samples": int(samples), "render_s": round(total_s, 2),
            "msamples_per_s": round(samples / max(total_s, 1e-9) / 1e6, 2),
            "final_ms_p50": pct(lat, 50), "final_ms_p99": pct(lat, 99),
            "note": "the first requests of the trace in arrival order, one render_pass each at full spp; latency "
                    "from arrival on a serial server (start at arrival or at the previous request's end)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seconds", type=float, default=20.0)
    ap.add_argument("--rate", type=float, default=4.0)
    ap.add_argument("--spp-hi", type=int, default=1024)
    ap.add_argument("--meshes", type=int, default=8)
    ap.add_argument("--max-wall", type=float, default=None,
                    help="stop the replay after this many seconds (default: 3 x --seconds + 30)")
    ap.add_argument("--step-log2", type=int, default=25, help="server step budget, log2 samples")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    trace = make_trace(args.seconds, args.rate, spp_hi=args.spp_hi)
    log(f"{len(trace)} requests over {args.seconds} s")
    pieces = build_pieces(args.meshes)
    log("meshes:", [len(p.triangles) for p in pieces])
    with Context(0) as ctx:
        wall = args.max_wall if args.max_wall is not None else 3 * args.seconds + 30
        res = run_trace(ctx, trace, pieces, step_samples=1 << args.step_log2, max_wall=wall, log=log)
    res["config"] = dict(seconds=args.seconds, rate=args.rate, spp_hi=args.spp_hi, meshes=args.meshes,
                         step_samples=1 << args.step_log2)
    print(json.dumps(res))
    if args.out:
        json.dump(res, open(args.out, "w"), indent=1)


if __name__ == "__main__":
    main()
