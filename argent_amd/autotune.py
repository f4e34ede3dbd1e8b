"""Autotuned launch shapes: run the native autotuner (ag_autotune_render /
ag_autotune_rays) for a workload class, persist the choice, apply it later.

    choice = autotune.tune_render(ctx, scene, cam, w, h, spp, ps)
    autotune.save("still", choice)            # $ARGENT_TUNING_FILE or ~/.cache/argent_amd/tuning.json
    autotune.apply(ctx, autotune.load("still"))

Classes used by the benchmark and the server: "still" (a full-resolution
render), "preview" (a small pass: the live view's first look, a served
request's preview step), "rays" (a resident ray batch).  A choice measured
on an MI355X ships with the package (tuned_mi355x.json) and is what load()
returns when no file of the user's holds the class."""
import ctypes as C
import json
import os

from . import _abi as A
from ._abi import check

SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.json")
FIELDS = [k for k, _ in A.Tuning._fields_]


def default_path():
    return os.environ.get("ARGENT_TUNING_FILE") or os.path.join(os.path.expanduser("~"), ".cache", "argent_amd",
                                                                 "tuning.json")


def _result(r):
    return {"tuning": {k: int(getattr(r.best, k)) for k in FIELDS}, "ms_start": r.ms_start, "ms_best": r.ms_best,
            "trials": int(r.trials)}


def tune_render(ctx, scene, cam, width, height, spp, ps, repeats=3):
    """Autotune render passes of this workload; the best tuning is left on ctx.
    Returns {"tuning": {...}, "ms_start", "ms_best", "trials"}."""
    r = A.AutotuneResult()
    check(A.lib().ag_autotune_render(ctx.h, scene.h, C.byref(cam), width, height, spp, C.byref(ps.c()), repeats,
                                     C.byref(r)))
    return _result(r)


def tune_rays(ctx, scene, rays, repeats=3):
    """Autotune the closest-hit query on a resident batch (argent_amd.DeviceRays)."""
    r = A.AutotuneResult()
    check(A.lib().ag_autotune_rays(ctx.h, scene.h, C.byref(rays.rays), rays.n, C.byref(rays.hits), repeats,
                                   C.byref(r)))
    return _result(r)


def save(workload_class, choice, path=None, device="gfx950"):
    path = path or default_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    data = json.load(open(path)) if os.path.exists(path) else {}
    data.setdefault(device, {})[workload_class] = choice
    with open(path + ".tmp", "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(path + ".tmp", path)


def load(workload_class, path=None, device="gfx950"):
    """(tuning dict, source) for the class: the user's file first, then the
    shipped MI355X choice; (None, None) when neither has it."""
    for p, src in ((path or default_path(), "user"), (SHIPPED, "shipped")):
        if os.path.exists(p):
            entry = json.load(open(p)).get(device, {}).get(workload_class)
            if entry:
                return dict(entry["tuning"]), f"{src}:{p}"
    return None, None


def apply(ctx, tuning):
    """Set the given fields (a tuning dict) on the context; others keep their values."""
    if tuning:
        ctx.set_tuning(**{k: v for k, v in tuning.items() if k in FIELDS})


def main():
    """Tune the three classes on the benchmark workloads and write the choices:
    still (the signet at 1080 x 1920, trials at its 256 spp), preview (its live
    first look: 540 x 960 halved, 2 spp, depth 32) and rays (the 2^24-ray
    batch)."""
    import argparse
    import sys
    import time

    from . import Context, DeviceRays, PathSettings, Scene, synth, workloads
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="choice file (default: the user's tuning file)")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--classes", default="still,preview,rays")
    args = ap.parse_args()
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    wl = workloads.signet_still()
    piece = wl["piece"]
    with Context(0) as ctx:
        scene = Scene(ctx, piece.scene(wl["metal"], 0.07, 0.8, wl["studio"]))
        piece.prepare_patina(scene)
        defaults = ctx.tuning
        for cls in args.classes.split(","):
            t0 = time.time()
            if cls == "still":
                w, h = wl["width"], wl["height"]
                cam, turn = piece.shot("hero", 100.0, wl["f_stop"], w / h)
                scene.set_turntable(turn)
                res = tune_render(ctx, scene, cam, w, h, wl["spp"], workloads.path_settings(wl["depth"]), 1)
            elif cls == "preview":
                w, h = 270, 480
                cam, turn = piece.shot("hero", 100.0, wl["f_stop"], w / h)
                scene.set_turntable(turn)
                res = tune_render(ctx, scene, cam, w, h, 2, PathSettings(), args.repeats + 2)
            else:
                pos = piece.positions
                c = (pos.min(0) + pos.max(0)) * 0.5
                o, d = synth.sphere_rays(1 << 24, c, seed=3)
                rays = DeviceRays(ctx, o, d)
                res = tune_rays(ctx, scene, rays, args.repeats + 2)
                rays.close()
            res["seconds"] = round(time.time() - t0, 1)
            log(cls, res)
            save(cls, res, args.out)
            # each class starts from the library's defaults
            ctx.set_tuning(**defaults)
        scene.close()


if __name__ == "__main__":
    main()
