"""Writes tests/golden/xyz_refs.npz: reference renders in linear CIE XYZ made
by the oracle (the CPU restatement of the reference renderer) for reduced
versions of the three still workloads, and the oracle's own error at lower
sample counts.

    python tests/golden/make_xyz_refs.py [--threads 8]

Per scene (torus ring 128 x 128, depth 8, 14k gold, noir, f/16 with autofocus;
signet 54 x 96, depth 12, sterling, white; signet with 12 brilliants under the
HDRI and softboxes 54 x 96, depth 24):
  ref            16384 spp at seed 4096, XYZ = M^-1 * the film's colour
  camera         the ag_camera fields used (autofocus through the oracle's
                 closest-hit query, which the GPU's equals bit for bit)
  turn           the piece's turntable angle
  oracle_err_N   relative RMSE (XYZ) of the oracle's own N-spp render at seed 1
                 against ref, N in 256, 1024, 4096
Takes about ten minutes on eight cores."""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import argent_amd as ag  # noqa: E402
import oracle as O  # noqa: E402
from argent_amd import workloads  # noqa: E402

REF_SPP, REF_SEED, TEST_SEED = 16384, 4096, 1
ERR_SPP = (256, 1024, 4096)
CAM_FIELDS = ("pos", "forward", "right", "up", "half_w", "half_h", "lens_radius", "focus_dist", "shift")


NAMES = ("torus", "signet", "gems")


def scene_spec(name):
    """(workload dict, width, height) of a reduced still."""
    if name == "torus":
        return workloads.torus_still(), 128, 128
    if name == "signet":
        return workloads.signet_still(), 54, 96
    return workloads.signet_gems_still(hdri_size=(2048, 1024)), 54, 96


def build(wl):
    """(scene descriptor, path settings) of a still workload."""
    piece = wl["piece"]
    desc = piece.scene(wl["metal"], 0.07, 0.8, wl["studio"], gems=wl.get("gems", ()))
    if "hdri" in wl:
        desc.set_env_map(wl["hdri"])
    return desc, workloads.path_settings(wl["depth"])


def camera_to_array(cam):
    return np.concatenate([np.ravel(np.array(getattr(cam, f), np.float32)) for f in CAM_FIELDS])


def camera_from_array(a):
    cam = ag.A.Camera()
    k = 0
    for f in CAM_FIELDS:
        cur = getattr(cam, f)
        n = len(cur) if hasattr(cur, "__len__") else 1
        vals = a[k:k + n]
        if n == 1:
            setattr(cam, f, float(vals[0]))
        else:
            getattr(cam, f)[:] = [float(x) for x in vals]
        k += n
    return cam


def to_xyz(rgb_mean, desc):
    m = O.camera_matrix(np.array(desc.d.white_xyz, np.float32) if desc.d.has_white else None).astype(np.float64)
    return (rgb_mean.reshape(-1, 3).astype(np.float64) @ np.linalg.inv(m).T).reshape(rgb_mean.shape).astype(np.float32)


def oracle_xyz(osc, desc, cam, w, h, ps, spp, seed, threads):
    sums = np.zeros((3, h, w, 3), np.float32)
    done = 0
    while done < spp:  # passes of at most 1024 samples, continuing sample indices
        c = min(1024, spp - done)
        osc.render_pass(cam, w, h, sums, done, ps, c, seed, threads=threads)
        done += c
    return to_xyz(sums[0] / spp, desc)


def rel_rmse(a, b):
    return float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)) / np.sqrt(np.mean(b.astype(np.float64) ** 2)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "xyz_refs.npz"))
    args = ap.parse_args()
    out = {}
    for name in NAMES:
        wl, w, h = scene_spec(name)
        t0 = time.time()
        desc, ps = build(wl)
        osc = O.OracleScene(desc)
        cam, turn = wl["piece"].shot("hero", 100.0, wl["f_stop"], w / h,
                                     intersect=lambda o, d: osc.intersect(o, d, mode=O.MODE_ROBUST))
        osc.set_turntable(turn)
        ref = oracle_xyz(osc, desc, cam, w, h, ps, REF_SPP, REF_SEED, args.threads)
        out[f"{name}_ref"] = ref
        out[f"{name}_camera"] = camera_to_array(cam)
        out[f"{name}_turn"] = np.float32(turn)
        for n in ERR_SPP:
            e = rel_rmse(oracle_xyz(osc, desc, cam, w, h, ps, n, TEST_SEED, args.threads), ref)
            out[f"{name}_oracle_err_{n}"] = np.float64(e)
            print(f"{name}: oracle {n} spp relative RMSE {e:.4f}", flush=True)
        print(f"{name}: {time.time() - t0:.0f} s", flush=True)
        osc.close()
    np.savez_compressed(args.out, **out)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
