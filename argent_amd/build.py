"""Build libargent_hip.so in-tree with hipcc for gfx950.

The device code is compiled with -ffp-contract=off (expressions round as
written, so GPU and CPU evaluations of the same formula agree bit for bit)
and without packed-FP32 instructions.
"""
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libargent_hip.so")

HIP_SOURCES = ["kernels.hip", "post.hip", "mesh.hip", "bvh.hip", "normals.hip", "cavity.hip"]
CPP_SOURCES = ["runtime.cpp", "host_build.cpp", "host_spectra.cpp", "host_bvh.cpp", "host_post.cpp", "autotune.cpp"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def run(cmd):
    r = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build(verbose=False, jobs=8):
    hc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    common = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
              "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
    dev = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-fno-gpu-flush-denormals-to-zero"]
    procs = []
    objs = []
    t0 = time.time()
    for src in HIP_SOURCES + CPP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        if src.endswith(".hip"):
            cmd = [hc, "-x", "hip", *common, *dev, "-c", src, "-o", obj]
        else:
            cmd = [hc, "-x", "hip", *common, *dev, "-c", src, "-o", obj] if src in ("runtime.cpp", "host_post.cpp") else \
                  [hc, *common, "-c", src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, cwd=CSRC, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
    failed = []
    for src, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            failed.append((src, out + err))
        elif verbose and (out or err):
            sys.stderr.write(out + err)
    if failed:
        for src, msg in failed:
            sys.stderr.write(f"--- {src}\n{msg}\n")
        raise RuntimeError("compilation failed: " + ", ".join(s for s, _ in failed))
    tmp = OUT + ".tmp"
    run([hc, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp, "-lpthread"])
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT} in {time.time() - t0:.1f}s")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
