"""Closest-hit agreement between two traversals of different trees.

f32 Moller-Trumbore (bvh.rs:391-409) can report a hit a little outside its
triangle for a ray that grazes it; whether a traversal tests that triangle at
all depends on how its tree groups triangles, in the reference as here.  Two
correct traversals may therefore differ only on rays where the closer of the
two answers is such an artefact: in f64 the ray misses that triangle.  A
genuine closer hit is never allowed to be missed."""
import numpy as np


def f64_barycentric_min(pos, tris, prim, o, d):
    """Smallest exact (f64) barycentric coordinate of the ray's crossing of
    the triangle's plane: negative means the ray misses the triangle."""
    v = pos[tris[prim]].astype(np.float64)
    e1, e2 = v[1] - v[0], v[2] - v[0]
    dd, oo = d.astype(np.float64), o.astype(np.float64)
    p = np.cross(dd, e2)
    det = e1 @ p
    s = oo - v[0]
    u = (s @ p) / det
    w = (dd @ np.cross(s, e1)) / det
    return min(u, w, 1.0 - u - w)


def assert_same_closest(pos, tris, o, d, got, ref, max_rate=1e-6):
    """got/ref: (t, prim) arrays.  Equal ids everywhere except artefact rays,
    which must be rare and each verified as above."""
    gt, gp = got
    rt, rp = ref
    bad = np.nonzero(gp != rp)[0]
    if len(bad) == 0:
        return 0
    print(f"{len(bad)} rays differ; checking each in f64")
    assert len(bad) <= max(2, max_rate * len(gp)), len(bad)
    for i in bad:
        cands = [(float(t), int(p)) for t, p in ((gt[i], gp[i]), (rt[i], rp[i])) if p != 0xFFFFFFFF]
        assert cands, i
        t_close, p_close = min(cands)
        m = f64_barycentric_min(pos, tris, p_close, o[i], d[i])
        print(f"  ray {i}: got {gp[i]} t={gt[i]}, ref {rp[i]} t={rt[i]}; closer prim {p_close} "
              f"f64 barycentric min {m:.3g}")
        assert m < 0.0, ("a genuine closer hit was missed", i, p_close)
    return len(bad)


def classify_rule_differences(pos, tris, o, d, ref, robust):
    """Rays where the reference's traversal rules (bvh.rs:159-245: exact box
    test, first hit found wins an equal t) and the robust rules (boxes widened
    by 2^-16, equal t to the lower triangle id) disagree.  Each must be either
    an exact tie -- both report the same t for two triangles, a ray through
    their shared edge -- or a grazing ray whose closer answer is an f32
    artefact outside its triangle in f64.  Returns (ties, artefacts)."""
    xt, xp = ref
    rt, rp = robust
    ties = arts = 0
    for i in np.nonzero((xp != rp) | (xt != rt))[0]:
        if xt[i] == rt[i] and xp[i] != 0xFFFFFFFF and rp[i] != 0xFFFFFFFF:
            ties += 1
            continue
        cands = [(float(t), int(p)) for t, p in ((xt[i], xp[i]), (rt[i], rp[i])) if p != 0xFFFFFFFF]
        t_close, p_close = min(cands)
        m = f64_barycentric_min(pos, tris, p_close, o[i], d[i])
        assert m < 0.0, ("the rules differ on a ray whose closer hit is genuine", i, p_close, m)
        arts += 1
    return ties, arts
