"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle (liboracle.so).

The oracle is the CPU restatement of the reference's render core (see
oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg import this module, as the checker of the HIP path; argent_amd never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from argent_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MODE_REFERENCE, MODE_ROBUST = 0, 1
STRATEGY_MIS, STRATEGY_NEE, STRATEGY_BSDF, STRATEGY_DOUBLE = 0, 1, 2, 3

_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)
_u8 = C.POINTER(C.c_uint8)
_u64 = C.POINTER(C.c_uint64)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB)
        rays = [_f] * 7
        sigs = [
            ("or_scene_create", C.c_void_p, [C.POINTER(A.SceneDesc)]),
            ("or_scene_destroy", None, [C.c_void_p]),
            ("or_scene_set_turntable", None, [C.c_void_p, C.c_float]),
            ("or_scene_set_material", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(A.Material)]),
            ("or_scene_stats", C.c_int, [C.c_void_p, _u32, _u32]),
            ("or_scene_set_strategy", C.c_int, [C.c_void_p, C.c_int32]),
            ("or_intersect", C.c_int, [C.c_void_p, *rays, C.c_uint32, C.c_int32, C.c_int32, _f, _f, _f, _u32]),
            ("or_occluded", C.c_int, [C.c_void_p, *rays, C.c_uint32, C.c_int32, C.c_int32, _u8]),
            ("or_intersect_stats", C.c_int, [C.c_void_p, *rays, C.c_uint32, C.c_int32, _u64]),
            ("or_ambient_occlusion", C.c_int, [C.c_void_p, C.c_uint32, C.c_float, C.c_int32, C.c_int32, _f]),
            ("or_tonemap_pixel", C.c_int, [_f, C.c_uint32, C.c_float, C.c_int32, C.c_float, _f]),
            ("or_develop_encoded", C.c_int, [_f, C.c_uint32, C.c_float, C.c_int32, C.c_float, _f]),
            ("or_srgb_encode", C.c_float, [C.c_float]),
            ("or_develop", C.c_int, [_f, C.c_uint32, C.c_uint32, C.c_float, C.c_int32, C.c_float, _u8]),
            ("or_atrous", C.c_int, [C.c_uint32, C.c_uint32, _f, _f, _f, _f]),
            ("or_brute_intersect", C.c_int, [_f, _u32, C.c_uint32, *rays, C.c_uint32, C.c_int32, _f, _f, _f, _u32]),
            ("or_render_pass", C.c_int, [C.c_void_p, C.POINTER(A.Camera), C.c_uint32, C.c_uint32, _f, C.c_uint32,
                                         C.POINTER(A.PathSettingsC), C.c_uint32, C.c_uint32, C.c_int32, C.c_uint32,
                                         C.c_uint32, _u64]),
            ("or_energy_table", C.c_int, [_f]),
            ("or_fresnel_dielectric", C.c_float, [C.c_float, C.c_float]),
            ("or_fresnel_conductor1", C.c_float, [C.c_float, C.c_float, C.c_float]),
            ("or_sample_wavelengths", C.c_int, [C.c_float, _f, _f]),
            ("or_camera_matrix", C.c_int, [C.c_int32, _f, _f]),
            ("or_rng_dim", C.c_float, [C.c_uint32, C.c_uint32, C.c_uint32]),
            ("or_camera_ray", C.c_int, [C.POINTER(A.Camera), C.c_float, C.c_float, C.c_float, C.c_float, _f, _f]),
            ("or_bsdf_probe", C.c_int, [C.c_void_p, C.c_int32, _f, C.c_float, _f, _f, _f, _f]),
            ("or_linear_rgb", C.c_int, [_f, _f]),
            ("or_env_build", C.c_int, [_f, C.c_int32, C.c_int32, C.c_float, _f, _f, _f, _f]),
            ("or_env_sample", C.c_int, [_f, _f, C.c_int32, C.c_int32, C.c_float, _f, C.c_uint32, _f, _f]),
            ("or_env_pdf", C.c_int, [_f, _f, C.c_int32, C.c_int32, C.c_float, _f, C.c_uint32, _f]),
            ("or_stl_count", C.c_int64, [C.c_char_p, C.c_size_t]),
            ("or_relative_change", C.c_double, [_f, _f, C.c_uint32]),
            ("or_normals", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, C.c_float, _f, _f, _u32, _u32, _u32, _u64]),
            ("or_weld", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, C.c_float, _f, _u32, _u32, _u32, _u64, _f]),
            ("or_compact", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, _f, _u32, _u32, _u32, _u64]),
            ("or_concavity", C.c_int, [_f, _f, C.c_uint32, _u32, C.c_uint32, C.c_float, _f]),
            ("or_smooth_vertex_normals", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, _f]),
            ("or_smooth", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, _f, C.c_uint32, C.c_float]),
            ("or_patina_finish", C.c_int, [_f, C.c_uint32, _u32, C.c_uint32, _f, _f, C.c_float, C.c_uint32, _f]),
            ("or_load_stl", C.c_int, [C.c_char_p, C.c_size_t, C.c_int32, C.c_float, C.c_float, _f, _u32, _u32,
                                      _u32, _u64, _f]),
        ]
        for name, res, args in sigs:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t=C.c_float):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _rays(o, d, tmax):
    o = np.ascontiguousarray(np.asarray(o, np.float32).reshape(-1, 3))
    d = np.ascontiguousarray(np.asarray(d, np.float32).reshape(-1, 3))
    cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + [np.ascontiguousarray(d[:, k]) for k in range(3)]
    tm = None if tmax is None else np.ascontiguousarray(np.asarray(tmax, np.float32))
    return cols, tm, len(o)


class OracleScene:
    def __init__(self, desc):
        desc._bind()
        self.desc = desc
        self.h = lib().or_scene_create(C.byref(desc.d))
        if not self.h:
            raise RuntimeError("oracle scene creation failed")

    def close(self):
        if self.h:
            lib().or_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_turntable(self, deg):
        lib().or_scene_set_turntable(self.h, deg)

    def set_material(self, m, mat):
        assert lib().or_scene_set_material(self.h, m, C.byref(mat)) == 0

    def set_strategy(self, strategy):
        """STRATEGY_MIS (the reference's), STRATEGY_NEE, STRATEGY_BSDF, or
        STRATEGY_DOUBLE (a planted double count the closed forms must reject)."""
        assert lib().or_scene_set_strategy(self.h, strategy) == 0

    def intersect(self, o, d, tmax=None, mode=MODE_ROBUST, threads=0):
        cols, tm, n = _rays(o, d, tmax)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        prim = np.empty(n, np.uint32)
        lib().or_intersect(self.h, *[_p(c) for c in cols], _p(tm), n, mode, threads, _p(t), _p(u), _p(v),
                           _p(prim, C.c_uint32))
        return t, u, v, prim

    def intersect_stats(self, o, d, tmax=None, mode=MODE_ROBUST):
        """(Node4 visits, triangle tests) of the reference's closest-hit query over these rays."""
        cols, tm, n = _rays(o, d, tmax)
        out = np.zeros(2, np.uint64)
        lib().or_intersect_stats(self.h, *[_p(c) for c in cols], _p(tm), n, mode, _p(out, C.c_uint64))
        return int(out[0]), int(out[1])

    def occluded(self, o, d, tmax=None, mode=MODE_ROBUST, threads=0):
        cols, tm, n = _rays(o, d, tmax)
        out = np.empty(n, np.uint8)
        lib().or_occluded(self.h, *[_p(c) for c in cols], _p(tm), n, mode, threads, _p(out, C.c_uint8))
        return out.astype(bool)

    def ambient_occlusion(self, n_vertices, samples=48, distance=1.2, mode=MODE_ROBUST, threads=0):
        """Per-vertex cavity AO (argent-mesh cavity.rs:95-128), float32 [n_vertices]."""
        out = np.zeros(n_vertices, np.float32)
        rc = lib().or_ambient_occlusion(self.h, samples, distance, mode, threads, _p(out))
        if rc:
            raise ValueError(f"or_ambient_occlusion: rc {rc}")
        return out

    def render_pass(self, cam, width, height, sums, first, ps, count, seed, threads=0, rows=None):
        """sums: float32 (3, h, w, 3) array updated in place.  Returns (segments, shadows, samples)."""
        assert sums.dtype == np.float32 and sums.flags.c_contiguous and sums.size == 9 * width * height
        r0, r1 = rows if rows else (0, height)
        st = np.zeros(3, np.uint64)
        lib().or_render_pass(self.h, C.byref(cam), width, height, _p(sums), first, C.byref(ps.c()), count, seed,
                             threads, r0, r1, _p(st, C.c_uint64))
        return tuple(int(x) for x in st)

    def bsdf_probe(self, m, lam, mask, u, wo, wi):
        out = np.empty(18, np.float32)
        args = [np.ascontiguousarray(np.asarray(a, np.float32)) for a in (lam, u, wo, wi)]
        lib().or_bsdf_probe(self.h, m, _p(args[0]), float(mask), _p(args[1]), _p(args[2]), _p(args[3]), _p(out))
        return out


def brute_intersect(positions, triangles, o, d, tmax=None, threads=0):
    pos = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1))
    tri = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1))
    cols, tm, n = _rays(o, d, tmax)
    t, u, v = (np.empty(n, np.float32) for _ in range(3))
    prim = np.empty(n, np.uint32)
    lib().or_brute_intersect(_p(pos), _p(tri, C.c_uint32), len(tri) // 3, *[_p(c) for c in cols], _p(tm), n, threads,
                             _p(t), _p(u), _p(v), _p(prim, C.c_uint32))
    return t, u, v, prim


def energy_table():
    out = np.empty(1024, np.float32)
    lib().or_energy_table(_p(out))
    return out


def fresnel_dielectric(c, eta):
    return lib().or_fresnel_dielectric(c, eta)


def fresnel_conductor1(c, eta, k):
    return lib().or_fresnel_conductor1(c, eta, k)


def sample_wavelengths(u):
    lam = np.empty(4, np.float32)
    pdf = np.empty(4, np.float32)
    lib().or_sample_wavelengths(u, _p(lam), _p(pdf))
    return lam, pdf


def camera_matrix(white=None):
    m = np.empty(9, np.float32)
    w = np.zeros(3, np.float32) if white is None else np.asarray(white, np.float32)
    lib().or_camera_matrix(0 if white is None else 1, _p(w), _p(m))
    return m.reshape(3, 3)


def rng_dim(ps, si, dim):
    return lib().or_rng_dim(ps, si, dim)


def linear_rgb():
    """LinearRgb (envmap.rs:15-105): bases (3, 471) and inverse response (3, 3)."""
    b = np.empty((3, 471), np.float32)
    m = np.empty((3, 3), np.float32)
    lib().or_linear_rgb(_p(b), _p(m))
    return b, m


def env_build(rgb, exposure_ev=0.0):
    """EnvData::load (envmap.rs:141-178) tables for a linear-RGB image (h, w, 3)."""
    rgb = np.ascontiguousarray(np.asarray(rgb, np.float32))
    h, w = rgb.shape[:2]
    weights = np.empty((h, w, 3), np.float32)
    marg = np.empty(h + 1, np.float32)
    cond = np.empty((h, w + 1), np.float32)
    total = np.empty(1, np.float32)
    lib().or_env_build(_p(rgb), w, h, exposure_ev, _p(weights), _p(marg), _p(cond), _p(total))
    return weights, marg, cond, float(total[0])


def env_sample(marg, cond, u, rotation_deg=0.0):
    """EnvMap::sample (envmap.rs:224-239) for (n, 2) uniforms: directions, pdfs."""
    h, w1 = cond.shape
    u = np.ascontiguousarray(np.asarray(u, np.float32))
    n = len(u)
    dirs = np.empty((n, 3), np.float32)
    pdfs = np.empty(n, np.float32)
    lib().or_env_sample(_p(np.ascontiguousarray(marg)), _p(np.ascontiguousarray(cond)), w1 - 1, h, rotation_deg,
                        _p(u), n, _p(dirs), _p(pdfs))
    return dirs, pdfs


def env_pdf(marg, cond, dirs, rotation_deg=0.0):
    """EnvMap::pdf (envmap.rs:208-221) of (n, 3) directions."""
    h, w1 = cond.shape
    dirs = np.ascontiguousarray(np.asarray(dirs, np.float32))
    pdfs = np.empty(len(dirs), np.float32)
    lib().or_env_pdf(_p(np.ascontiguousarray(marg)), _p(np.ascontiguousarray(cond)), w1 - 1, h, rotation_deg,
                     _p(dirs), len(dirs), _p(pdfs))
    return pdfs



# ------------------------------------------------------------------ preview development (post.cpp)
TONEMAPS = {"agx": 0, "agx_punchy": 1, "agx_golden": 2, "aces": 3, "khronos": 4, "reinhard": 5, "clamp": 6}


def _img(a):
    return np.ascontiguousarray(np.asarray(a, np.float32).reshape(-1, 3))


def tonemap_pixel(rgb, op, ev=0.0, sat=1.0):
    """tonemap.rs:207-218 (exact, display-linear), rgb (n, 3)."""
    x = _img(rgb)
    out = np.empty_like(x)
    if lib().or_tonemap_pixel(_p(x), len(x), ev, TONEMAPS.get(op, op), sat, _p(out)):
        raise ValueError(op)
    return out


def develop_encoded(rgb, op, ev=0.0, sat=1.0):
    """tonemap.rs:262-338 (the table fast path): sRGB-encoded 0..1, rgb (n, 3)."""
    x = _img(rgb)
    out = np.empty_like(x)
    if lib().or_develop_encoded(_p(x), len(x), ev, TONEMAPS.get(op, op), sat, _p(out)):
        raise ValueError(op)
    return out


def srgb_encode(v):
    return np.array([lib().or_srgb_encode(float(x)) for x in np.ravel(v)], np.float32).reshape(np.shape(v))


def develop(img, op, ev=0.0, sat=1.0):
    """tonemap.rs:370-388: (h, w, 3) float32 -> (h, w, 3) uint8."""
    img = np.ascontiguousarray(np.asarray(img, np.float32))
    h, w = img.shape[:2]
    out = np.empty((h, w, 3), np.uint8)
    if lib().or_develop(_p(img), w, h, ev, TONEMAPS.get(op, op), sat, _p(out, C.c_uint8)):
        raise ValueError(op)
    return out


def atrous(color, albedo=None, normal=None):
    """denoise.rs:660-709: (h, w, 3) float32 images -> (h, w, 3)."""
    color = np.ascontiguousarray(np.asarray(color, np.float32))
    h, w = color.shape[:2]
    a = None if albedo is None else np.ascontiguousarray(np.asarray(albedo, np.float32))
    n = None if normal is None else np.ascontiguousarray(np.asarray(normal, np.float32))
    out = np.empty_like(color)
    lib().or_atrous(w, h, _p(color), _p(a) if a is not None else None, _p(n) if n is not None else None, _p(out))
    return out


WELD_FIELDS = ("input_vertices", "output_vertices", "exact_merges", "tolerance_merges", "dropped_triangles")


def _mesh_in(pos, tris):
    pos = np.ascontiguousarray(np.asarray(pos, np.float32).reshape(-1, 3))
    tris = np.ascontiguousarray(np.asarray(tris, np.uint32).reshape(-1, 3))
    return pos, tris


def weld(pos, tris, rel_tolerance=1e-6):
    """weld.rs:38-62: (positions, triangles, report dict)."""
    pos, tris = _mesh_in(pos, tris)
    op = np.empty_like(pos)
    ot = np.empty_like(tris)
    nv, nt = C.c_uint32(), C.c_uint32()
    rep = np.zeros(5, np.uint64)
    tol = C.c_float()
    if lib().or_weld(_p(pos), len(pos), _p(tris, C.c_uint32), len(tris), rel_tolerance, _p(op), C.byref(nv),
                     _p(ot, C.c_uint32), C.byref(nt), _p(rep, C.c_uint64), C.byref(tol)):
        raise IndexError("vertex index out of range")
    r = dict(zip(WELD_FIELDS, (int(x) for x in rep)))
    r["tolerance_mm"] = tol.value
    return op[:nv.value].copy(), ot[:nt.value].copy(), r


def compact(pos, tris):
    """weld.rs:66-91: (positions, triangles, dropped triangles)."""
    pos, tris = _mesh_in(pos, tris)
    op = np.empty_like(pos)
    ot = np.empty_like(tris)
    nv, nt = C.c_uint32(), C.c_uint32()
    dropped = C.c_uint64()
    if lib().or_compact(_p(pos), len(pos), _p(tris, C.c_uint32), len(tris), _p(op), C.byref(nv),
                        _p(ot, C.c_uint32), C.byref(nt), C.byref(dropped)):
        raise IndexError("vertex index out of range")
    return op[:nv.value].copy(), ot[:nt.value].copy(), dropped.value


LOAD_ERRORS = {1: "binary STL shorter than its 84-byte header", 2: "file contains no triangles",
               3: "file has only degenerate triangles"}


def load_stl(data, weld=True, weld_tolerance=1e-6, unit_scale=1.0):
    """io.rs:96-122 for a binary STL: (positions, triangles, report dict)."""
    data = bytes(data)
    n = lib().or_stl_count(data, len(data))
    cap = max(int(n), 1)
    op = np.empty((3 * cap, 3), np.float32)
    ot = np.empty((cap, 3), np.uint32)
    nv, nt = C.c_uint32(), C.c_uint32()
    rep = np.zeros(10, np.uint64)
    tol = C.c_float()
    rc = lib().or_load_stl(data, len(data), int(bool(weld)), weld_tolerance, unit_scale, _p(op), C.byref(nv),
                           _p(ot, C.c_uint32), C.byref(nt), _p(rep, C.c_uint64), C.byref(tol))
    if rc:
        raise ValueError(LOAD_ERRORS[rc])
    r = dict(zip(("raw_triangles", "raw_vertices", "invalid_triangles", "trailing_bytes", "declared_triangles"),
                 (int(x) for x in rep[:5])))
    r["weld"] = dict(zip(WELD_FIELDS, (int(x) for x in rep[5:]))) if weld else None
    if weld:
        r["weld"]["tolerance_mm"] = tol.value
    return op[:nv.value].copy(), ot[:nt.value].copy(), r


def relative_change(a, b):
    """live.rs:208-218."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    b = np.ascontiguousarray(np.asarray(b, np.float32))
    return lib().or_relative_change(_p(a), _p(b), a.size // 3)


NORMAL_FIELDS = ("shells", "flipped_shells", "cavity_shells", "repaired_faces", "boundary_edges",
                 "nonmanifold_edges", "inconsistent_edges", "degenerate_faces", "split_vertices")


def normals(pos, tris, crease_angle_deg=50.0):
    """normals.rs:47-49: (positions, normals, triangles, report dict)."""
    pos, tris = _mesh_in(pos, tris)
    cap = len(pos) + 3 * len(tris)
    op = np.empty((cap, 3), np.float32)
    on = np.empty((cap, 3), np.float32)
    ot = np.empty((max(len(tris), 1), 3), np.uint32)
    nv, nt = C.c_uint32(), C.c_uint32()
    rep = np.zeros(9, np.uint64)
    if lib().or_normals(_p(pos), len(pos), _p(tris, C.c_uint32), len(tris), crease_angle_deg, _p(op), _p(on),
                        C.byref(nv), _p(ot, C.c_uint32), C.byref(nt), _p(rep, C.c_uint64)):
        raise IndexError("vertex index out of range")
    return (op[:nv.value].copy(), on[:nv.value].copy(), ot[:nt.value].copy(),
            dict(zip(NORMAL_FIELDS, (int(x) for x in rep))))


def concavity(pos, nrm, tris, radius):
    """cavity.rs:130-243: signed concavity per vertex, float32 [n_vertices]."""
    pos, tris = _mesh_in(pos, tris)
    nrm = np.ascontiguousarray(np.asarray(nrm, np.float32).reshape(-1, 3))
    out = np.zeros(len(pos), np.float32)
    lib().or_concavity(_p(pos), _p(nrm), len(pos), _p(tris, C.c_uint32), len(tris), radius, _p(out))
    return out


def smooth(pos, tris, values, iterations, lam=0.5):
    """cavity.rs:245-296: Laplacian smoothing; returns the smoothed copy."""
    pos, tris = _mesh_in(pos, tris)
    v = np.ascontiguousarray(np.asarray(values, np.float32)).copy()
    assert v.size == len(pos), "smooth: one value per vertex"
    lib().or_smooth(_p(pos), len(pos), _p(tris, C.c_uint32), len(tris), _p(v), iterations, lam)
    return v


def patina_mask(pos, nrm, tris, ao_samples=48, ao_distance=1.2, curvature_radius=0.4, concavity_weight=0.35,
                smooth_iterations=3, mode=MODE_ROBUST, ao=None):
    """cavity.rs:52-80 with the mesh's own BVH as occluder (the scene's
    Bvh::occluded); `ao` may carry a precomputed AO term.  float32 [n_vertices]."""
    pos, tris = _mesh_in(pos, tris)
    nv = len(pos)
    if nv == 0 or len(tris) == 0:
        return np.zeros(nv, np.float32)
    # cavity.rs:82-87: the mesh's normals, else smooth_vertex_normals
    nrm = smooth_vertex_normals(pos, tris) if nrm is None else np.ascontiguousarray(
        np.asarray(nrm, np.float32).reshape(-1, 3))
    if ao is None:
        if ao_samples > 0 and ao_distance > 0.0:
            d = _piece_desc(pos, tris, nrm)
            sc = OracleScene(d)
            ao = sc.ambient_occlusion(nv, ao_samples, ao_distance, mode)
            sc.close()
        else:
            ao = np.zeros(nv, np.float32)
    ao = np.ascontiguousarray(np.asarray(ao, np.float32))
    w = min(max(concavity_weight, 0.0), 1.0)
    cav = concavity(pos, nrm, tris, curvature_radius) if w > 0.0 and curvature_radius > 0.0 else np.zeros(nv, np.float32)
    out = np.zeros(nv, np.float32)
    lib().or_patina_finish(_p(pos), nv, _p(tris, C.c_uint32), len(tris), _p(ao), _p(cav), concavity_weight,
                           smooth_iterations, _p(out))
    return out


def _piece_desc(pos, tris, nrm):
    """A scene holding only the piece (one sterling material), for AO."""
    import argent_amd as ag
    d = ag.SceneDesc(pos, tris, normals=nrm)
    d.add_material(ag.Material.metal("sterling"))
    return d


def smooth_vertex_normals(pos, tris):
    """normals.rs:226-242: angle-weighted vertex normals, float32 (n_vertices, 3)."""
    pos, tris = _mesh_in(pos, tris)
    out = np.zeros_like(pos)
    if lib().or_smooth_vertex_normals(_p(pos), len(pos), _p(tris, C.c_uint32), len(tris), _p(out)):
        raise IndexError("vertex index out of range")
    return out
