"""argent_amd — MI355X-native render core of argent.

A thin Python face over the C ABI in include/argent.h (libargent_hip.so,
hand-written gfx950 HIP kernels).  The names mirror argent-core's API:

    Scene / Piece          crates/argent-core/src/scene.rs:10-86
    Material constructors  crates/argent-core/src/material.rs:39-79
    Studio rigs            crates/argent-core/src/studio.rs:87-197
    Camera.look_at, frame  crates/argent-core/src/camera.rs:38-55, views.rs:45-83
    Film                   crates/argent-core/src/render.rs:13-64
    render_pass            crates/argent-core/src/render.rs:67-115
    Scene.intersect/occluded  crates/argent-core/src/bvh.rs:159-245

Every compute call runs on the GPU through the native library; there is no
CPU fallback in this package.
"""
import ctypes as C
import time

import numpy as np

from . import _abi as A
from ._abi import ArgentError, check

__all__ = ["Live", "Mesh", "load_stl", "Context", "SceneDesc", "Scene", "Film", "PathSettings", "Material", "render_pass", "render_pass_async", "Pass", "render_views",
           "look_at", "frame", "VIEWS", "ArgentError", "alloy_f0_srgb", "alloy_ior", "grade", "develop_image",
           "denoise_image"]

VIEWS = {  # views.rs:17-31: (azimuth, elevation, fill)
    "hero": (34.0, 26.0, 0.72), "front": (0.0, 6.0, 0.70), "face": (0.0, 6.0, 0.70),
    "side": (90.0, 6.0, 0.72), "profile": (90.0, 6.0, 0.72), "top": (0.0, 62.0, 0.70),
    "three-quarter": (45.0, 24.0, 0.72), "34": (45.0, 24.0, 0.72), "back": (160.0, 16.0, 0.72),
    "low": (22.0, 3.0, 0.72), "detail": (14.0, 34.0, 0.92), "macro": (14.0, 34.0, 0.92),
}


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a, t=C.c_float):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class PathSettings:
    """integrator.rs:10-24; defaults (32, 24.0, 6)."""

    def __init__(self, max_depth=32, clamp_indirect=24.0, rr_depth=6):
        self.max_depth, self.clamp_indirect, self.rr_depth = int(max_depth), float(clamp_indirect), int(rr_depth)

    def c(self):
        return A.PathSettingsC(self.max_depth, self.clamp_indirect, self.rr_depth)


class Material:
    """Constructors of material.rs; each returns the baked ag_material."""

    @staticmethod
    def metal(alloy="sterling", roughness=0.07, patina=0.8):
        m = A.Material()
        check(A.lib().ag_material_metal(alloy.encode(), roughness, -1.0 if patina is None else patina, C.byref(m)))
        return m

    @staticmethod
    def gem(name):
        m = A.Material()
        check(A.lib().ag_material_gem(name.encode(), C.byref(m)))
        return m

    @staticmethod
    def plastic(rgb, eta, roughness):
        m = A.Material()
        check(A.lib().ag_material_plastic(_ptr(_f32(rgb)), eta, roughness, C.byref(m)))
        return m

    @staticmethod
    def diffuse(rgb):
        m = A.Material()
        check(A.lib().ag_material_diffuse(_ptr(_f32(rgb)), C.byref(m)))
        return m

    @staticmethod
    def black():
        m = A.Material()
        m.kind = A.MAT_BLACK
        return m


def alloy_ior(name):
    n = np.zeros(A.GRID_N, np.float32)
    k = np.zeros(A.GRID_N, np.float32)
    check(A.lib().ag_alloy_ior(name.encode(), _ptr(n), _ptr(k)))
    return n, k


def alloy_f0_srgb(name):
    rgb = np.zeros(3, np.float32)
    check(A.lib().ag_alloy_f0_srgb(name.encode(), _ptr(rgb)))
    return rgb


class Context:
    """One HIP device and stream (ag_ctx)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(A.lib().ag_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            A.lib().ag_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def last_query_ms(self):
        """Kernel time of the last ray query, cavity AO or patina mask call (HIP events), ms."""
        return A.lib().ag_last_query_ms(self.h)

    def graph_stats(self):
        """(captures, replays) of render passes in HIP graphs (tuning launch_graph)."""
        a, b = C.c_uint64(), C.c_uint64()
        check(A.lib().ag_ctx_graph_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def synchronize(self):
        check(A.lib().ag_ctx_synchronize(self.h))

    @property
    def tuning(self):
        t = A.Tuning()
        check(A.lib().ag_ctx_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.Tuning._fields_}

    def set_tuning(self, **kw):
        t = A.Tuning()
        check(A.lib().ag_ctx_get_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(A.lib().ag_ctx_set_tuning(self.h, C.byref(t)))

    def debug_fail_seed(self, seed):
        """Failure injection for tests (ag_debug_fail_seed): render passes that
        include a view rendered with `seed` fail before launching anything;
        None turns it off."""
        check(A.lib().ag_debug_fail_seed(self.h, -1 if seed is None else int(seed)))

    def enable_timing(self, on=True):
        check(A.lib().ag_ctx_enable_timing(self.h, 1 if on else 0))

    def stats(self):
        s = A.Stats()
        check(A.lib().ag_ctx_stats(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in A.Stats._fields_}

    def depth_stats(self, n=64):
        """Per bounce depth of the last render: extension and shadow rays, and
        (timing enabled) ms in the extend, shade and shadow kernels."""
        e, sh = (C.c_uint64 * n)(), (C.c_uint64 * n)()
        me, ms, mh = (C.c_double * n)(), (C.c_double * n)(), (C.c_double * n)()
        check(A.lib().ag_ctx_depth_stats(self.h, n, e, sh, me, ms, mh))
        return {"extension_rays": list(e), "shadow_rays": list(sh), "ms_extend": list(me), "ms_shade": list(ms),
                "ms_shadow": list(mh)}


class SceneDesc:
    """Host-side scene description (Scene::new inputs, scene.rs:82-86).

    Holds the mesh arrays, materials, panels, environment and floor; passes
    to both Scene (device) and the oracle unchanged."""

    def __init__(self, positions, triangles, normals=None, tri_material=None, cavity=None):
        self.positions = _f32(positions, (-1, 3))
        self.triangles = np.ascontiguousarray(np.asarray(triangles, dtype=np.uint32).reshape(-1, 3))
        self.normals = None if normals is None else _f32(normals, (-1, 3))
        self.tri_material = (np.zeros(len(self.triangles), np.uint16) if tri_material is None
                             else np.ascontiguousarray(np.asarray(tri_material, dtype=np.uint16)))
        self.cavity = None if cavity is None else _f32(cavity)
        self.d = A.SceneDesc()
        check(A.lib().ag_desc_init(C.byref(self.d)))
        self._env_owned = False
        self._bind()

    def _bind(self):
        d = self.d
        d.positions = _ptr(self.positions)
        d.triangles = _ptr(self.triangles, C.c_uint32)
        d.normals = _ptr(self.normals) if self.normals is not None else None
        d.tri_material = _ptr(self.tri_material, C.c_uint16)
        d.cavity = _ptr(self.cavity) if self.cavity is not None else None
        d.n_vertices = len(self.positions)
        d.n_triangles = len(self.triangles)

    def add_material(self, m):
        i = self.d.n_materials
        if i >= A.MAX_MATERIALS:
            raise ArgentError("too many materials")
        self.d.materials[i] = m
        self.d.n_materials = i + 1
        return i

    def apply_studio(self, studio, radius, floor_y):
        """Studio::build(radius) + floor_at(floor_y - 0.005) (studio.rs:87-197, shot.rs:226-241)."""
        check(A.lib().ag_desc_apply_studio(C.byref(self.d), studio.encode(), radius, floor_y))

    def set_env_map(self, rgb, exposure_ev=0.0, rotation_deg=0.0):
        """EnvData::load on in-memory linear RGB texels (envmap.rs:141-178)."""
        rgb = _f32(rgb)
        h, w = rgb.shape[0], rgb.shape[1]
        keep = (self.d.env.top[:], self.d.env.horizon[:], self.d.env.bottom[:], self.d.env.power)
        self.free_env()
        env = A.Env()
        check(A.lib().ag_env_from_rgb(_ptr(rgb.reshape(-1)), w, h, exposure_ev, rotation_deg, C.byref(env)))
        env.top[:], env.horizon[:], env.bottom[:], env.power = keep
        self.d.env = env
        self._env_owned = True

    def light_cdf(self):
        """Scene::rebuild_lights (scene.rs:88-110): (number of emitting panels,
        their indices, the CDF by power, the probability of picking the
        environment map)."""
        n = C.c_int32()
        idx = np.zeros(A.MAX_PANELS, np.int32)
        cdf = np.zeros(A.MAX_PANELS, np.float32)
        pick = C.c_float()
        check(A.lib().ag_desc_light_cdf(C.byref(self.d), C.byref(n), idx.ctypes.data_as(A.i32p),
                                        cdf.ctypes.data_as(A.f32p), C.byref(pick)))
        return n.value, idx[:n.value], cdf[:n.value], pick.value

    def free_env(self):
        if self._env_owned:
            A.lib().ag_env_free(C.byref(self.d.env))
            self._env_owned = False

    def __del__(self):
        try:
            self.free_env()
        except Exception:
            pass

    def add_panel(self, center, hu, hv, emission=None, camera_visible=False, feather=0.35):
        i = self.d.n_panels
        p = self.d.panels[i]
        p.center[:] = list(map(float, center))
        p.hu[:] = list(map(float, hu))
        p.hv[:] = list(map(float, hv))
        p.has_emission = 0 if emission is None else 1
        if emission is not None:
            p.emission[:] = list(map(float, _f32(emission)))
        p.camera_visible = 1 if camera_visible else 0
        p.feather = feather
        self.d.n_panels = i + 1
        return i

    def set_gradient_env(self, top, horizon, bottom, power):
        for name, rgb in (("top", top), ("horizon", horizon), ("bottom", bottom)):
            out = np.zeros(A.SPEC_N, np.float32)
            check(A.lib().ag_spectrum_rgb_illuminant(_ptr(_f32(rgb)), _ptr(out)))
            getattr(self.d.env, name)[:] = out.tolist()
        self.d.env.power = power

    def set_white_kelvin(self, kelvin):
        xyz = np.zeros(3, np.float32)
        check(A.lib().ag_blackbody_white(kelvin, _ptr(xyz)))
        self.d.has_white = 1
        self.d.white_xyz[:] = xyz.tolist()

    def bounds(self):
        return self.positions.min(axis=0), self.positions.max(axis=0)


class Scene:
    """Device scene (ag_scene): BVH, mesh attributes, tables resident in HBM."""

    def __init__(self, ctx, desc, share_mesh_with=None, gpu_bvh=False):
        """With `share_mesh_with` (a Scene of the same mesh), only the look --
        materials, panels, environment, floor -- comes from `desc`; the mesh
        and BVH on the device are shared (ag_scene_create_shared).  With
        `gpu_bvh` the BVH is built on the device (ag_scene_create_ex)."""
        self.ctx = ctx
        self.desc = desc
        self.base = share_mesh_with
        h = C.c_void_p()
        desc._bind()
        if share_mesh_with is None and gpu_bvh:
            check(A.lib().ag_scene_create_ex(ctx.h, C.byref(desc.d), 1, C.byref(h)))
        elif share_mesh_with is None:
            check(A.lib().ag_scene_create(ctx.h, C.byref(desc.d), C.byref(h)))
        else:
            check(A.lib().ag_scene_create_shared(ctx.h, share_mesh_with.h, C.byref(desc.d), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            A.lib().ag_scene_destroy(self.h)
            self.h = None

    def set_turntable(self, deg):
        check(A.lib().ag_scene_set_turntable(self.h, deg))

    def set_material(self, m, mat):
        check(A.lib().ag_scene_set_material(self.h, m, C.byref(mat)))

    @classmethod
    def from_mesh(cls, ctx, mesh, desc, gpu_bvh=True):
        """Scene over a device Mesh (ag_scene_create_from_mesh): the look comes
        from `desc`, the geometry stays on the device."""
        sc = cls.__new__(cls)
        sc.ctx, sc.desc, sc.base = ctx, desc, None
        h = C.c_void_p()
        desc._bind()
        check(A.lib().ag_scene_create_from_mesh(ctx.h, mesh.h, C.byref(desc.d), 1 if gpu_bvh else 0, C.byref(h)))
        sc.h = h
        sc._nv = mesh.counts[0]
        return sc

    @property
    def n_vertices(self):
        root = self
        while root.base is not None:
            root = root.base
        nv = getattr(root, "_nv", None)
        return nv if nv is not None else len(root.desc.positions)

    def stats(self):
        n, l, dp = C.c_uint32(), C.c_uint32(), C.c_uint32()
        b = C.c_size_t()
        check(A.lib().ag_scene_stats(self.h, C.byref(n), C.byref(l), C.byref(dp), C.byref(b)))
        return {"nodes": n.value, "leaves": l.value, "depth": dp.value, "bytes": b.value}

    def intersect(self, o, d, tmax=None):
        """Bvh::intersect on the GPU for rays (n,3) o and d: returns t, u, v, prim."""
        o, d = _f32(o, (-1, 3)), _f32(d, (-1, 3))
        n = len(o)
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + [np.ascontiguousarray(d[:, k]) for k in range(3)]
        tm = None if tmax is None else _f32(tmax)
        rays = A.RaysSoA(*[_ptr(c) for c in cols], _ptr(tm) if tm is not None else None)
        t, u, v = (np.empty(n, np.float32) for _ in range(3))
        prim = np.empty(n, np.uint32)
        hits = A.HitsSoA(_ptr(t), _ptr(u), _ptr(v), _ptr(prim, C.c_uint32))
        check(A.lib().ag_intersect(self.ctx.h, self.h, C.byref(rays), n, C.byref(hits)))
        return t, u, v, prim

    def occluded(self, o, d, tmax=None):
        o, d = _f32(o, (-1, 3)), _f32(d, (-1, 3))
        n = len(o)
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + [np.ascontiguousarray(d[:, k]) for k in range(3)]
        tm = None if tmax is None else _f32(tmax)
        rays = A.RaysSoA(*[_ptr(c) for c in cols], _ptr(tm) if tm is not None else None)
        out = np.empty(n, np.uint8)
        check(A.lib().ag_occluded(self.ctx.h, self.h, C.byref(rays), n, _ptr(out, C.c_uint8)))
        return out.astype(bool)

    def ambient_occlusion(self, samples=48, distance=1.2):
        """Cavity AO per vertex on the GPU (ag_ambient_occlusion; argent-mesh
        cavity.rs:95-128 with the piece's BVH as occluder): float32 [n_vertices]."""
        root = self
        while root.base is not None:
            root = root.base
        out = np.zeros(len(root.desc.positions), np.float32)
        check(A.lib().ag_ambient_occlusion(self.ctx.h, self.h, samples, distance, _ptr(out)))
        return out

    def patina_settings(self, **overrides):
        """PatinaSettings sized by the piece as Prepared::build does
        (shot.rs:133-141), with `overrides` applied: a dict."""
        ps = A.PatinaSettings()
        check(A.lib().ag_patina_for_scene(self.h, C.byref(ps)))
        d = {k: getattr(ps, k) for k, _ in A.PatinaSettings._fields_}
        d.update(overrides)
        return d

    def patina_mask(self, settings=None, store=False, parts=False):
        """patina_mask (cavity.rs:52-80) on the GPU (ag_patina_mask): float32
        [n_vertices] in [0, 1]; `settings` a dict (default: PatinaSettings::default()).
        store: the mask becomes the piece's cavity for shading (Piece::cavity).
        parts: also return the AO fraction and the concavity."""
        ps = A.PatinaSettings()
        check(A.lib().ag_patina_default(C.byref(ps)))
        for k, v in (settings or {}).items():
            setattr(ps, k, v)
        nv = self.n_vertices
        mask = np.zeros(nv, np.float32)
        ao = np.zeros(nv, np.float32) if parts else None
        cav = np.zeros(nv, np.float32) if parts else None
        check(A.lib().ag_patina_mask(self.ctx.h, self.h, C.byref(ps), 1 if store else 0, _ptr(mask),
                                     _ptr(ao) if parts else None, _ptr(cav) if parts else None))
        return (mask, ao, cav) if parts else mask

    def smooth_values(self, values, iterations, lam=0.5):
        """smooth (cavity.rs:245-296) over the piece on the GPU: a smoothed copy."""
        v = np.ascontiguousarray(np.asarray(values, np.float32)).copy()
        assert v.size == self.n_vertices, "one value per vertex"
        check(A.lib().ag_smooth_values(self.ctx.h, self.h, _ptr(v), iterations, lam))
        return v

    def bsdf_probe(self, mat, lam, mask, u, wo, wi):
        mat = np.ascontiguousarray(np.asarray(mat, np.int32))
        n = len(mat)
        out = np.empty((n, 18), np.float32)
        args = [_f32(lam, (n, 4)), _f32(mask, (n,)), _f32(u, (n, 4)), _f32(wo, (n, 3)), _f32(wi, (n, 3))]
        check(A.lib().ag_bsdf_probe(self.ctx.h, self.h, n, _ptr(mat, C.c_int32), *[_ptr(a) for a in args],
                                    _ptr(out)))
        return out


class Film:
    """render.rs:13-64; sums live in device memory."""

    def __init__(self, ctx, width, height):
        self.ctx, self.width, self.height = ctx, int(width), int(height)
        h = C.c_void_p()
        check(A.lib().ag_film_create(ctx.h, self.width, self.height, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            A.lib().ag_film_destroy(self.h)
            self.h = None

    @property
    def samples(self):
        return A.lib().ag_film_samples(self.h)

    def clear(self):
        check(A.lib().ag_film_clear(self.h))

    def _read(self, which, sums=False):
        out = np.empty((self.height, self.width, 3), np.float32)
        f = A.lib().ag_film_read_sums if sums else A.lib().ag_film_read
        check(f(self.h, which, _ptr(out)))
        return out

    def color(self):
        return self._read(A.FILM_COLOR)

    def albedo(self):
        return self._read(A.FILM_ALBEDO)

    def normal(self):
        return self._read(A.FILM_NORMAL)

    def xyz(self):
        """Mean colour in linear CIE XYZ: the inverse of the camera matrix
        (color.rs:48-55) applied to the white-balanced linear sRGB mean."""
        return self._read(A.FILM_XYZ)

    def xyz(self):
        return self._read(A.FILM_XYZ)

    def sums(self):
        return np.stack([self._read(k, sums=True) for k in (A.FILM_COLOR, A.FILM_ALBEDO, A.FILM_NORMAL)])

    def denoise(self):
        """À-trous of the film's colour with its albedo and normal guides, on the GPU (ag_film_denoise)."""
        out = np.empty((self.height, self.width, 3), np.float32)
        check(A.lib().ag_film_denoise(self.ctx.h, self.h, _ptr(out)))
        return out

    def develop(self, tonemap="agx_punchy", exposure_ev=0.0, saturation=1.0, denoise=False):
        """The live preview's development on the GPU (ag_film_develop): (h, w, 3) uint8."""
        out = np.empty((self.height, self.width, 3), np.uint8)
        g = grade(tonemap, exposure_ev, saturation)
        check(A.lib().ag_film_develop(self.ctx.h, self.h, C.byref(g), 1 if denoise else 0, _ptr(out, C.c_uint8)))
        return out


def grade(tonemap="agx_punchy", exposure_ev=0.0, saturation=1.0):
    """ag_grade; tonemap: agx, agx_punchy (the reference's default), agx_golden, aces, khronos, reinhard, clamp."""
    return A.Grade(exposure_ev, A.TONEMAPS[tonemap] if isinstance(tonemap, str) else int(tonemap), saturation)


class Live:
    """The live view's render worker (argent serve live.rs:637-830) on the GPU:
    half-resolution first look, ~70 ms refinement passes, paced frames,
    à-trous + develop on the device, settle test.  `now_ms` defaults to the
    monotonic clock."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        check(A.lib().ag_live_create(ctx.h, C.byref(h)))
        self.h = h
        self.size = (2, 2)

    def close(self):
        if self.h:
            A.lib().ag_live_destroy(self.h)
            self.h = None

    def attach(self, server):
        """Run this session's passes inside `server`'s steps (argent_amd.serve.Server); None detaches."""
        check(A.lib().ag_live_attach(self.h, server.h if server is not None else None))

    @staticmethod
    def _now(now_ms):
        return time.monotonic() * 1e3 if now_ms is None else float(now_ms)

    def set(self, scene, cam, width=540, height=960, max_spp=256, denoise=True, tonemap="agx_punchy",
            exposure_ev=0.0, saturation=1.0, ps=None, now_ms=None):
        """A change of scene, camera or size: a new generation."""
        p = A.LiveParams(width, height, max_spp, 1 if denoise else 0, grade(tonemap, exposure_ev, saturation),
                         (ps or PathSettings()).c())
        self.size = (int(width), int(height))
        self._scene = scene
        check(A.lib().ag_live_set(self.h, scene.h, C.byref(cam), C.byref(p), self._now(now_ms)))

    def set_look(self, tonemap="agx_punchy", exposure_ev=0.0, saturation=1.0, denoise=True, now_ms=None):
        """A look-only change: the last frame is re-developed."""
        g = grade(tonemap, exposure_ev, saturation)
        check(A.lib().ag_live_set_look(self.h, C.byref(g), 1 if denoise else 0, self._now(now_ms)))

    def step(self, now_ms=None):
        """One worker iteration: (frame dict, (h, w, 3) uint8 image or None)."""
        w, h = self.size
        out = np.empty(w * h * 3, np.uint8)
        f = A.LiveFrame()
        check(A.lib().ag_live_step(self.h, self._now(now_ms), C.byref(f), _ptr(out, C.c_uint8)))
        d = {k: getattr(f, k) for k, _ in A.LiveFrame._fields_}
        img = out[:f.width * f.height * 3].reshape(f.height, f.width, 3).copy() if f.published else None
        return d, img


def _weld_dict(w):
    return {k: getattr(w, k) for k, _ in A.WeldReport._fields_}


class Mesh:
    """A device-resident triangle mesh (argent-mesh lib.rs:33-39, without parts)
    with the reference's welding (weld.rs:38-62) and compaction (weld.rs:66-91)
    run on the GPU."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle

    @classmethod
    def from_arrays(cls, ctx, positions, triangles):
        p = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1, 3))
        t = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1, 3))
        h = C.c_void_p()
        check(A.lib().ag_mesh_create(ctx.h, _ptr(p), len(p), _ptr(t, C.c_uint32), len(t), C.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self.h:
            A.lib().ag_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def counts(self):
        nv, nt = C.c_uint32(), C.c_uint32()
        check(A.lib().ag_mesh_info(self.h, C.byref(nv), C.byref(nt)))
        return nv.value, nt.value

    def read(self, triangles=True):
        """(positions (nv, 3) float32, triangles (nt, 3) uint32 or None) in host memory."""
        nv, nt = self.counts
        p = np.empty((nv, 3), np.float32)
        t = np.empty((nt, 3), np.uint32) if triangles else None
        check(A.lib().ag_mesh_read(self.h, _ptr(p), _ptr(t, C.c_uint32) if triangles else None))
        return p, t

    def frame_sample(self, stride=None):
        """(hull (n, 3) float32: every stride-th vertex, lo (3,), hi (3,)) from
        the device, without copying the mesh back; stride defaults to the one
        Piece uses for its hull (max(n_vertices // 4000, 1))."""
        nv, _ = self.counts
        stride = max(nv // 4000, 1) if stride is None else int(stride)
        cap = (nv + stride - 1) // stride if stride > 0 else 0
        hull = np.empty((max(cap, 1), 3), np.float32)
        lo, hi = np.empty(3, np.float32), np.empty(3, np.float32)
        n = C.c_uint32()
        check(A.lib().ag_mesh_frame_sample(self.h, stride, _ptr(hull), cap, C.byref(n), _ptr(lo), _ptr(hi)))
        return hull[:n.value], lo, hi

    def weld(self, rel_tolerance=1e-6):
        """weld.rs:38-62; returns the WeldReport as a dict."""
        r = A.WeldReport()
        check(A.lib().ag_mesh_weld(self.h, rel_tolerance, C.byref(r)))
        return _weld_dict(r)

    def normals(self, crease_angle_deg=50.0):
        """normals::compute (normals.rs:47-49) on the GPU; returns the NormalReport as a dict."""
        r = A.NormalReport()
        check(A.lib().ag_mesh_normals(self.h, crease_angle_deg, C.byref(r)))
        return {k: getattr(r, k) for k, _ in A.NormalReport._fields_}

    def read_normals(self):
        nv, _ = self.counts
        n = np.empty((nv, 3), np.float32)
        check(A.lib().ag_mesh_read_normals(self.h, _ptr(n)))
        return n

    def compact(self):
        """weld.rs:66-91; returns the number of triangles dropped."""
        d = C.c_uint64()
        check(A.lib().ag_mesh_compact(self.h, C.byref(d)))
        return d.value


def load_stl(ctx, data, weld=True, weld_tolerance=1e-6, unit_scale=1.0, crease_angle_deg=50.0, device_ptr=None):
    """load_bytes (io.rs:87-126) for a binary STL: returns (Mesh, report dict).
    `data` is the file's bytes in host memory; for bytes already in device
    memory pass their address as device_ptr and their length as data."""
    o = A.LoadOptions(int(bool(weld)), weld_tolerance, crease_angle_deg, unit_scale)
    r = A.LoadReport()
    h = C.c_void_p()
    if device_ptr is not None:
        check(A.lib().ag_mesh_load_stl(ctx.h, C.c_void_p(device_ptr), int(data), 1, C.byref(o), C.byref(h),
                                       C.byref(r)))
    else:
        buf = bytes(data)
        check(A.lib().ag_mesh_load_stl(ctx.h, buf, len(buf), 0, C.byref(o), C.byref(h), C.byref(r)))
    rep = {k: getattr(r, k) for k, _ in A.LoadReport._fields_ if k != "weld"}
    rep["weld"] = _weld_dict(r.weld) if r.welded else None
    return Mesh(ctx, h), rep


def develop_image(ctx, img, tonemap="agx_punchy", exposure_ev=0.0, saturation=1.0):
    """develop (tonemap.rs:370-388) on the GPU: (h, w, 3) float32 -> (h, w, 3) uint8."""
    img = np.ascontiguousarray(np.asarray(img, np.float32))
    h, w = img.shape[:2]
    out = np.empty((h, w, 3), np.uint8)
    g = grade(tonemap, exposure_ev, saturation)
    check(A.lib().ag_develop_image(ctx.h, _ptr(img), w, h, C.byref(g), _ptr(out, C.c_uint8)))
    return out


def relative_change(ctx, a, b):
    """live.rs:208-218 on the GPU: mean |dY| / mean Y over every 4th pixel of two (h, w, 3) images."""
    a = np.ascontiguousarray(np.asarray(a, np.float32))
    b = np.ascontiguousarray(np.asarray(b, np.float32))
    if a.shape != b.shape:
        raise ValueError("images differ in size")
    h, w = a.shape[:2]
    out = C.c_double()
    check(A.lib().ag_relative_change(ctx.h, _ptr(a), _ptr(b), w, h, C.byref(out)))
    return out.value


def denoise_image(ctx, color, albedo=None, normal=None):
    """The à-trous denoiser (denoise.rs:660-709) on the GPU: (h, w, 3) float32 images."""
    color = np.ascontiguousarray(np.asarray(color, np.float32))
    h, w = color.shape[:2]
    a = None if albedo is None else np.ascontiguousarray(np.asarray(albedo, np.float32))
    n = None if normal is None else np.ascontiguousarray(np.asarray(normal, np.float32))
    out = np.empty_like(color)
    check(A.lib().ag_denoise_image(ctx.h, _ptr(color), _ptr(a) if a is not None else None,
                                   _ptr(n) if n is not None else None, w, h, _ptr(out)))
    return out


def look_at(pos, target, up_hint=(0.0, 1.0, 0.0), focal_mm=100.0, f_stop=0.0, aspect=1.0):
    cam = A.Camera()
    check(A.lib().ag_camera_look_at(_ptr(_f32(pos)), _ptr(_f32(target)), _ptr(_f32(up_hint)), focal_mm, f_stop,
                                    aspect, C.byref(cam)))
    return cam


def frame(points, view="hero", focal_mm=100.0, f_stop=0.0, aspect=1.0):
    az, el, fill = VIEWS[view] if isinstance(view, str) else view
    pts = _f32(points, (-1, 3))
    cam = A.Camera()
    check(A.lib().ag_camera_frame(_ptr(pts.reshape(-1)), len(pts), az, el, fill, focal_mm, f_stop, aspect,
                                  C.byref(cam)))
    return cam


def render_pass(ctx, scene, cam, film, ps, count, seed):
    """render.rs:67-115 on the GPU: add `count` samples per pixel to `film`."""
    check(A.lib().ag_render_pass(ctx.h, scene.h, C.byref(cam), film.h, C.byref(ps.c()), count, seed))


class Pass:
    """An enqueued render pass (ag_render_pass_async): done() polls, wait()
    blocks and fills the context statistics."""

    def __init__(self, h):
        self.h = h

    def done(self):
        rc = A.lib().ag_event_query(self.h)
        if rc == A.PENDING:
            return False
        check(rc)
        return True

    def wait(self):
        check(A.lib().ag_event_wait(self.h))

    def close(self):
        if self.h:
            A.lib().ag_event_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_pass_async(ctx, scene, cam, film, ps, count, seed):
    """render_pass, enqueued; returns a Pass.  Keep scene, camera and film
    alive until it is done."""
    h = C.c_void_p()
    check(A.lib().ag_render_pass_async(ctx.h, scene.h, C.byref(cam), film.h, C.byref(ps.c()), count, seed,
                                       C.byref(h)))
    return Pass(h)


def render_views(ctx, scene, views, ps, materials=()):
    """Many views of one scene in one wavefront schedule.  views: list of
    dicts with camera, film, count, seed, turntable_deg, material0 (index
    into `materials` or -1)."""
    arr = (A.View * len(views))()
    for i, v in enumerate(views):
        arr[i].camera = v["camera"]
        arr[i].turntable_deg = v.get("turntable_deg", 0.0)
        arr[i].material0 = v.get("material0", -1)
        arr[i].film = v["film"].h
        arr[i].count = v["count"]
        arr[i].seed = v["seed"]
        arr[i].scene = v["scene"].h if v.get("scene") is not None else None
    mats = (A.Material * max(1, len(materials)))(*materials) if materials else None
    check(A.lib().ag_render_views(ctx.h, scene.h, mats, len(materials), arr, len(views), C.byref(ps.c())))


class DeviceBytes:
    """Bytes copied into device memory (e.g. a file for ag_mesh_load_stl)."""

    def __init__(self, ctx, data):
        self.ctx, self.n = ctx, len(data)
        p = C.c_void_p()
        check(A.lib().ag_device_alloc(ctx.h, max(self.n, 4), C.byref(p)))
        self.p = p
        buf = (C.c_char * self.n).from_buffer_copy(bytes(data))
        check(A.lib().ag_memcpy_to_device(ctx.h, p, C.cast(buf, C.c_void_p), self.n))

    @property
    def ptr(self):
        return self.p.value

    def close(self):
        if self.p:
            A.lib().ag_device_free(self.ctx.h, self.p)
            self.p = None


class DeviceRays:
    """A ray batch resident in HBM (structure of arrays) with room for hits:
    the input of Scene.intersect_device / occluded_device."""

    def __init__(self, ctx, o, d, tmax=None):
        o, d = _f32(o, (-1, 3)), _f32(d, (-1, 3))
        self.ctx, self.n = ctx, len(o)
        cols = [np.ascontiguousarray(o[:, k]) for k in range(3)] + [np.ascontiguousarray(d[:, k]) for k in range(3)]
        if tmax is not None:
            cols.append(_f32(tmax))
        self.bufs = []
        for c in cols:
            p = C.c_void_p()
            check(A.lib().ag_device_alloc(ctx.h, c.nbytes, C.byref(p)))
            check(A.lib().ag_memcpy_to_device(ctx.h, p, c.ctypes.data_as(C.c_void_p), c.nbytes))
            self.bufs.append(p)
        self.out = []
        for _ in range(4):
            p = C.c_void_p()
            check(A.lib().ag_device_alloc(ctx.h, 4 * self.n, C.byref(p)))
            self.out.append(p)
        f = lambda p: C.cast(p, A.f32p)
        self.rays = A.RaysSoA(*[f(p) for p in self.bufs[:6]], f(self.bufs[6]) if tmax is not None else None)
        self.hits = A.HitsSoA(f(self.out[0]), f(self.out[1]), f(self.out[2]), C.cast(self.out[3], A.u32p))

    def hits_host(self):
        res = []
        for k, dt in enumerate((np.float32, np.float32, np.float32, np.uint32)):
            a = np.empty(self.n, dt)
            check(A.lib().ag_memcpy_to_host(self.ctx.h, a.ctypes.data_as(C.c_void_p), self.out[k], a.nbytes))
            res.append(a)
        return tuple(res)

    def close(self):
        for p in self.bufs + self.out:
            A.lib().ag_device_free(self.ctx.h, p)
        self.bufs, self.out = [], []


def intersect_device(ctx, scene, rays):
    """Closest hit for a resident batch; returns the kernel time in ms (HIP events)."""
    check(A.lib().ag_intersect_device(ctx.h, scene.h, C.byref(rays.rays), rays.n, C.byref(rays.hits)))
    return A.lib().ag_last_query_ms(ctx.h)


def occluded_device(ctx, scene, rays):
    check(A.lib().ag_occluded_device(ctx.h, scene.h, C.byref(rays.rays), rays.n, C.cast(rays.out[3], A.u8p)))
    return A.lib().ag_last_query_ms(ctx.h)
