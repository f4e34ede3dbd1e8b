"""ctypes mirror of include/argent.h and the loader of libargent_hip.so.

The library is loaded from this package directory; a missing or unloadable
library raises at import of the native API — there is no fallback path.
"""
import ctypes as C
import os

SPEC_N = 471
GRID_N = 95
MAX_PANELS = 16
PENDING = 6  # AG_PENDING
MAX_MATERIALS = 32

MAT_METAL, MAT_GEM, MAT_GEM_OPAQUE, MAT_PLASTIC, MAT_DIFFUSE, MAT_BLACK, MAT_OPAL = range(7)
ENV_GRADIENT, ENV_MAP = 0, 1
FILM_COLOR, FILM_ALBEDO, FILM_NORMAL, FILM_XYZ = range(4)

f32p = C.POINTER(C.c_float)
u32p = C.POINTER(C.c_uint32)
u16p = C.POINTER(C.c_uint16)
u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)


class Material(C.Structure):
    _fields_ = [("kind", C.c_int32), ("alpha", C.c_float), ("eta", C.c_float),
                ("metal_n", C.c_float * GRID_N), ("metal_k", C.c_float * GRID_N),
                ("has_patina", C.c_int32), ("patina_strength", C.c_float), ("patina_lo", C.c_float),
                ("patina_hi", C.c_float), ("patina_alpha", C.c_float), ("patina_eta", C.c_float),
                ("patina_base", C.c_float * SPEC_N), ("spec_a", C.c_float * SPEC_N),
                ("spec_b", C.c_float * SPEC_N), ("opal_cell_mm", C.c_float), ("opal_strength", C.c_float)]


class Panel(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("hu", C.c_float * 3), ("hv", C.c_float * 3),
                ("has_emission", C.c_int32), ("camera_visible", C.c_int32), ("feather", C.c_float),
                ("emission", C.c_float * SPEC_N)]


class Env(C.Structure):
    _fields_ = [("kind", C.c_int32), ("top", C.c_float * SPEC_N), ("horizon", C.c_float * SPEC_N),
                ("bottom", C.c_float * SPEC_N), ("power", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("weights", f32p), ("marginal", f32p), ("conditional", f32p),
                ("total", C.c_float), ("rotation_deg", C.c_float), ("basis", (C.c_float * SPEC_N) * 3)]


class Floor(C.Structure):
    _fields_ = [("present", C.c_int32), ("y", C.c_float), ("material", C.c_int32),
                ("fade_start", C.c_float), ("fade_end", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("positions", f32p), ("normals", f32p), ("triangles", u32p), ("tri_material", u16p),
                ("cavity", f32p), ("n_vertices", C.c_uint32), ("n_triangles", C.c_uint32),
                ("n_materials", C.c_int32), ("materials", Material * MAX_MATERIALS),
                ("n_panels", C.c_int32), ("panels", Panel * MAX_PANELS), ("env", Env), ("floor", Floor),
                ("has_white", C.c_int32), ("white_xyz", C.c_float * 3), ("turntable_deg", C.c_float),
                ("energy", C.c_float * 1024)]


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("half_w", C.c_float), ("half_h", C.c_float),
                ("lens_radius", C.c_float), ("focus_dist", C.c_float), ("shift", C.c_float * 2)]


class PathSettingsC(C.Structure):
    _fields_ = [("max_depth", C.c_uint32), ("clamp_indirect", C.c_float), ("rr_depth", C.c_uint32)]


class Grade(C.Structure):
    """ag_grade (tonemap.rs:53-58)."""
    _fields_ = [("exposure_ev", C.c_float), ("tonemap", C.c_int32), ("saturation", C.c_float)]


TONEMAPS = {"agx": 0, "agx_punchy": 1, "agx_golden": 2, "aces": 3, "khronos": 4, "reinhard": 5, "clamp": 6}


class LiveParams(C.Structure):
    """ag_live_params (params.rs: preview size, max_spp, denoise, grade, path settings)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("max_spp", C.c_uint32), ("denoise", C.c_int32),
                ("grade", Grade), ("ps", PathSettingsC)]


class LiveFrame(C.Structure):
    """ag_live_frame."""
    _fields_ = [("published", C.c_int32), ("idle", C.c_int32), ("generation", C.c_uint64), ("width", C.c_uint32),
                ("height", C.c_uint32), ("spp", C.c_uint32), ("max_spp", C.c_uint32), ("half", C.c_int32),
                ("converged", C.c_int32), ("settled", C.c_int32), ("denoised", C.c_int32),
                ("redeveloped", C.c_int32), ("pass_spp", C.c_uint32), ("pass_ms", C.c_float),
                ("ms_per_spp", C.c_float), ("relative_change", C.c_float)]


class LoadOptions(C.Structure):
    """ag_load_options (argent-mesh lib.rs:53-74)."""
    _fields_ = [("weld", C.c_int32), ("weld_tolerance", C.c_float), ("crease_angle_deg", C.c_float),
                ("unit_scale", C.c_float)]


class WeldReport(C.Structure):
    """ag_weld_report (weld.rs:21-33)."""
    _fields_ = [("input_vertices", C.c_uint64), ("output_vertices", C.c_uint64), ("exact_merges", C.c_uint64),
                ("tolerance_merges", C.c_uint64), ("dropped_triangles", C.c_uint64), ("tolerance_mm", C.c_float)]


class NormalReport(C.Structure):
    """ag_normal_report (normals.rs:23-42)."""
    _fields_ = [(k, C.c_uint64) for k in ("shells", "flipped_shells", "cavity_shells", "repaired_faces",
                                           "boundary_edges", "nonmanifold_edges", "inconsistent_edges",
                                           "degenerate_faces", "split_vertices")]


class LoadReport(C.Structure):
    """ag_load_report (io.rs:21-40, binary STL)."""
    _fields_ = [("file_bytes", C.c_uint64), ("raw_triangles", C.c_uint64), ("raw_vertices", C.c_uint64),
                ("invalid_triangles", C.c_uint64), ("trailing_bytes", C.c_uint64),
                ("declared_triangles", C.c_uint64), ("welded", C.c_int32), ("weld", WeldReport),
                ("upload_ms", C.c_float), ("parse_ms", C.c_float), ("weld_ms", C.c_float)]


class RaysSoA(C.Structure):
    _fields_ = [("ox", f32p), ("oy", f32p), ("oz", f32p), ("dx", f32p), ("dy", f32p), ("dz", f32p),
                ("tmax", f32p)]


class HitsSoA(C.Structure):
    _fields_ = [("t", f32p), ("u", f32p), ("v", f32p), ("prim", u32p)]


class View(C.Structure):
    _fields_ = [("camera", Camera), ("turntable_deg", C.c_float), ("material0", C.c_int32),
                ("film", C.c_void_p), ("count", C.c_uint32), ("seed", C.c_uint32), ("scene", C.c_void_p)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("turntable_deg", C.c_float),
                ("has_material0", C.c_int32), ("material0", Material), ("width", C.c_uint32),
                ("height", C.c_uint32), ("spp", C.c_uint32), ("seed", C.c_uint32), ("preview_spp", C.c_uint32)]


class RequestStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("samples", C.c_uint32), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_first_preview", C.c_double), ("t_done", C.c_double)]


REQ_QUEUED, REQ_ACTIVE, REQ_DONE, REQ_REJECTED, REQ_FAILED = range(5)


class Tuning(C.Structure):
    _fields_ = [("paths_per_wave", C.c_uint32), ("block_traverse", C.c_uint32),
                ("blocks_per_cu", C.c_uint32), ("sort_by_material", C.c_uint32),
                ("refill_min", C.c_uint32), ("batch_blocks_per_cu", C.c_uint32), ("batch_refill_min", C.c_uint32),
                ("shadow_refill_min", C.c_uint32), ("spread_min_lanes", C.c_uint32),
                ("ordered_emission", C.c_uint32), ("cursor_slices", C.c_uint32),
                ("launch_graph", C.c_uint32), ("shade_waves", C.c_uint32), ("raygen_bands", C.c_uint32)]


class PatinaSettings(C.Structure):  # ag_patina_settings (cavity.rs:24-42)
    _fields_ = [("ao_samples", C.c_uint32), ("ao_distance", C.c_float), ("curvature_radius", C.c_float),
                ("concavity_weight", C.c_float), ("smooth_iterations", C.c_uint32)]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("ms_start", C.c_double), ("ms_best", C.c_double), ("trials", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("ms_total", C.c_double), ("ms_extend", C.c_double), ("ms_shade", C.c_double),
                ("ms_shadow", C.c_double), ("ms_raygen", C.c_double), ("ms_accum", C.c_double),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("samples", C.c_uint64),
                ("extend_launches", C.c_uint32), ("shadow_launches", C.c_uint32)]


LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libargent_hip.so")

# (name, restype, argtypes)
_SIGS = [
    ("ag_last_error", C.c_char_p, []),
    ("ag_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("ag_ctx_destroy", None, [C.c_void_p]),
    ("ag_ctx_synchronize", C.c_int, [C.c_void_p]),
    ("ag_material_metal", C.c_int, [C.c_char_p, C.c_float, C.c_float, C.POINTER(Material)]),
    ("ag_material_gem", C.c_int, [C.c_char_p, C.POINTER(Material)]),
    ("ag_material_plastic", C.c_int, [f32p, C.c_float, C.c_float, C.POINTER(Material)]),
    ("ag_material_diffuse", C.c_int, [f32p, C.POINTER(Material)]),
    ("ag_alloy_ior", C.c_int, [C.c_char_p, f32p, f32p]),
    ("ag_alloy_f0_srgb", C.c_int, [C.c_char_p, f32p]),
    ("ag_karat_gold", C.c_int, [C.c_float, C.c_float, C.c_float, f32p, f32p]),
    ("ag_gem_optics", C.c_int, [C.c_char_p, i32p, f32p, f32p, f32p, i32p]),
    ("ag_spectrum_emitter", C.c_int, [C.c_float, C.c_float, f32p]),
    ("ag_spectrum_rgb_albedo", C.c_int, [f32p, f32p]),
    ("ag_spectrum_rgb_illuminant", C.c_int, [f32p, f32p]),
    ("ag_luminance_y", C.c_float, [f32p]),
    ("ag_blackbody_white", C.c_int, [C.c_float, f32p]),
    ("ag_camera_matrix", C.c_int, [C.c_int32, f32p, f32p]),
    ("ag_cie_xyz_bar", C.c_int, [C.c_float, f32p]),
    ("ag_reflectance_srgb", C.c_int, [f32p, f32p]),
    ("ag_fresnel_conductor1", C.c_float, [C.c_float, C.c_float, C.c_float]),
    ("ag_fresnel_dielectric", C.c_float, [C.c_float, C.c_float]),
    ("ag_energy_table", C.c_int, [f32p]),
    ("ag_desc_init", C.c_int, [C.POINTER(SceneDesc)]),
    ("ag_desc_apply_studio", C.c_int, [C.POINTER(SceneDesc), C.c_char_p, C.c_float, C.c_float]),
    ("ag_env_from_rgb", C.c_int, [f32p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.POINTER(Env)]),
    ("ag_env_free", None, [C.POINTER(Env)]),
    ("ag_desc_light_cdf", C.c_int, [C.POINTER(SceneDesc), i32p, i32p, f32p, f32p]),
    ("ag_camera_look_at", C.c_int, [f32p, f32p, f32p, C.c_float, C.c_float, C.c_float, C.POINTER(Camera)]),
    ("ag_camera_frame", C.c_int, [f32p, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                  C.c_float, C.POINTER(Camera)]),
    ("ag_scene_create", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("ag_scene_create_from_mesh", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SceneDesc), C.c_uint32,
                                            C.POINTER(C.c_void_p)]),
    ("ag_scene_create_ex", C.c_int, [C.c_void_p, C.POINTER(SceneDesc), C.c_uint32, C.POINTER(C.c_void_p)]),
    ("ag_scene_create_shared", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]),
    ("ag_scene_destroy", None, [C.c_void_p]),
    ("ag_scene_set_turntable", C.c_int, [C.c_void_p, C.c_float]),
    ("ag_scene_set_material", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(Material)]),
    ("ag_scene_bounds", C.c_int, [C.c_void_p, f32p, f32p]),
    ("ag_scene_stats", C.c_int, [C.c_void_p, u32p, u32p, u32p, C.POINTER(C.c_size_t)]),
    ("ag_film_create", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("ag_film_destroy", None, [C.c_void_p]),
    ("ag_film_clear", C.c_int, [C.c_void_p]),
    ("ag_film_read", C.c_int, [C.c_void_p, C.c_int32, f32p]),
    ("ag_develop_image", C.c_int, [C.c_void_p, f32p, C.c_uint32, C.c_uint32, C.POINTER(Grade), u8p]),
    ("ag_denoise_image", C.c_int, [C.c_void_p, f32p, f32p, f32p, C.c_uint32, C.c_uint32, f32p]),
    ("ag_film_denoise", C.c_int, [C.c_void_p, C.c_void_p, f32p]),
    ("ag_film_develop", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Grade), C.c_int32, u8p]),
    ("ag_film_read_sums", C.c_int, [C.c_void_p, C.c_int32, f32p]),
    ("ag_film_samples", C.c_uint32, [C.c_void_p]),
    ("ag_render_pass", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_void_p,
                                 C.POINTER(PathSettingsC), C.c_uint32, C.c_uint32]),
    ("ag_render_pass_async", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_void_p, C.POINTER(PathSettingsC),
                                       C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("ag_event_query", C.c_int, [C.c_void_p]),
    ("ag_event_wait", C.c_int, [C.c_void_p]),
    ("ag_event_destroy", None, [C.c_void_p]),
    ("ag_intersect", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32, C.POINTER(HitsSoA)]),
    ("ag_occluded", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32, u8p]),
    ("ag_ambient_occlusion", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, f32p]),
    ("ag_patina_default", C.c_int, [C.POINTER(PatinaSettings)]),
    ("ag_ctx_graph_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("ag_patina_for_scene", C.c_int, [C.c_void_p, C.POINTER(PatinaSettings)]),
    ("ag_patina_mask", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(PatinaSettings), C.c_int32, f32p, f32p, f32p]),
    ("ag_smooth_values", C.c_int, [C.c_void_p, C.c_void_p, f32p, C.c_uint32, C.c_float]),
    ("ag_intersect_device", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32,
                                      C.POINTER(HitsSoA)]),
    ("ag_intersect_device_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32,
                                              C.POINTER(HitsSoA)]),
    ("ag_occluded_device", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32, u8p]),
    ("ag_bsdf_probe", C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, i32p, f32p, f32p, f32p, f32p, f32p, f32p]),
    ("ag_render_views", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Material), C.c_int32, C.POINTER(View),
                                  C.c_int32, C.POINTER(PathSettingsC)]),
    ("ag_server_create", C.c_int, [C.c_void_p, C.c_size_t, C.c_uint64, C.POINTER(PathSettingsC),
                                   C.POINTER(C.c_void_p)]),
    ("ag_server_destroy", None, [C.c_void_p]),
    ("ag_server_submit", C.c_int, [C.c_void_p, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    ("ag_server_step", C.c_int, [C.c_void_p, u32p, u32p]),
    ("ag_server_status", C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(RequestStatus)]),
    ("ag_server_read", C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, f32p]),
    ("ag_server_release", C.c_int, [C.c_void_p, C.c_uint64]),
    ("ag_server_error", C.c_char_p, [C.c_void_p, C.c_uint64]),
    ("ag_server_budget", C.c_int, [C.c_void_p, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    ("ag_debug_fail_seed", C.c_int, [C.c_void_p, C.c_int64]),
    ("ag_device_alloc", C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    ("ag_device_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("ag_memcpy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("ag_memcpy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    ("ag_last_query_ms", C.c_double, [C.c_void_p]),
    ("ag_ctx_get_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("ag_ctx_set_tuning", C.c_int, [C.c_void_p, C.POINTER(Tuning)]),
    ("ag_autotune_render", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(PathSettingsC), C.c_uint32, C.POINTER(AutotuneResult)]),
    ("ag_autotune_rays", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(RaysSoA), C.c_uint32, C.POINTER(HitsSoA),
                                   C.c_uint32, C.POINTER(AutotuneResult)]),
    ("ag_ctx_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("ag_ctx_enable_timing", C.c_int, [C.c_void_p, C.c_int32]),
    ("ag_mesh_load_stl", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int32, C.POINTER(LoadOptions),
                                   C.POINTER(C.c_void_p), C.POINTER(LoadReport)]),
    ("ag_mesh_create", C.c_int, [C.c_void_p, f32p, C.c_uint32, u32p, C.c_uint32, C.POINTER(C.c_void_p)]),
    ("ag_mesh_destroy", None, [C.c_void_p]),
    ("ag_mesh_info", C.c_int, [C.c_void_p, u32p, u32p]),
    ("ag_mesh_read", C.c_int, [C.c_void_p, f32p, u32p]),
    ("ag_mesh_frame_sample", C.c_int, [C.c_void_p, C.c_uint32, f32p, C.c_uint32, C.POINTER(C.c_uint32), f32p, f32p]),
    ("ag_mesh_device_arrays", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("ag_mesh_weld", C.c_int, [C.c_void_p, C.c_float, C.POINTER(WeldReport)]),
    ("ag_mesh_compact", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    ("ag_mesh_normals", C.c_int, [C.c_void_p, C.c_float, C.POINTER(NormalReport)]),
    ("ag_mesh_read_normals", C.c_int, [C.c_void_p, f32p]),
    ("ag_live_create", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("ag_live_destroy", None, [C.c_void_p]),
    ("ag_live_attach", C.c_int, [C.c_void_p, C.c_void_p]),
    ("ag_live_set", C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Camera), C.POINTER(LiveParams), C.c_double]),
    ("ag_live_set_look", C.c_int, [C.c_void_p, C.POINTER(Grade), C.c_int32, C.c_double]),
    ("ag_live_step", C.c_int, [C.c_void_p, C.c_double, C.POINTER(LiveFrame), u8p]),
    ("ag_relative_change", C.c_int, [C.c_void_p, f32p, f32p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    ("ag_ctx_depth_stats", C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def lib():
    """Load libargent_hip.so (built in-tree by argent_amd.build); raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class ArgentError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        msg = lib().ag_last_error()
        raise ArgentError(f"argent error {rc}: {msg.decode() if msg else ''}")
    return rc
