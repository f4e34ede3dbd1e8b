/*
 * argent.h — C ABI of the MI355X-native argent render core.
 *
 * The reference's seam is a Rust function call: argent_core::render::render_pass
 * (crates/argent-core/src/render.rs:67), with Bvh::build / intersect / occluded
 * (crates/argent-core/src/bvh.rs:117, :159, :216) for geometry queries and
 * Scene::new (crates/argent-core/src/scene.rs:82) for scene construction.  Each
 * entry point below names the reference item it replaces.  All functions are
 * synchronous unless named *_async, return 0 on success and a nonzero AG_E_*
 * code on failure; ag_last_error() returns a thread-local message for the last
 * failure on the calling thread.
 *
 * Units follow the reference: millimetres, nanometres, degrees.  Spectra are
 * tabulated at 1 nm from 360 to 830 nm (471 values, argent-core tab.rs:8-32);
 * metal optical constants on the 5 nm grid 360..830 (95 rows,
 * argent-spectra metals.rs:53-56).
 */
#ifndef ARGENT_H
#define ARGENT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AG_SPEC_N 471  /* 1 nm table, 360..830 nm */
#define AG_GRID_N 95   /* 5 nm grid, 360..830 nm */
#define AG_MAX_PANELS 16
#define AG_MAX_MATERIALS 32
/* Largest ag_tuning.paths_per_wave: 2^29 path slots (about 150 GB of path
 * state and queues) hold a whole 1080 x 1920 x 256-spp still in one wave. */
#define AG_MAX_PATHS_PER_WAVE (1u << 29)

enum {
    AG_OK = 0,
    AG_E_INVALID = 1,   /* bad argument or descriptor */
    AG_E_DEVICE = 2,    /* HIP runtime error */
    AG_E_NOMEM = 3,     /* allocation failed or over budget */
    AG_E_NOTFOUND = 4,  /* unknown name */
    AG_E_STATE = 5,     /* call not valid in the object's state */
    AG_PENDING = 6      /* ag_event_query: the pass has not finished */
};

/* ---------------------------------------------------------------- materials
 * argent-core material.rs:27-36 (Material) and bsdf.rs:10-27 (Bsdf). */
enum {
    AG_MAT_METAL = 0,    /* Conductor: GGX + exact complex Fresnel + multi-scatter */
    AG_MAT_GEM = 1,      /* transparent dispersive dielectric with Beer-Lambert */
    AG_MAT_GEM_OPAQUE = 2, /* opaque stone: coated, GGX r 0.06, eta = n(550) */
    AG_MAT_PLASTIC = 3,  /* coated diffuse */
    AG_MAT_DIFFUSE = 4,
    AG_MAT_BLACK = 5,
    AG_MAT_OPAL = 6
};

typedef struct ag_material {
    int32_t kind;            /* AG_MAT_* */
    float alpha;             /* GGX alpha = max(r^2, 1e-4) (microfacet.rs:22-25) */
    float eta;               /* coat index (plastic, opaque gem) */
    /* metal: n and k on the 5 nm grid (metals.rs ComplexIor) */
    float metal_n[AG_GRID_N];
    float metal_k[AG_GRID_N];
    /* patina layer of a metal (material.rs:15-24); has_patina = 0 disables */
    int32_t has_patina;
    float patina_strength, patina_lo, patina_hi, patina_alpha, patina_eta;
    float patina_base[AG_SPEC_N];
    /* spec_a: gem ior / plastic base / diffuse albedo / opal body (1 nm table)
     * spec_b: gem absorption sigma_a in 1/mm (1 nm table) */
    float spec_a[AG_SPEC_N];
    float spec_b[AG_SPEC_N];
    /* opal (material.rs:33, :62-66) */
    float opal_cell_mm, opal_strength;
} ag_material;

/* ---------------------------------------------------------------- lights
 * argent-core light.rs:8-98 (Panel).  has_emission = 0 is a flag (black card). */
typedef struct ag_panel {
    float center[3], hu[3], hv[3];
    int32_t has_emission;
    int32_t camera_visible;
    float feather;
    float emission[AG_SPEC_N];
} ag_panel;

enum { AG_ENV_GRADIENT = 0, AG_ENV_MAP = 1 };

/* light.rs:110-127 (Env::Gradient) and envmap.rs:107-125 (EnvData / EnvMap).
 * For a map, weights are the three LinearRgb basis weights per texel
 * (envmap.rs:147), marginal has height+1 entries and conditional
 * height*(width+1) (envmap.rs:149-175).  basis holds the three 1 nm basis
 * spectra of LinearRgb (envmap.rs:31-47). */
typedef struct ag_env {
    int32_t kind;
    float top[AG_SPEC_N], horizon[AG_SPEC_N], bottom[AG_SPEC_N];
    float power;
    int32_t width, height;
    const float* weights;      /* width*height*3 */
    const float* marginal;     /* height+1 */
    const float* conditional;  /* height*(width+1) */
    float total;
    float rotation_deg;
    float basis[3][AG_SPEC_N];
} ag_env;

/* light.rs:130-162 (Floor). */
typedef struct ag_floor {
    int32_t present;
    float y;
    int32_t material;
    float fade_start, fade_end;
} ag_floor;

/* scene.rs:47-64 (Scene) with scene.rs:10-19 (Piece).  Arrays are host memory;
 * ag_scene_create copies them.  normals and cavity may be NULL. */
typedef struct ag_scene_desc {
    const float* positions;      /* n_vertices*3 */
    const float* normals;        /* n_vertices*3 or NULL */
    const uint32_t* triangles;   /* n_triangles*3 */
    const uint16_t* tri_material;/* n_triangles */
    const float* cavity;         /* n_vertices or NULL */
    uint32_t n_vertices, n_triangles;
    int32_t n_materials;
    ag_material materials[AG_MAX_MATERIALS];
    int32_t n_panels;
    ag_panel panels[AG_MAX_PANELS];
    ag_env env;
    ag_floor floor;
    int32_t has_white;
    float white_xyz[3];
    float turntable_deg;
    /* 32x32 GGX energy table E(mu, sqrt(alpha)) (microfacet.rs:93-140) */
    float energy[32 * 32];
} ag_scene_desc;

/* camera.rs:15-27 (Camera), 15 floats. */
typedef struct ag_camera {
    float pos[3], forward[3], right[3], up[3];
    float half_w, half_h, lens_radius, focus_dist;
    float shift[2];
} ag_camera;

/* Camera::look_at (camera.rs:38-55): 36 mm long side, lens radius
 * 0.5 * focal / f_stop (f_stop <= 0 disables depth of field). */
int ag_camera_look_at(const float pos[3], const float target[3], const float up_hint[3], float focal_mm,
                      float f_stop, float aspect, ag_camera* out);
/* views.rs:45-83 (frame): a camera along the view direction (azimuth from +Z
 * toward +X, elevation above the horizon) that fits the points (n x 3) to
 * `fill` of the frame, centred. */
int ag_camera_frame(const float* points, uint32_t n, float azimuth_deg, float elevation_deg, float fill,
                    float focal_mm, float f_stop, float aspect, ag_camera* out);

/* integrator.rs:10-24 (PathSettings). */
typedef struct ag_path_settings {
    uint32_t max_depth;
    float clamp_indirect;
    uint32_t rr_depth;
} ag_path_settings;

/* Ray batches in structure-of-arrays form (bvh.rs:159 Ray + t_max). */
typedef struct ag_rays_soa {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
} ag_rays_soa;

/* Closest-hit output (bvh.rs TriHit): prim = original triangle index,
 * 0xFFFFFFFF on a miss. */
typedef struct ag_hits_soa {
    float *t, *u, *v;
    uint32_t* prim;
} ag_hits_soa;

typedef struct ag_ctx ag_ctx;
typedef struct ag_scene ag_scene;
typedef struct ag_film ag_film;

const char* ag_last_error(void);

/* One context per device: a HIP stream and the device's scratch. */
int ag_ctx_create(int device, ag_ctx** out);
void ag_ctx_destroy(ag_ctx* ctx);
int ag_ctx_synchronize(ag_ctx* ctx);

/* ------------------------------------------------------------ scene build
 * Host-side builders that fill descriptor parts the way the reference's
 * constructors do.  They bake spectra to the 1 nm tables the kernels read. */

/* Material::metal(alloy, roughness).with_patina(default_patina(strength, alloy))
 * (material.rs:39-48, :142-149).  patina_strength < 0 leaves has_patina = 0. */
int ag_material_metal(const char* alloy, float roughness, float patina_strength, ag_material* out);
/* Material::gem(name) (material.rs:50-60). */
int ag_material_gem(const char* name, ag_material* out);
/* Material::plastic / ::diffuse (material.rs:73-79). */
int ag_material_plastic(const float rgb[3], float eta, float roughness, ag_material* out);
int ag_material_diffuse(const float rgb[3], ag_material* out);
/* Alloy::ior (alloys.rs:142-171) on the 5 nm grid; F0 linear sRGB (alloys.rs:174-176). */
int ag_alloy_ior(const char* alloy, float n_out[AG_GRID_N], float k_out[AG_GRID_N]);
int ag_alloy_f0_srgb(const char* alloy, float rgb_out[3]);
/* karat_gold(au, ag, cu) (alloys.rs:423-460). */
int ag_karat_gold(float au, float ag, float cu, float n_out[AG_GRID_N], float k_out[AG_GRID_N]);
/* Gem optics (gems.rs:395-440): kind 0 transparent, 1 opal, 2 opaque. */
int ag_gem_optics(const char* name, int32_t* kind, float ior_out[AG_SPEC_N],
                  float abs_out[AG_SPEC_N], float body_rgb_out[3], int32_t* has_body);
/* studio.rs:15-25 emitter(kelvin, intensity), baked. */
int ag_spectrum_emitter(float kelvin, float intensity, float out[AG_SPEC_N]);
/* RgbAlbedo::new (spectrum.rs:194-203) and RgbIlluminant::new (spectrum.rs:395-402), baked. */
int ag_spectrum_rgb_albedo(const float rgb[3], float out[AG_SPEC_N]);
int ag_spectrum_rgb_illuminant(const float rgb[3], float out[AG_SPEC_N]);
/* luminance_y over a 1 nm table (spectrum.rs:413-420). */
float ag_luminance_y(const float spec[AG_SPEC_N]);
/* color.rs:32-45 blackbody_white and color.rs:48-55 camera_matrix (row-major 3x3). */
int ag_blackbody_white(float kelvin, float xyz_out[3]);
int ag_camera_matrix(int32_t has_white, const float white_xyz[3], float m_out[9]);
/* CIE helpers (cie.rs:38-49, :154-172). */
int ag_cie_xyz_bar(float lambda, float out[3]);
int ag_reflectance_srgb(const float spec[AG_SPEC_N], float rgb_out[3]);
/* Complex Fresnel of one wavelength and the dielectric form (fresnel.rs:9-25, :81-93). */
float ag_fresnel_conductor1(float cos_i, float eta, float k);
float ag_fresnel_dielectric(float cos_i, float eta);
/* The GGX energy table (microfacet.rs:112-140) into out[32*32]. */
int ag_energy_table(float out[32 * 32]);

/* Scene descriptor defaults: zeroed, gradient env of zero, no floor, energy
 * table filled.  Studio::build(radius) + floor_at (studio.rs:87-190, :194-197)
 * fill panels, env, floor and white; the floor material is appended to the
 * material list as Prepared::scene does (shot.rs:226-230).  studio is
 * "noir" | "white" | "velvet" | "ember" | "window". */
int ag_desc_init(ag_scene_desc* d);
int ag_desc_apply_studio(ag_scene_desc* d, const char* studio, float radius, float floor_y);
/* Equirectangular HDR environment from linear RGB texels (EnvData::load,
 * envmap.rs:141-178, without the file decode).  The arrays are owned by the
 * library until ag_env_free. */
int ag_env_from_rgb(const float* rgb, int32_t width, int32_t height, float exposure_ev,
                    float rotation_deg, ag_env* out);
void ag_env_free(ag_env* env);
/* Rebuilds the scene's light CDF (scene.rs:88-110); exposed for tests. */
int ag_desc_light_cdf(const ag_scene_desc* d, int32_t* n_lights, int32_t lights[AG_MAX_PANELS],
                      float cdf[AG_MAX_PANELS], float* env_pick);

/* Scene::new + Piece::new (scene.rs:21-32, :82-86): copies the arrays to the
 * device and builds the BVH.  The scene owns its device memory. */
int ag_scene_create(ag_ctx* ctx, const ag_scene_desc* desc, ag_scene** out);
/* Scene::new with a choice of BVH builder.  flags: AG_SCENE_GPU_BVH builds the
 * tree on the device (Morton-ordered radix tree collapsed to the same 8-wide
 * format: about two orders of magnitude faster to build than the host's
 * binned SAH, somewhat slower to traverse), for a new mesh whose first
 * picture matters more than the render's last percent.  0 = ag_scene_create.
 * Closest-hit results are the same with either tree. */
enum { AG_SCENE_GPU_BVH = 1 };
int ag_scene_create_ex(ag_ctx* ctx, const ag_scene_desc* desc, uint32_t flags, ag_scene** out);
/* Scene::new over a device mesh (see ag_mesh_load_stl below): the mesh's
 * positions, triangles and normals (if ag_mesh_normals ran) are copied on the
 * device, every triangle takes material slot 0, and the BVH is built on the
 * device with AG_SCENE_GPU_BVH (else on the host after a read-back).
 * Materials, panels, environment, floor and white come from `desc`, whose
 * mesh arrays are ignored.  The mesh may be destroyed afterwards. */
typedef struct ag_mesh ag_mesh;
int ag_scene_create_from_mesh(ag_ctx* ctx, const ag_mesh* mesh, const ag_scene_desc* desc, uint32_t flags,
                              ag_scene** out);
/* Scene::new(Arc<Piece>, ...) over the piece of `base` (scene.rs:48, :82-86;
 * jobs.rs:690-696 and shot.rs:206-221 clone the Arc per look): materials,
 * panels, environment, floor and white come from `desc`, the mesh and BVH are
 * shared with `base` (its mesh arrays in `desc` are ignored) and stay alive
 * until the last scene over them is destroyed. */
int ag_scene_create_shared(ag_ctx* ctx, const ag_scene* base, const ag_scene_desc* desc, ag_scene** out);
void ag_scene_destroy(ag_scene* s);
/* Scene::turntable_deg (scene.rs:53-55). */
int ag_scene_set_turntable(ag_scene* s, float deg);
/* Replace material slot m (metal variants without a rebuild). */
int ag_scene_set_material(ag_scene* s, int32_t m, const ag_material* mat);
/* Bvh::bounds (bvh.rs:145-148) and statistics of the device BVH. */
int ag_scene_bounds(const ag_scene* s, float lo[3], float hi[3]);
int ag_scene_stats(const ag_scene* s, uint32_t* nodes, uint32_t* leaves, uint32_t* depth, size_t* bytes);

/* ------------------------------------------------------------ film
 * render.rs:13-64 (Film): colour, albedo and normal running sums plus a
 * uniform sample count; storage is device memory. */
enum { AG_FILM_COLOR = 0, AG_FILM_ALBEDO = 1, AG_FILM_NORMAL = 2, AG_FILM_XYZ = 3 };
int ag_film_create(ag_ctx* ctx, uint32_t width, uint32_t height, ag_film** out);
void ag_film_destroy(ag_film* f);
int ag_film_clear(ag_film* f);
/* Averages as Film::color/albedo/normal do (render.rs:41-64); AG_FILM_XYZ
 * returns M^-1 * colour.  out holds width*height*3 floats, host memory. */
int ag_film_read(const ag_film* f, int32_t which, float* out);
/* Raw running sums (no division), host memory, width*height*3 floats. */
int ag_film_read_sums(const ag_film* f, int32_t which, float* out);
uint32_t ag_film_samples(const ag_film* f);

/* ------------------------------------------------------------ rendering
 * render_pass(scene, cam, film, ps, count, seed) (render.rs:67-115): adds
 * `count` samples per pixel to `film`, continuing sample indices from
 * film->samples.  Synchronous. */
int ag_render_pass(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, ag_film* film,
                   const ag_path_settings* ps, uint32_t count, uint32_t seed);
/* The same pass, enqueued: returns once the work is on the context's stream.
 * The film's sample count advances at once, so further passes (async or
 * not) on the context run after it and continue its sample indices; the
 * film, scene and camera must stay alive until the event is done.  No per-
 * kernel timing is recorded; ag_event_wait fills the context statistics.
 * The reference has no asynchronous pass; this is for callers that overlap
 * host work or other contexts with rendering. */
typedef struct ag_event ag_event;
int ag_render_pass_async(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, ag_film* film,
                         const ag_path_settings* ps, uint32_t count, uint32_t seed, ag_event** ev);
int ag_event_query(const ag_event* ev);  /* AG_OK done, AG_PENDING not yet */
int ag_event_wait(ag_event* ev);
void ag_event_destroy(ag_event* ev);

/* Geometry queries over the scene's piece BVH in object space:
 * Bvh::intersect (bvh.rs:159-213) and Bvh::occluded (bvh.rs:216-245).
 * Host arrays; a ray's t_max is rays->tmax[i] (NULL = +inf). */
int ag_intersect(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, ag_hits_soa* out);
int ag_occluded(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, uint8_t* out);
/* The same on device pointers already resident in HBM; asynchronous on the
 * context stream.  Returns after the launch. */
int ag_intersect_device(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays_dev, uint32_t n,
                        ag_hits_soa* out_dev);
int ag_occluded_device(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays_dev, uint32_t n,
                       uint8_t* out_dev);
/* ag_intersect_device without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * traversal's scratch -- ray cursors and per-thread stack spill areas -- lives
 * in one static device array of 96 MiB + 4 KiB per device, sized for a grid
 * of at most 512 traversal blocks; any n fits, since the kernel pulls rays.
 * Results equal ag_intersect_device's bit for bit.  Enqueues on the context
 * stream and returns without waiting (ag_ctx_synchronize waits).  Every call
 * on a device shares that device's array: calls that overlap on one device
 * (other contexts, other threads) race and must be serialised by the caller. */
int ag_intersect_device_noalloc(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays_dev, uint32_t n,
                                ag_hits_soa* out_dev);

/* Cavity ambient occlusion of the scene's piece: ambient_occlusion
 * (argent-mesh cavity.rs:95-128) with the piece as its Occluder (argent-core
 * scene.rs:40-44, i.e. Bvh::occluded).  For each vertex, the fraction of
 * `samples` (at least 1) cosine-weighted rays about its normal, leaving
 * p + n * max(2e-4 * |diagonal of the vertex bounds|, 1e-3), that are blocked
 * within `distance` (object space).  The scene's vertex normals, or without
 * them the angle-weighted smooth normals (normals.rs:226-242) as
 * cavity.rs:82-87 does; a zero normal gives 0.  out: n_vertices floats (host).  This is the ray-traced step
 * of the reference's patina mask (cavity.rs:52-79); ag_patina_mask runs the
 * whole mask. */
int ag_ambient_occlusion(ag_ctx* ctx, const ag_scene* s, uint32_t samples, float distance, float* out);

/* The patina mask: PatinaSettings (argent-mesh cavity.rs:24-42). */
typedef struct ag_patina_settings {
    uint32_t ao_samples;        /* AO rays per vertex (48); 0 = no AO term */
    float ao_distance;          /* how far a recess sees, object units (1.2); <= 0 = no AO term */
    float curvature_radius;     /* concavity neighbourhood radius (0.4) */
    float concavity_weight;     /* blend of AO (0) and concavity (1), clamped to [0, 1] (0.35) */
    uint32_t smooth_iterations; /* Laplacian passes over the mask (3) */
} ag_patina_settings;
/* PatinaSettings::default() (cavity.rs:38-42). */
int ag_patina_default(ag_patina_settings* out);
/* The settings Prepared::build sizes by the piece (shot.rs:133-141):
 * radius = max(|half the vertex-bounds diagonal|, 1e-3), ao_distance =
 * clamp(0.09 radius, 0.4, 2.5), curvature_radius = clamp(0.03 radius, 0.15,
 * 1), the rest default. */
int ag_patina_for_scene(const ag_scene* s, ag_patina_settings* out);
/* patina_mask (cavity.rs:52-80) of the scene's piece with the piece itself as
 * the Occluder (scene.rs:40-44), all on the device: one value in [0, 1] per
 * vertex, 1 = deep recess.  The AO term is ag_ambient_occlusion's; the
 * concavity (cavity.rs:130-243) is radius x mean curvature from a Gaussian
 * neighbourhood over grid cells; the blend is smoothstep-shaped; the
 * Laplacian smoothing (cavity.rs:245-296) treats copies of a vertex at one
 * position (crease splits) as one.  Without the scene's vertex normals it
 * uses the angle-weighted smooth normals, as ag_ambient_occlusion does.
 * store != 0: the mask becomes the piece's cavity (Piece::cavity, shot.rs:142)
 * that shading reads (the patina layer's weight, material.rs:113-121), for
 * this scene and for scenes created over the same piece afterwards
 * (ag_scene_create_shared); scenes created over it earlier keep the cavity
 * they had.  mask / ao / concavity: host arrays of n_vertices floats (the
 * mask, the AO fraction, the clamped concavity), each may be NULL. */
int ag_patina_mask(ag_ctx* ctx, ag_scene* s, const ag_patina_settings* settings, int32_t store, float* mask,
                   float* ao, float* concavity);
/* smooth (cavity.rs:245-296) of per-vertex values over the scene's piece, in
 * place: x <- x + lambda (mean of the neighbours - x), `iterations` times. */
int ag_smooth_values(ag_ctx* ctx, const ag_scene* s, float* values, uint32_t iterations, float lambda);

/* Preview development (argent-post), on the device.
 * Tone operators and looks of tonemap.rs:41-67 (Grade): */
enum {
    AG_TONEMAP_AGX = 0,        /* ToneMap::AgX(AgxLook::Base) */
    AG_TONEMAP_AGX_PUNCHY = 1, /* AgX Punchy, the reference's default grade */
    AG_TONEMAP_AGX_GOLDEN = 2,
    AG_TONEMAP_ACES = 3,       /* AcesFitted */
    AG_TONEMAP_KHRONOS = 4,    /* KhronosNeutral */
    AG_TONEMAP_REINHARD = 5,
    AG_TONEMAP_CLAMP = 6
};
typedef struct ag_grade {  /* tonemap.rs:53-58 */
    float exposure_ev;
    int32_t tonemap;       /* AG_TONEMAP_* */
    float saturation;      /* 1 = unchanged */
} ag_grade;
/* develop (tonemap.rs:370-388): exposure x 2^ev, tone operator, saturation,
 * sRGB encode (through the reference's per-grade tables), TPDF dither, 8 bit.
 * rgb: w*h*3 floats, scene-linear (host); out: w*h*3 bytes (host), row-major. */
int ag_develop_image(ag_ctx* ctx, const float* rgb, uint32_t w, uint32_t h, const ag_grade* grade, uint8_t* out);
/* The built-in à-trous denoiser (denoise.rs:660-709, the DenoiseBackend::Atrous
 * path of denoise_with): 5 edge-avoiding passes, albedo demodulation and
 * albedo/normal edge stops when those are given (NULL: not used).  Host
 * arrays of w*h*3 floats. */
int ag_denoise_image(ag_ctx* ctx, const float* color, const float* albedo, const float* normal, uint32_t w,
                     uint32_t h, float* out);
/* The same from a film on the device (Film::color/albedo/normal,
 * render.rs:35-64), without a host round trip of the linear images: the live
 * preview's development (live.rs:429-496: à-trous with both guides when
 * `denoise` is nonzero, then develop).  out: w*h*3 floats / bytes (host). */
int ag_film_denoise(ag_ctx* ctx, const ag_film* film, float* out);
int ag_film_develop(ag_ctx* ctx, const ag_film* film, const ag_grade* grade, int32_t denoise, uint8_t* out);

/* ------------------------------------------------------------ mesh ingest
 * A device-resident triangle mesh: positions (n_vertices x 3 floats) and
 * triangles (n_triangles x 3 indices), argent-mesh lib.rs:33-39 (Mesh) without
 * parts.  Welding and compaction run on the device and give the reference's
 * output bit for bit. */

/* LoadOptions (argent-mesh lib.rs:53-74); defaults weld 1, weld_tolerance
 * 1e-6, crease_angle_deg 50, unit_scale 1. */
typedef struct ag_load_options {
    int32_t weld;
    float weld_tolerance;    /* fraction of the bounding-box diagonal */
    float crease_angle_deg;  /* reserved for crease normals */
    float unit_scale;
} ag_load_options;

/* WeldReport (weld.rs:21-33). */
typedef struct ag_weld_report {
    uint64_t input_vertices, output_vertices, exact_merges, tolerance_merges, dropped_triangles;
    float tolerance_mm;
} ag_weld_report;

/* LoadReport (io.rs:21-40) for a binary STL: record counts, the truncation
 * and padding io/stl.rs:38-50 warns about, the invalid triangles io.rs:206-224
 * drops, and device time of each stage (upload_ms is host wall time of the
 * copy when the bytes come from host memory). */
typedef struct ag_load_report {
    uint64_t file_bytes, raw_triangles, raw_vertices, invalid_triangles, trailing_bytes, declared_triangles;
    int32_t welded;
    ag_weld_report weld;
    float upload_ms, parse_ms, weld_ms;
} ag_load_report;

/* load_bytes (io.rs:87-126) for a binary STL (io/stl.rs:34-62): parse, drop
 * triangles with non-finite corners, scale by unit_scale, then weld
 * (weld.rs:38-62) or, with weld == 0, compact (weld.rs:66-91).  bytes is host
 * memory, or device memory (4-byte aligned) when bytes_on_device is nonzero.
 * Format detection (io.rs:149-180) is the caller's: ASCII STL, OBJ, PLY and
 * 3MF are not read here.  Errors: AG_E_INVALID for a file shorter than its
 * header, with no triangles, or with only degenerate ones (io.rs:101-103,
 * :121-123). */
int ag_mesh_load_stl(ag_ctx* ctx, const void* bytes, size_t len, int32_t bytes_on_device,
                     const ag_load_options* opts, ag_mesh** out, ag_load_report* report);
/* A device mesh from host arrays. */
int ag_mesh_create(ag_ctx* ctx, const float* positions, uint32_t n_vertices, const uint32_t* triangles,
                   uint32_t n_triangles, ag_mesh** out);
void ag_mesh_destroy(ag_mesh* m);
int ag_mesh_info(const ag_mesh* m, uint32_t* n_vertices, uint32_t* n_triangles);
/* Copies the mesh to host arrays (either may be NULL). */
int ag_mesh_read(const ag_mesh* m, float* positions, uint32_t* triangles);
/* What framing a shot takes from a mesh, without copying the mesh back: the
 * vertex bounds (lo, hi: 3 floats each; zeros for an empty mesh) and every
 * stride-th vertex position (vertices 0, stride, 2 stride, ... below the
 * vertex count, at most cap of them, into hull_xyz; *n_hull = how many).
 * Replaces Prepared's host pass over the positions (shot.rs:73-147: radius
 * and floor height from the bounds, the framing hull from a subsample). */
int ag_mesh_frame_sample(const ag_mesh* m, uint32_t stride, float* hull_xyz, uint32_t cap, uint32_t* n_hull,
                         float* lo, float* hi);
/* The mesh's device arrays, valid until the next call that changes the mesh. */
int ag_mesh_device_arrays(const ag_mesh* m, const float** positions, const uint32_t** triangles);
/* weld (weld.rs:38-62): merges bit-identical vertices, then vertices within
 * rel_tolerance x the bbox diagonal; drops collapsed triangles and
 * unreferenced vertices.  A triangle index out of range is AG_E_INVALID. */
int ag_mesh_weld(ag_mesh* m, float rel_tolerance, ag_weld_report* report);
/* compact (weld.rs:66-91): drops collapsed triangles and unreferenced vertices. */
int ag_mesh_compact(ag_mesh* m, uint64_t* dropped_triangles);

/* NormalReport (normals.rs:23-42). */
typedef struct ag_normal_report {
    uint64_t shells, flipped_shells, cavity_shells, repaired_faces, boundary_edges, nonmanifold_edges,
        inconsistent_edges, degenerate_faces, split_vertices;
} ag_normal_report;

/* normals::compute (normals.rs:47-49): drops collapsed triangles, repairs
 * inconsistent winding (majority per edge-connected component), flips closed
 * inside-out shells that no other shell encloses, splits vertices along
 * creases sharper than crease_angle_deg (0: faceted) and gives every vertex
 * the angle-weighted normal of its fan.  Split copies are appended after the
 * original vertices; triangle order is kept.  The mesh gains normals
 * (ag_mesh_read_normals); a later weld or compact drops them. */
int ag_mesh_normals(ag_mesh* m, float crease_angle_deg, ag_normal_report* report);
int ag_mesh_read_normals(const ag_mesh* m, float* normals);

/* ------------------------------------------------------------ live view
 * The live view's render worker (argent serve live.rs:637-830) on the device:
 * after a change, a half-resolution first look at 2 spp; then refinement
 * passes sized toward 70 ms (1 spp while the set is "in motion", younger than
 * 350 ms); frames published at most every 83 ms in motion, else every
 * 166 ms x sqrt(spp / 32) up to 1.5 s; each frame denoised (à-trous, the
 * film's albedo and normal guides) and developed to dithered 8-bit on the
 * device; refinement stops at max_spp, or once a denoised frame differs from
 * the previous full-resolution one by less than 0.5 % (mean |dY| / mean Y
 * over every 4th pixel) from 64 spp on.  The caller's thread drives the loop
 * (one ag_live_step per worker iteration) and passes its clock in
 * milliseconds; the session adds the measured time of its own passes. */
typedef struct ag_live ag_live;

/* The Params fields the worker reads (params.rs): preview size (540 x 960
 * for the default portrait format), max_spp (default 256), denoise, the grade
 * of the look, and the path settings (PathSettings::default(): 32, 24, 6). */
typedef struct ag_live_params {
    uint32_t width, height;
    uint32_t max_spp;
    int32_t denoise;
    ag_grade grade;
    ag_path_settings ps;
} ag_live_params;

/* What one step did. */
typedef struct ag_live_frame {
    int32_t published;    /* a frame was developed into out */
    int32_t idle;         /* nothing to do until the next change */
    uint64_t generation;  /* of the frame (published) or of the session */
    uint32_t width, height, spp, max_spp;
    int32_t half;         /* the half-resolution first look */
    int32_t converged;    /* final: max_spp reached, or settled */
    int32_t settled;      /* final before max_spp by the settle test */
    int32_t denoised;
    int32_t redeveloped;  /* a look-only change re-developed the last frame */
    uint32_t pass_spp;    /* samples this step rendered (0: none) */
    float pass_ms, ms_per_spp;
    float relative_change; /* the settle test's value, -1 when not evaluated */
} ag_live_frame;

int ag_live_create(ag_ctx* ctx, ag_live** out);
void ag_live_destroy(ag_live* live);
/* Attach the session to a continuous-batching server (NULL detaches): its
 * passes then run inside that server's steps, in the same wavefront schedule
 * as the queued requests' chunks, which take what the pass leaves of the step
 * budget.  The reference instead pauses the live view while a queued job
 * renders (live.rs:344-352, queue.rs:511-515).  The session's frames are the
 * same either way; its pacing is charged the pass's share of the step's time
 * by samples.  The caller drives both from one thread. */
int ag_live_attach(ag_live* live, struct ag_server* server);
/* A change of scene, camera or params: a new generation (accumulation restarts).
 * The session keeps a pointer to `scene`, which must outlive its use (until
 * the next ag_live_set or ag_live_destroy). */
int ag_live_set(ag_live* live, const ag_scene* scene, const ag_camera* cam, const ag_live_params* params,
                double now_ms);
/* A look-only change (grade, denoise): the last frame is re-developed at the
 * next step.  now_ms is accepted for symmetry with ag_live_set; the look
 * change does not restart the motion window. */
int ag_live_set_look(ag_live* live, const ag_grade* grade, int32_t denoise, double now_ms);
/* One worker iteration.  out: width * height * 3 bytes of the params' size. */
int ag_live_step(ag_live* live, double now_ms, ag_live_frame* frame, uint8_t* out);
/* relative_change (live.rs:208-218) of two linear images (host, w*h*3 floats),
 * computed on the device: sum |lum a - lum b| / sum |lum b| over every 4th
 * pixel in f64, 0 when the base is below 1e-9. */
int ag_relative_change(ag_ctx* ctx, const float* a, const float* b, uint32_t w, uint32_t h, double* out);

/* Material::bsdf + Bsdf::{eval,pdf,sample} (material.rs:96-138, bsdf.rs:53-239)
 * evaluated on the device for n given inputs, for checking the shading code:
 * material index, 4 wavelengths, cavity mask, u = [layer, u0, u1, u2], wo and
 * wi in the local frame.  out: n x 18 floats = eval(4), pdf, sampled(0/1),
 * wi(3), f(4), pdf, delta, transmit, terminate_secondary, is_delta. */
int ag_bsdf_probe(ag_ctx* ctx, const ag_scene* s, uint32_t n, const int32_t* mat, const float* lam4,
                  const float* mask, const float* u4, const float* wo3, const float* wi3, float* out18);

/* Device buffers for callers that keep ray batches resident in HBM (used with
 * the *_device queries).  Copies are synchronous on the context stream. */
int ag_device_alloc(ag_ctx* ctx, size_t bytes, void** out);
int ag_device_free(ag_ctx* ctx, void* p);
int ag_memcpy_to_device(ag_ctx* ctx, void* dst, const void* src, size_t bytes);
int ag_memcpy_to_host(ag_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Elapsed device time of the last *_device query on the context stream (HIP
 * events around the kernel), milliseconds. */
double ag_last_query_ms(const ag_ctx* ctx);

/* ------------------------------------------------------------ batching
 * Many uneven views of one scene in one wavefront schedule: the turntable
 * loop of jobs.rs:689-710 and metal variants, with each view's own camera,
 * turntable angle, metal slot-0 material, film and sample count.  Each
 * view's film equals what ag_render_pass would give for it alone. */
typedef struct ag_view {
    ag_camera camera;
    float turntable_deg;
    int32_t material0;      /* index into the materials array, -1 = the scene's slot 0 */
    ag_film* film;
    uint32_t count;
    uint32_t seed;
    const ag_scene* scene;  /* NULL = the call's scene; views of different scenes share waves */
} ag_view;
int ag_render_views(ag_ctx* ctx, const ag_scene* scene, const ag_material* materials, int32_t n_materials,
                    const ag_view* views, int32_t n_views, const ag_path_settings* ps);

/* ------------------------------------------------------------ serving
 * A continuous-batching render server: the role of the studio's live worker
 * and job queue (live.rs:637-823, queue.rs:475-691) for many clients at once.
 * Requests are admitted FIFO while their films fit the HBM budget.  Every
 * ag_server_step renders, in one shared wavefront schedule (requests on
 * different scenes share waves), the next sample chunk of as many active
 * requests as fit in step_samples samples, those with the least work left
 * first (at least one request per step).  A step in which any request still
 * waits for its first preview renders only such requests, so a request
 * arriving while others run gets its preview after the running step plus one
 * short preview step.  A request's chunk sizes depend only on its own spp and the
 * server's settings (first chunk preview_spp, then doubling up to a cap of
 * step_samples / pixels), so its film is bit-identical whether it is served
 * alone or in a mix.  Films stay resident until released; ag_server_read
 * gives the progressive image.
 *
 * Admission charges the HBM budget with the request's film, the device bytes
 * of its scene look and piece (each charged once while any admitted request
 * uses it; ag_scene_stats gives them) and the context's wave state; a request
 * that would not fit an idle server is REJECTED at submit.  Preview-only
 * steps run at most 4 in a row; then the requests in progress go first.
 * A request whose film cannot be created, or whose chunk fails to render when
 * retried alone, becomes FAILED (ag_server_error tells why) and frees its
 * film; the other requests carry on, as the reference's job queue isolates
 * each job (queue.rs:514-557). */
enum { AG_REQ_QUEUED = 0, AG_REQ_ACTIVE = 1, AG_REQ_DONE = 2, AG_REQ_REJECTED = 3, AG_REQ_FAILED = 4 };
typedef struct ag_request {
    const ag_scene* scene;
    ag_camera camera;
    float turntable_deg;
    int32_t has_material0;   /* 1: material0 replaces the scene's slot 0 (a metal variant) */
    ag_material material0;
    uint32_t width, height, spp, seed;
    uint32_t preview_spp;    /* samples in the first chunk; 0 = 2 (the live view's first look, live.rs:745) */
} ag_request;
typedef struct ag_request_status {
    int32_t state;
    uint32_t samples;
    /* milliseconds since the server was created; -1 until reached */
    double t_submit, t_admit, t_first_preview, t_done;
} ag_request_status;
typedef struct ag_server ag_server;
int ag_server_create(ag_ctx* ctx, size_t hbm_budget_bytes, uint64_t step_samples, const ag_path_settings* ps,
                     ag_server** out);
void ag_server_destroy(ag_server* s);
int ag_server_submit(ag_server* s, const ag_request* req, uint64_t* id);
/* One scheduling round: admit, render one chunk of each request chosen as
 * above, retire finished ones.  Returns the counts left active and queued. */
int ag_server_step(ag_server* s, uint32_t* n_active, uint32_t* n_queued);
int ag_server_status(const ag_server* s, uint64_t id, ag_request_status* out);
/* Progressive image of a request (as ag_film_read), host memory w*h*3. */
int ag_server_read(const ag_server* s, uint64_t id, int32_t which, float* out);
int ag_server_release(ag_server* s, uint64_t id);
/* Why a request failed ("" otherwise); valid until the request is released. */
const char* ag_server_error(const ag_server* s, uint64_t id);
/* Bytes charged against the budget now (films, scenes, wave state), and the budget. */
int ag_server_budget(const ag_server* s, size_t* used, size_t* budget);
/* Failure injection for tests of failure handling (no counterpart in the
 * reference): from now on,
 * every render pass on ctx that includes a view rendered with this seed fails
 * with AG_E_DEVICE before launching anything; -1 turns it off.  Lets a test
 * make one request's chunks fail from the host, with every kernel still
 * exiting normally. */
int ag_debug_fail_seed(ag_ctx* ctx, int64_t seed);

/* ------------------------------------------------------------ tuning
 * Launch shape of the wavefront kernels (the autotuner's knobs). */
typedef struct ag_tuning {
    uint32_t paths_per_wave;   /* path slots in flight (1024..AG_MAX_PATHS_PER_WAVE,
                                  about 280 bytes each) */
    uint32_t block_traverse;   /* threads per block, traversal kernels */
    uint32_t blocks_per_cu;    /* persistent traversal blocks per CU, render passes */
    uint32_t sort_by_material; /* sort granularity of shading: 1 = material-sorted
                                  order; 0 = arrival order; 2 = arrival order in
                                  runs of one traversal wavefront's hits, each
                                  run ordered by material; 3 = misses and panel
                                  hits first, then every surface hit in arrival
                                  order; 4 = by shading family (misses and
                                  panel hits, metals, diffuse surfaces, the
                                  rest), each family shaded by its own kernel */
    uint32_t refill_min;       /* render passes, extension rays: idle lanes (1..64)
                                  before a wavefront takes new rays; 64 = whole
                                  coherent batches */
    uint32_t batch_blocks_per_cu; /* the same two for ag_intersect / ag_occluded */
    uint32_t batch_refill_min;
    uint32_t shadow_refill_min;   /* render passes, shadow rays (1..64) */
    uint32_t spread_min_lanes;    /* every traversal (render passes, ray batches,
                                     cavity AO): a ray queue too short to give
                                     every wavefront of the grid 64 rays is
                                     spread over all of them, ceil(n / wavefronts)
                                     lanes each but at least this many (1..64;
                                     64 = never spread; default 1) */
    uint32_t ordered_emission;    /* 1: shading writes each hit's continuation ray
                                     to the hit's own slot of the next queue
                                     (dead slots where paths end), so the next
                                     queue keeps the hits' order exactly; 0:
                                     appended as shading wavefronts finish */
    uint32_t cursor_slices;       /* render traversal: a queue is cut into this
                                     many slices (1..8), each pulled through its
                                     own cursor, block b starting on slice
                                     b mod slices (8 = one per XCD) */
    uint32_t launch_graph;        /* 1: a render pass that fits one wave
                                     replays a captured HIP graph of its
                                     launches (captured again when the pass's
                                     shape -- films, depth, wave, buffers --
                                     changes; a few shapes are kept); 0: every
                                     kernel is launched on its own */
    uint32_t shade_waves;         /* shading kernel compiled for this many waves
                                     per SIMD (2: every value in registers; 3:
                                     168 registers, a few values spilled) */
    uint32_t raygen_bands;        /* camera rays queued in this many horizontal
                                     bands of the image, band-major (1..8; with
                                     cursor_slices 8, one band per XCD) */
} ag_tuning;
int ag_ctx_get_tuning(const ag_ctx* ctx, ag_tuning* out);
int ag_ctx_set_tuning(ag_ctx* ctx, const ag_tuning* t);
/* Render passes captured into a HIP graph and passes replayed from one since
 * the context was created (launch_graph); no counterpart in the reference. */
int ag_ctx_graph_stats(const ag_ctx* ctx, uint64_t* captures, uint64_t* replays);

/* Autotuner (no counterpart in the reference, whose CPU renderer has no launch
 * shape): times the caller's workload under candidate tunings and leaves the
 * fastest set on the context.  Coordinate descent from the current tuning,
 * each candidate the median of `repeats` timed runs after one untimed run; a
 * change is kept when at least 1 % faster.  Every candidate renders the same
 * film bit for bit.  ag_autotune_render times render passes of width x height
 * at spp samples (paths_per_wave, blocks_per_cu, refill_min,
 * shadow_refill_min, spread_min_lanes, sort_by_material); ag_autotune_rays
 * times ag_intersect_device on a resident batch (batch_blocks_per_cu,
 * batch_refill_min, spread_min_lanes). */
typedef struct ag_autotune_result {
    ag_tuning best;     /* also set on the context */
    double ms_start;    /* the starting tuning's median time */
    double ms_best;     /* the best tuning's median time */
    uint32_t trials;    /* tunings timed */
} ag_autotune_result;
int ag_autotune_render(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, uint32_t width, uint32_t height,
                       uint32_t spp, const ag_path_settings* ps, uint32_t repeats, ag_autotune_result* out);
int ag_autotune_rays(ag_ctx* ctx, const ag_scene* scene, const ag_rays_soa* rays_dev, uint32_t n,
                     ag_hits_soa* hits_dev, uint32_t repeats, ag_autotune_result* out);

/* Kernel timings of the last ag_render_pass / ag_render_views (HIP events on
 * the context stream), milliseconds, and ray counts. */
typedef struct ag_stats {
    double ms_total, ms_extend, ms_shade, ms_shadow, ms_raygen, ms_accum;
    uint64_t extension_rays, shadow_rays, samples;
    uint32_t extend_launches, shadow_launches;
} ag_stats;
int ag_ctx_stats(const ag_ctx* ctx, ag_stats* out);
/* The last render's statistics per bounce depth d < n (n <= 64): extension
 * and shadow rays traced, and (with timing enabled) milliseconds in the
 * extend, shade and shadow kernels of that depth; any pointer may be null. */
int ag_ctx_depth_stats(const ag_ctx* ctx, int32_t n, uint64_t* extension_rays, uint64_t* shadow_rays,
                       double* ms_extend, double* ms_shade, double* ms_shadow);
/* Per-stage event timing for synchronous render passes (off by default).
 * With it on, every stage keeps a launch of its own so that it can be timed;
 * with it off, the shadow rays of each bounce and the extension rays of the
 * next run as one launch.  Films are the same either way. */
int ag_ctx_enable_timing(ag_ctx* ctx, int32_t on);

#ifdef __cplusplus
}
#endif
#endif /* ARGENT_H */
