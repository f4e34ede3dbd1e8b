/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.  A CPU restatement of the reference's
 * render core (argent-core render.rs, integrator.rs, scene.rs, light.rs,
 * envmap.rs, bvh.rs, bsdf.rs, microfacet.rs, material.rs, camera.rs,
 * sampler.rs, color.rs and argent-spectra spec.rs / fresnel.rs / cie.rs).
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg load
 * it, as the checker of the HIP path and as the timed CPU baseline; the
 * product library never links it.
 *
 * It consumes the same ag_scene_desc as ag_scene_create.  The stratified
 * sample dimensions use the same counter-based stream as the GPU path
 * (hash of pixel seed, sample index, dimension) in place of the reference's
 * Owen-scrambled Sobol (sobol_burley 0.5.0, not available here).
 */
#ifndef ARGENT_ORACLE_H
#define ARGENT_ORACLE_H
#include <stdint.h>

#include "../include/argent.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* traversal modes of or_intersect */
enum {
    OR_MODE_REFERENCE = 0, /* bvh.rs:159-213 exactly: SSE min/max slab, first hit wins ties */
    OR_MODE_ROBUST = 1     /* same tree, boxes widened by 2^-21, ties to the lower triangle id */
};

or_scene* or_scene_create(const ag_scene_desc* d);
void or_scene_destroy(or_scene* s);
void or_scene_set_turntable(or_scene* s, float deg);
int or_scene_set_material(or_scene* s, int32_t m, const ag_material* mat);
int or_scene_stats(const or_scene* s, uint32_t* binary_nodes, uint32_t* wide_nodes);
/* Estimator strategy (checking knob, not in the reference): 0 = NEE + BSDF
 * sampling combined by the power heuristic as integrator.rs does; 1 = NEE
 * alone; 2 = BSDF sampling alone; 3 = both unweighted (a planted double
 * count). */
int or_scene_set_strategy(or_scene* s, int32_t strategy);

int or_intersect(const or_scene* s, const float* ox, const float* oy, const float* oz, const float* dx, const float* dy,
                 const float* dz, const float* tmax, uint32_t n, int32_t mode, int32_t threads, float* t, float* u,
                 float* v, uint32_t* prim);
// reference-tree work of closest-hit queries: totals = {Node4 visits, triangle tests}
int or_intersect_stats(const or_scene* s, const float* ox, const float* oy, const float* oz, const float* dx,
                       const float* dy, const float* dz, const float* tmax, uint32_t n, int32_t mode, uint64_t* totals);
// cavity AO per vertex (argent-mesh cavity.rs:95-128); without the scene's
// normals, smooth_vertex_normals' (cavity.rs:82-87)
int or_ambient_occlusion(const or_scene* s, uint32_t samples, float distance, int32_t mode, int32_t threads,
                         float* out);
int or_occluded(const or_scene* s, const float* ox, const float* oy, const float* oz, const float* dx, const float* dy,
                const float* dz, const float* tmax, uint32_t n, int32_t mode, int32_t threads, uint8_t* out);
/* Brute-force closest hit over every triangle (bvh.rs:557-570), ties to the lower id. */
int or_brute_intersect(const float* positions, const uint32_t* triangles, uint32_t n_tris, const float* ox,
                       const float* oy, const float* oz, const float* dx, const float* dy, const float* dz,
                       const float* tmax, uint32_t n, int32_t threads, float* t, float* u, float* v, uint32_t* prim);

/* render_pass (render.rs:67-115) over rows [row0, row1) of a w x h film.
 * sums: 9*w*h floats (colour, albedo, normal planes), updated in place;
 * first = samples already in the film.  stats[0] = extension segments,
 * stats[1] = shadow rays, stats[2] = samples. */
int or_render_pass(const or_scene* s, const ag_camera* cam, uint32_t w, uint32_t h, float* sums, uint32_t first,
                   const ag_path_settings* ps, uint32_t count, uint32_t seed, int32_t threads, uint32_t row0,
                   uint32_t row1, uint64_t* stats);

/* Known-answer helpers */
int or_energy_table(float* out1024);
float or_fresnel_dielectric(float cos_i, float eta);
float or_fresnel_conductor1(float cos_i, float eta, float k);
int or_sample_wavelengths(float u, float* lam4, float* pdf4);
int or_camera_matrix(int32_t has_white, const float* white, float* m9);
float or_rng_dim(uint32_t pix_seed, uint32_t si, uint32_t dim);
int or_camera_ray(const ag_camera* cam, float sx, float sy, float lu, float lv, float* o3, float* d3);
/* BSDF probe (material.rs:96-138 + bsdf.rs:53-239) for material m of the
 * scene at wavelengths lam4: out = [eval(4), pdf, sample ok, wi(3), f(4),
 * pdf, delta, transmit, term2] (18 floats).  u = [layer, u0, u1, u2]. */
int or_bsdf_probe(const or_scene* s, int32_t m, const float* lam4, float mask, const float* u4, const float* wo3,
                  const float* wi3, float* out18);

/* Environment map (envmap.rs): the LinearRgb bases (3 x 471) and inverse
 * response (3 x 3); EnvData::load's tables from linear RGB (w*h*3);
 * EnvMap::sample for n (u1, u2) pairs and EnvMap::pdf for n directions. */
int or_linear_rgb(float* basis3x471, float* inv9);
int or_env_build(const float* rgb, int32_t w, int32_t h, float exposure_ev, float* weights, float* marginal,
                 float* conditional, float* total);
int or_env_sample(const float* marginal, const float* conditional, int32_t w, int32_t h, float rotation_deg,
                  const float* u, uint32_t n, float* dirs, float* pdfs);
int or_env_pdf(const float* marginal, const float* conditional, int32_t w, int32_t h, float rotation_deg,
               const float* dirs, uint32_t n, float* pdfs);

// Preview development (post.cpp): argent-post tonemap.rs and denoise.rs.
// op: 0-2 AgX Base/Punchy/Golden, 3 ACES fitted, 4 Khronos Neutral, 5 Reinhard, 6 Clamp.
int or_tonemap_pixel(const float* rgb, uint32_t n, float ev, int32_t op, float sat, float* out);
int or_develop_encoded(const float* rgb, uint32_t n, float ev, int32_t op, float sat, float* out);
float or_srgb_encode(float x);
int or_develop(const float* rgb, uint32_t w, uint32_t h, float ev, int32_t op, float sat, uint8_t* out);
int or_atrous(uint32_t w, uint32_t h, const float* color, const float* albedo, const float* normal, float* out);
double or_relative_change(const float* a, const float* b, uint32_t n);

// Mesh ingest (mesh.cpp): argent-mesh io/stl.rs, io.rs, weld.rs.
// rep (weld): input vertices, output vertices, exact merges, tolerance merges,
// dropped triangles.  or_load_stl's rep: raw triangles, raw vertices, invalid
// triangles dropped, trailing bytes, declared count, then the weld's five.
int64_t or_stl_count(const uint8_t* b, size_t len);
int or_weld(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float rel_tolerance, float* out_pos,
            uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t* rep5, float* tol_mm);
int or_compact(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_pos, uint32_t* out_nv,
               uint32_t* out_tris, uint32_t* out_nt, uint64_t* dropped);
// Crease-aware normals (normals.cpp): normals::compute with one crease angle.
// rep9: shells, flipped shells, cavity shells, repaired faces, boundary,
// non-manifold and inconsistent edges, degenerate faces, split vertices.
int or_normals(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float crease_deg, float* out_pos,
               float* out_nrm, uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t* rep9);
/* Patina mask (cavity.cpp): argent-mesh cavity.rs.  or_concavity: signed
 * concavity per vertex (cavity.rs:130-243) from the given vertex normals;
 * or_smooth: Laplacian smoothing in place with crease copies as one
 * (cavity.rs:245-296); or_patina_finish: the blend of AO and concavity,
 * smoothing and clamp of patina_mask (cavity.rs:64-79). */
int or_concavity(const float* pos, const float* nrm, uint32_t nv, const uint32_t* tris, uint32_t nt, float radius,
                 float* out);
int or_smooth(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* values, uint32_t iterations,
              float lambda);
int or_patina_finish(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, const float* ao,
                     const float* cav, float concavity_weight, uint32_t smooth_iterations, float* out);
/* normals::smooth_vertex_normals (normals.rs:226-242), normals.cpp */
int or_smooth_vertex_normals(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out);
int or_load_stl(const uint8_t* b, size_t len, int32_t do_weld, float rel_tolerance, float unit_scale, float* out_pos,
                uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t* rep10, float* tol_mm);

#ifdef __cplusplus
}
#endif
#endif
