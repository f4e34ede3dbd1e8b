// Launch wrappers of the device kernels (defined in kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_types.hpp"

namespace ag {

struct RaysSoA {
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmax;
};
struct HitsSoA {
    float *t, *u, *v;
    uint32_t* prim;
};

struct ProbeIn {
    const int32_t* mat;
    const float* lam;   // 4 per probe
    const float* mask;
    const float* u;     // 4 per probe: layer, u0, u1, u2
    const float* wo;    // 3 per probe
    const float* wi;    // 3 per probe
};
void launch_bsdf_probe(hipStream_t s, const DScene* S, ProbeIn in, float* out, uint32_t n);
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n);
// `cursor` must hold 0 on entry (the persistent kernels pull rays from it).
// `refill`: idle lanes (1..64) before a wavefront takes new rays; `spread`:
// the least lanes per wavefront when a short queue is spread over the grid.
void launch_rays_intersect(hipStream_t s, int grid, const float4* nodes, const float4* tris, RaysSoA r, HitsSoA h,
                           uint32_t n, uint32_t* cursor, uint32_t refill, uint32_t spread, uint2* spill);
void launch_rays_occluded(hipStream_t s, int grid, const float4* nodes, const float4* tris, RaysSoA r, uint8_t* out,
                          uint32_t n, uint32_t* cursor, uint32_t refill, uint32_t spread, uint2* spill);
// cavity ambient occlusion: n = n_vertices * samples rays, hit counts added to counts[vertex]
void launch_ao(hipStream_t s, int grid, const float4* nodes, const float4* tris, const float* pos, const float* nrm,
               uint32_t samples, float dist, float eps, uint32_t* counts, uint32_t n, uint32_t* cursor,
               uint32_t refill, uint32_t spread, uint2* spill);
// Resident blocks per CU of the traversal kernels (occupancy limit).
int traversal_blocks_per_cu(bool batch);  // resident blocks per CU: ray-batch or render kernels
void launch_raygen(hipStream_t s, int grid, const DSeg* segs, int nseg, uint32_t npaths, const DView* views, DPaths P,
                   DQueues Q);
// spread: the least lanes per wavefront when a short queue is spread over the grid (64: never)
void launch_extend(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth, uint32_t refill,
                   uint32_t spread, uint2* spill);
// waves: the k_shade instance compiled for 2 or 3 waves per SIMD; tab_floats:
// the first view's scene table floats to stage in LDS (dynamic LDS, at most
// SHADE_TAB_FLOATS).  Q.sort 4 launches one kernel per shading family.
void launch_shade(hipStream_t s, int grid, const DView* views, const DMat* vmats, uint32_t max_depth,
                  const float* energy, DPaths P, DQueues Q, int depth, int waves, uint32_t tab_floats);
void launch_shadow(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth, uint32_t refill,
                   uint32_t spread, uint2* spill);
// k_shadow then k_extend of depth + 1 as one launch (same results)
void launch_shadow_extend(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth,
                          uint32_t refill_shadow, uint32_t refill_extend, uint32_t spread, uint2* spill);
void launch_accumulate(hipStream_t s, int grid, const DSeg* segs, int nseg, const DView* views, const float* cmf,
                       DPaths P);
void launch_tally(hipStream_t s, const uint32_t* counters, uint32_t depth_slots, unsigned long long* totals);
void launch_film_commit(hipStream_t s, float* film, float* acc, size_t n);

}  // namespace ag
