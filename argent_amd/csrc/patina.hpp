// The patina mask's device pipeline (cavity.hip), called by the runtime.
#pragma once
#include <cstdint>

#include "../../include/argent.h"

namespace ag {

struct PatinaInputs {
    const float* pos;             // device, 3 * nv
    const float* nrm;             // device, 3 * nv (needed for the concavity)
    const uint32_t* tris;         // device, 3 * nt
    uint32_t nv, nt;
    float lo[3];                  // Mesh::bounds minimum (the cell grid origin)
    const uint32_t* ao_counts;    // device, blocked AO rays per vertex, or null (no AO term)
    uint32_t ao_samples;          // rays per vertex behind ao_counts
    const float* values;          // device: smooth these instead of building the mask (or null)
    float lambda;                 // smoothing step (0.5 for the mask)
    int clamp01;                  // clamp the result to [0, 1] (the mask does)
};

// normals::smooth_vertex_normals (normals.rs:226-242) into `out` (device,
// 3 * nv floats), the normals cavity.rs:82-87 falls back to.  Synchronizes.
int smooth_normals_device(ag_ctx* ctx, const float* pos, const uint32_t* tris, uint32_t nv, uint32_t nt,
                          float* out);

// Writes nv floats to `mask` (device); ao_out / cav_out (device, or null)
// get the AO fraction and the concavity per vertex.  Synchronizes.
int patina_device(ag_ctx* ctx, const PatinaInputs& in, const ag_patina_settings& s, float* mask, float* ao_out,
                  float* cav_out);

}  // namespace ag
