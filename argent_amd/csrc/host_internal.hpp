// Internal declarations shared by the host translation units.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/argent.h"

namespace ag {

int fail(int code, const std::string& msg);
void light_cdf(const ag_scene_desc* d, int* n_lights, int* lights, float* cdf, float* env_pick);

// Wide BVH in the device layout.  Node: 8 x 16 B = 128 B
//   [0..5] lo_x, lo_y, lo_z, hi_x, hi_y, hi_z of the four children (float4 each)
//   [6]    child entries (uint4): inner node index, or LEAF | (count-1) << 27 | first triangle,
//          or EMPTY
//   [7]    reserved
// Triangle: 3 x 16 B = 48 B
//   (v0.x, v0.y, v0.z, e1.x) (e1.y, e1.z, e2.x, e2.y) (e2.z, original index bits, 0, 0)
constexpr uint32_t BVH_LEAF = 0x80000000u;
constexpr uint32_t BVH_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t BVH_COUNT_SHIFT = 27;
constexpr uint32_t BVH_INDEX_MASK = (1u << 27) - 1u;
constexpr int BVH_MAX_LEAF = 4;

struct WideBvh {
    std::vector<float> nodes;  // 32 floats per node
    std::vector<float> tris;   // 12 floats per triangle
    uint32_t n_nodes = 0, n_leaves = 0, depth = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

void build_wide_bvh(const float* positions, const uint32_t* triangles, uint32_t n_tris, WideBvh* out);

}  // namespace ag
