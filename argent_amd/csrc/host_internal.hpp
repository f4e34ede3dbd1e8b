// Internal declarations shared by the host translation units.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/argent.h"

struct ihipStream_t;  // hipStream_t is ihipStream_t*

namespace ag {

int fail(int code, const std::string& msg);

// The context's device and HIP stream (runtime.cpp).
::ihipStream_t* ctx_stream(const ag_ctx* c);
int ctx_device(const ag_ctx* c);
// A device workspace of at least `bytes`, owned by the context and reused by
// later calls (grow-only; its contents do not survive a call that grows it).
int ctx_workspace(ag_ctx* c, size_t bytes, void** p);

struct DevGrade;  // post.hpp
// The tables of a tone-mapping grade (tonemap.rs:284-313), host_post.cpp:
// blob = curves (n_curves x curve_len floats) then the encode table
// (encode_len floats); `g` gets everything but the device pointers.
int build_grade(float exposure_ev, int32_t op, float saturation, std::vector<float>* blob, DevGrade* g, int* n_curves,
                size_t* curve_len, size_t* encode_len);
void light_cdf(const ag_scene_desc* d, int* n_lights, int* lights, float* cdf, float* env_pick);

// Compressed 8-wide BVH in the device layout.  Node: 5 x 16 B = 80 B
//   [0] origin p.xyz (float), then one word: exponent bytes ex | ey << 8 |
//       ez << 16 (axis scale 2^(e-127), e in 1..254) | imask << 24 (bit k:
//       child k is an inner node)
//   [1] child_base (inner children k-th inner at child_base + k), tri_base,
//       meta bytes of children 0-3, of children 4-7: inner -> its rank among
//       the inner children; leaf -> count << 5 | offset of its first triangle
//       from tri_base (count 1..4, offset 0..31); 0 and not inner -> empty
//   [2] qlo_x[0..7], qlo_y[0..7]   [3] qlo_z[0..7], qhi_x[0..7]
//   [4] qhi_y[0..7], qhi_z[0..7]   (bytes; box = p + q * scale, float fma)
// The quantized boxes contain the exact child boxes as the kernel evaluates
// them (checked on the host in the same float arithmetic).
// Stack entries: inner node index, or LEAF | (count-1) << 27 | first triangle.
// Triangle: 3 x 16 B = 48 B
//   (v0.x, v0.y, v0.z, e1.x) (e1.y, e1.z, e2.x, e2.y) (e2.z, original index bits, 0, 0)
constexpr uint32_t BVH_LEAF = 0x80000000u;
constexpr uint32_t BVH_EMPTY = 0xFFFFFFFFu;
constexpr uint32_t BVH_COUNT_SHIFT = 27;
constexpr uint32_t BVH_INDEX_MASK = (1u << 27) - 1u;
constexpr int BVH_MAX_LEAF = 4;
constexpr int BVH_WIDTH = 8;
#ifndef AG_NODE_FLOATS
#define AG_NODE_FLOATS 20  // 80-byte nodes; the traversal kernels read the first 20 floats
#endif
constexpr int BVH_NODE_FLOATS = AG_NODE_FLOATS;

struct WideBvh {
    std::vector<float> nodes;  // BVH_NODE_FLOATS per node
    std::vector<float> tris;   // 12 floats per triangle
    uint32_t n_nodes = 0, n_leaves = 0, depth = 0;
    uint32_t max_stack = 0;    // most stack entries any traversal can hold
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};

// Quantize child box [lo, hi] against node origin p and scale s (exact power
// of two) so that fmaf(qlo, s, p) <= lo and fmaf(qhi, s, p) >= hi.
void quantize_axis(float p, float s, float lo, float hi, uint8_t* qlo, uint8_t* qhi);

void build_wide_bvh(const float* positions, const uint32_t* triangles, uint32_t n_tris, WideBvh* out);

// The same tree format built on the device (bvh.hip) from device positions
// and triangle indices: nodes needs room for n_tris + 1 nodes, tris for
// n_tris triangles; temporaries live in the context workspace.
struct GpuBvhInfo {
    uint32_t n_nodes = 0, n_leaves = 0, depth = 0, max_stack = 0;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};
int gpu_build_wide_bvh(ag_ctx* ctx, const float* positions, const uint32_t* triangles, uint32_t n_tris, float* nodes,
                       float* tris, GpuBvhInfo* info);

}  // namespace ag
