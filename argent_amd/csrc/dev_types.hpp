// Types shared by the HIP kernels and the host runtime that launches them.
// Everything here is plain data laid out for the device: scalars of a scene
// live in one DScene block in device memory (read with scalar loads), the
// spectral tables in one float blob, and path state in structure-of-arrays
// buffers sized for one wave of paths.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/argent.h"

namespace ag {

// Address spaces for device pointers whose home is known (LDS, global): reads
// through them are ds_read / global_load, never flat loads, which wait on the
// LDS and scalar counters as well.  Empty in the host pass.
#if defined(__HIP_DEVICE_COMPILE__)
#define AG_LDS __attribute__((address_space(3)))
#define AG_GLOBAL __attribute__((address_space(1)))
#else
#define AG_LDS
#define AG_GLOBAL
#endif

constexpr int MAX_KEYS = 2 + AG_MAX_MATERIALS;  // miss, panel, one per material
constexpr int TRAV_BLOCK = 256;                  // threads per traversal block
#ifndef AG_STACK_LDS
#define AG_STACK_LDS 8
#endif
constexpr int STACK_LDS = AG_STACK_LDS;          // stack entries per lane in LDS
constexpr int STACK_SPILL = 96;                  // further entries per thread in HBM
constexpr int MAX_SEGMENTS = 4096;               // work segments per wave

// Hit codes written by the extension kernel.
constexpr uint32_t HIT_MISS = 0xFFFFFFFFu;
constexpr uint32_t HIT_PANEL = 0x80000000u;  // | panel index
constexpr uint32_t HIT_FLOOR = 0xC0000000u;

struct DMat {
    int32_t kind;
    float alpha, eta;
    const float* tn;       // metal n grid (95)
    const float* tk;       // metal k grid (95)
    int32_t has_patina;
    float p_strength, p_lo, p_hi, p_alpha, p_eta;
    const float* tpatina;  // 1 nm tables
    const float* ta;
    const float* tb;
    float opal_cell, opal_strength;
};

struct DPanel {
    float c[3], hu[3], hv[3], n[3];
    float area, hu2, hv2;  // area, |hu|^2, |hv|^2
    const float* emit;     // 1 nm table, null = flag
    int32_t camera_visible;
    float feather;
};

struct DScene {
    const float4* nodes;
    const float4* tris;
    const float* pos;       // n_vertices * 3
    const float* nrm;       // n_vertices * 3 or null
    const uint32_t* idx;    // n_triangles * 3
    const uint16_t* tmat;   // n_triangles
    const float* cavity;    // n_vertices or null
    uint32_t n_tris;
    int32_t n_mat;
    DMat mat[AG_MAX_MATERIALS];
    int32_t n_panels;
    DPanel panel[AG_MAX_PANELS];
    int32_t n_lights;
    int32_t lights[AG_MAX_PANELS];
    float light_cdf[AG_MAX_PANELS];
    float env_pick;
    int32_t env_kind;
    float env_power;
    const float* env_top;
    const float* env_hor;
    const float* env_bot;
    int32_t env_w, env_h;
    const float* env_weights;
    const float* env_marg;
    const float* env_cond;
    float env_rs, env_rc, env_rsn, env_rcn;  // sin/cos of +rotation and of -rotation (radians)
    const float* env_basis;  // 3 x 471
    int32_t floor_present;
    float floor_y;
    int32_t floor_mat;
    float floor_f0, floor_f1;
    const float* energy;     // 32 x 32
    const float* cmf;        // 471 x 3
    // The table blob's first win_n floats (materials, panel emission, the
    // gradient sky's spectra; at most SHADE_TAB_FLOATS), which the shading
    // kernel stages into LDS.
    const float* win;
    uint32_t win_n;
};
constexpr uint32_t SHADE_TAB_FLOATS = 8192;

// One view of a scene: camera, turntable, slot-0 material, film buffers.
struct DView {
    const DScene* scene;  // the scene this view renders (views of different scenes share waves)
    float pos[3], fwd[3], right[3], up[3];
    float half_w, half_h, lens_radius, focus_dist, shift0, shift1;
    float turn_deg;      // 0 = identity (no rotation, as the reference skips it)
    float ts, tc;        // sin, cos of turn_deg in radians (host libm)
    float tsn, tcn;      // sin, cos of -turn_deg
    int32_t mat0;        // material override index in the view-material block, -1 none
    uint32_t width, height;
    uint32_t seed_h;     // hash32(seed ^ 0xa511e9b3)
    float aspect_sx, aspect_sy;
    float cam_m[9];
    float* acc;          // pass accumulators: colour, albedo, normal, each w*h*3
    uint32_t max_depth, rr_depth;  // this view's PathSettings (integrator.rs:10-24): views of
    float clamp_indirect;          // one schedule may differ (a live view beside served finals)
    uint64_t fam;        // 2 bits per material index: its shading family (0 metal, 1 diffuse, 2 other)
};

// A contiguous range of path ids: `ns` samples starting at `s0` for pixels
// [p0, p0 + npix) of view `view`.  Path id = begin + (s - s0) * npix + (p - p0).
struct DSeg {
    uint32_t begin;
    uint32_t view;
    uint32_t p0, npix;
    uint32_t s0, ns;
};

// Path state, one slot per path (the view a path belongs to travels in its
// ray and shadow records instead).
struct DPaths {
    float4* beta;
    float4* L;
    float4* lam;
    float4* pdf;
    float* prev_pdf;
    uint32_t* flags;      // bit0 specular, bit1 rough_seen, bits 8..15 medium material (255 none)
    uint64_t* rng;        // next_random state
    uint32_t* pix_seed;
    uint32_t* si;
    float4* galb;         // guide albedo
    float4* gnrm;         // guide normal
};

struct DQueues {
    float4* ray_o[2];     // (o.xyz, path id bits)
    float4* ray_d[2];     // (d.xyz, view << 8 | depth bits)
    float4* hit;          // (t, u, v, code bits)
    uint32_t* bucket;     // MAX_KEYS-strided buckets of ray indices, capacity cap each
    float4* sh_o;         // (origin.xyz, path id bits)
    float4* sh_d;         // panel: (light point.xyz, ignore panel bits); sky: (dir.xyz, -1 bits)
    float4* sh_c;         // contribution spectrum
    float4* sh_x;         // (a, p_light, clamp or -1, view << 1 | kind bits)
    uint32_t* counters;   // per depth: [ray count, shadow count, key counts...]
    uint32_t* cursors;    // per depth: extension set, shadow set (CURSOR_SET u32 each)
    uint32_t cap;         // path slots (and each bucket's capacity)
    uint32_t qcap;        // slots of each ray, hit and shadow queue: cap and the chunk tails' room
    int32_t n_keys;
    uint32_t slices;      // render traversal: queue slices with a cursor each (1..CURSOR_SLICES)
    uint32_t bands;       // ray generation: a segment's camera rays queued in this many bands of
                          // its pixel range (1..CURSOR_SLICES), band-major
    int32_t ordered;      // 1: shading writes hit j's continuation ray to slot j of the next queue
                          // (DEAD_PATH where the path ended), keeping the hit order exactly
    int32_t sort;         // hits: 1 to their material's bucket; 0 all to bucket 0 in arrival order;
                          // 2 to bucket 0 in runs (one traversal wavefront's) ordered by material;
                          // 3 misses and panel hits to buckets 0 and 1, surface hits to bucket 2 in arrival order;
                          // 4 to the buckets of their shading family (KEY_*), one shading kernel per family
};

// Counters per depth slot, each on its own 128-byte line: device-scope
// atomics on different counters of one line serialise with each other (two
// bucket keys sharing a line made the traversal that appends to them 2.7x
// slower than one key alone).  Word offsets within a depth slot:
// [CNT_RAYS] extension queue length, [CNT_SHADOW] shadow queue length,
// [CNT_KEY + k * CNT_PAD] bucket k's size, [CNT_LIVE] live extension rays and
// [CNT_SHLIVE] shadow rays emitted (the queue lengths less their dead slots:
// slots of paths that ended when shading emits in order, the unused tails of
// reserved chunks).
constexpr int CNT_PAD = 32;  // u32 per counter (128 B)
constexpr int CNT_RAYS = 0;
constexpr int CNT_SHADOW = 1 * CNT_PAD;
constexpr int CNT_KEY = 2 * CNT_PAD;
constexpr int CNT_LIVE = (2 + MAX_KEYS) * CNT_PAD;
constexpr int CNT_SHLIVE = (3 + MAX_KEYS) * CNT_PAD;
constexpr int CNT_STRIDE = (4 + MAX_KEYS) * CNT_PAD;  // u32 per depth slot
// One 128-byte counter takes about 88 M device-scope atomics per second on
// the MI355X (11 ns each, returning or not): a queue that every shading
// wavefront appends to once per hit batch would bound a 2^29-hit launch at
// ~95 ms.  Shading therefore reserves queue slots SHADE_CHUNK at a time per
// wavefront once a launch shades at least SHADE_CHUNK_MIN hits, leaving dead
// slots only in the tail of each wavefront's last chunk; the queues carry
// room for those tails (SHADE_LAUNCHES_MAX shading launches per depth, each
// wavefront of each with one tail).
constexpr uint32_t SHADE_CHUNK = 512;
constexpr uint32_t SHADE_CHUNK_MIN = 1u << 26;
constexpr uint32_t SHADE_LAUNCHES_MAX = 4;
// Bucket keys of shading by family (sort mode 4): misses, panel hits, metal
// surfaces, diffuse surfaces (the floor when its material is diffuse), every
// other surface.
constexpr int KEY_MISS = 0, KEY_PANEL = 1, KEY_METAL = 2, KEY_DIFFUSE = 3, KEY_OTHER = 4;
constexpr int FAM_KEYS = 5;
// path id of a dead slot in an extension queue written in order
constexpr uint32_t DEAD_PATH = 0xFFFFFFFFu;
// Ray cursors of the persistent traversal kernels: a queue of n rays is cut
// into CURSOR_SLICES equal slices, one per XCD, each with its own cursor on
// its own 128-byte line; per depth slot there is one set for extension rays
// and one for shadow rays.
constexpr int CURSOR_SLICES = 8;
constexpr int CURSOR_STRIDE = 32;  // u32 per cursor (128 B)
constexpr int CURSOR_SET = CURSOR_SLICES * CURSOR_STRIDE;
constexpr int CURSORS_PER_DEPTH = 2 * CURSOR_SET;
constexpr int TALLY_DEPTHS = 64;  // per-depth ray counts kept for statistics

}  // namespace ag
