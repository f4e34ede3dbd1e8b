// Host runtime: contexts, device scenes, films and the wave scheduler behind
// ag_render_pass / ag_render_views, plus the ray-batch queries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/argent.h"
#include "dev_types.hpp"
#include "host_internal.hpp"
#include "host_spectra.hpp"
#include "kernels.hpp"
#include "mesh_internal.hpp"
#include "patina.hpp"
#include "post.hpp"
#include "tables/spectral_data.h"

using namespace ag;

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return fail(AG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t n) {
        if (n <= bytes && p) return AG_OK;
        release();
        if (n == 0) return AG_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            return fail(AG_E_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
        return AG_OK;
    }
    template <class T>
    T* as() const {
        return reinterpret_cast<T*>(p);
    }
};

uint32_t hash32_host(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

}  // namespace

// ------------------------------------------------------------------ allocation-free ray queries
// ag_intersect_device_noalloc runs the closest-hit traversal with its scratch
// -- the per-XCD ray cursors and the per-thread stack spill areas -- in this
// one array instead of library-allocated buffers.  The array is sized for a
// traversal grid of at most NOALLOC_BLOCKS blocks (the kernel pulls rays, so a
// smaller grid serves any batch size); every device that loads the library
// holds its own copy until the process ends.  Host code never uses the
// array's name as an address: each device's address comes from
// hipGetSymbolAddress at context creation (noalloc_scratch_setup).
constexpr uint32_t NOALLOC_BLOCKS = 512;
constexpr size_t NOALLOC_CURSOR_BYTES = 4096;  // CURSOR_SET u32 (1 KiB), padded
constexpr size_t NOALLOC_SPILL_BYTES = (size_t)NOALLOC_BLOCKS * TRAV_BLOCK * STACK_SPILL * sizeof(uint2);  // 96 MiB
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_CURSOR_BYTES + NOALLOC_SPILL_BYTES;
static_assert(CURSOR_SET * sizeof(uint32_t) <= NOALLOC_CURSOR_BYTES, "cursor area too small");
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Bytes of path state per path slot (DPaths: beta, L, lam, pdf, guide albedo
// and normal, previous pdf, flags, pixel seed, sample index, generator).
constexpr size_t PATH_STATE_BYTES = 6 * sizeof(float4) + sizeof(float) + 3 * sizeof(uint32_t) + sizeof(uint64_t);

struct ag_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 256;
    ag_tuning tuning{};
    // wave buffers
    uint32_t cap = 0;
    int n_keys_alloc = 0;
    DevBuf paths_buf, rays_buf, hit_buf, bucket_buf, shadow_buf, counter_buf, spill_buf, seg_buf, view_buf, vmat_buf,
        cursor_buf,
        vtab_buf;
    DPaths P{};
    DQueues Q{};
    int grid_trav = 0, grid_batch = 0, grid_work = 0;
    int max_trav_blocks = 1;   // resident render traversal blocks per CU (occupancy)
    int max_batch_blocks = 1;  // the same for the ray-batch kernels
    size_t spill_threads = 0;
    // timing: HIP event pairs around every stage launch on the context stream
    bool timing = false;
    ag_stats stats{};
    hipEvent_t ev[8] = {};
    std::vector<hipEvent_t> evpool;
    size_t ev_used = 0;
    struct Mark {
        int stage, depth;
        size_t ev;  // index of the first event of the pair
    };
    std::vector<Mark> ev_marks;
    std::vector<unsigned long long> depth_rays;  // [2 * TALLY_DEPTHS]: extension, then shadow
    std::vector<double> depth_ms;                // [3 * TALLY_DEPTHS]: extend, shade, shadow
    DevBuf tally_buf;
    DevBuf mesh_ws;  // mesh ingest temporaries (mesh.hip)
    DevBuf post_buf;  // preview development: images, à-trous work, grade tables, 8-bit output
    DevBuf energy_buf;  // the library's GGX energy table, staged into LDS by the shading kernel
    double last_query_ms = 0.0;
    void* noalloc_scratch = nullptr;  // this device's address of `scratch`
    // failure injection for tests (ag_debug_fail_seed): render passes with a
    // view whose seed hashes to this value fail before launching anything
    bool fail_seed_on = false;
    uint32_t fail_seed_h = 0;
    // captured render passes (launch_graph): the bytes of every launch
    // argument of a pass -> its executable graph, least recently used out
    struct Graph {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
        uint64_t used = 0;
    };
    std::vector<Graph> graphs;
    uint64_t graph_clock = 0;
    uint64_t graph_captures = 0, graph_replays = 0;
    hipEvent_t next_event() {
        if (ev_used == evpool.size()) {
            hipEvent_t e;
            (void)hipEventCreate(&e);
            evpool.push_back(e);
        }
        return evpool[ev_used++];
    }
};

// The piece (scene.rs:48 Arc<Piece>): BVH and mesh arrays on the device,
// shared by every scene built over the same mesh.
struct Geometry {
    DevBuf nodes, tris, pos, nrm, idx, tmat, cavity;
    uint32_t n_nodes = 0, n_leaves = 0, depth = 0, n_vertices = 0, n_triangles = 0;
    uint16_t max_material = 0;
    bool has_nrm = false, has_cavity = false;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};        // the BVH's bounds
    float vlo[3] = {0, 0, 0}, vhi[3] = {0, 0, 0};      // all vertex positions (Mesh::bounds)
    size_t bytes() const { return nodes.bytes + tris.bytes + pos.bytes + idx.bytes + tmat.bytes + nrm.bytes + cavity.bytes; }
};

struct ag_scene {
    ag_ctx* ctx = nullptr;
    std::shared_ptr<Geometry> geo;
    DScene hs{};  // host copy with device pointers
    DevBuf d_scene, tab, env_w, env_m, env_c;
    std::vector<ag_material> mats;
    size_t mat_stride = 0;  // floats per material in the table blob
    size_t mat_base = 0;
    size_t bytes = 0;
    float turntable_deg = 0.0f;
    int n_keys = 2;
    float cam_m[9];  // white-balanced XYZ -> linear sRGB (color.rs:48-55)
};

struct ag_film {
    ag_ctx* ctx = nullptr;
    uint32_t w = 0, h = 0;
    DevBuf sums, acc;
    uint32_t samples = 0;
    float cam_m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
};

// ------------------------------------------------------------------ context
// The GGX energy table (microfacet.rs:93-140) is a fixed function of GGX;
// every descriptor must carry the library's (ag_desc_init fills it).
static const std::vector<float>& library_energy() {
    static const std::vector<float> E = [] {
        std::vector<float> e(1024);
        ag_energy_table(e.data());
        return e;
    }();
    return E;
}

namespace ag {
ihipStream_t* ctx_stream(const ag_ctx* c) { return c->stream; }
int ctx_device(const ag_ctx* c) { return c->device; }
int ctx_workspace(ag_ctx* c, size_t bytes, void** p) {
    const int rc = c->mesh_ws.ensure(bytes);
    *p = c->mesh_ws.p;
    return rc;
}
}  // namespace ag

extern "C" int ag_ctx_create(int device, ag_ctx** out) {
    if (!out) return fail(AG_E_INVALID, "null output");
    int count = 0;
    HIPCHK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(AG_E_INVALID, "no such device " + std::to_string(device));
    HIPCHK(hipSetDevice(device));
    auto* c = new ag_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0) c->n_cu = prop.multiProcessorCount;
    {
        // this device's copy of the allocation-free queries' scratch array
        // (the device is current); its address is kept per context
        hipError_t e = hipGetSymbolAddress(&c->noalloc_scratch, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !c->noalloc_scratch) {
            delete c;
            return fail(AG_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(AG_E_DEVICE, "stream creation failed");
    }
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    {
        const std::vector<float>& E = library_energy();
        if (c->energy_buf.ensure(E.size() * 4) || hipMemcpy(c->energy_buf.p, E.data(), E.size() * 4,
                                                            hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipStreamDestroy(c->stream);
            delete c;
            return fail(AG_E_DEVICE, "energy table upload failed");
        }
    }
    // 2^26 path slots (~19 GB of path state, queues and buckets on a 288 GB
    // MI355X): few waves per render, so the long tail of the last bounces --
    // a handful of rays keeping the GPU for a dispatch each -- is paid rarely.
    c->tuning.paths_per_wave = 1u << 26;
    c->tuning.block_traverse = TRAV_BLOCK;
    c->max_trav_blocks = traversal_blocks_per_cu(false);
    c->max_batch_blocks = traversal_blocks_per_cu(true);
    // Render passes: extension rays in whole coherent batches of 64 rays per
    // wavefront at 5 blocks per CU; shadow rays, whose lengths within a batch
    // differ widely (some stop at a flag, some cross the piece), refill at 32
    // idle lanes; ray batches (often incoherent): refill at 32, full occupancy.
    // A render queue shorter than 64 rays per wavefront of the grid is spread
    // down to one ray per wavefront (a 270x480 2-spp pass at depth 32: 8.3 ->
    // 5.9 ms; a 64-spp still unchanged).
    c->tuning.blocks_per_cu = (uint32_t)std::min(5, c->max_trav_blocks);
    c->tuning.refill_min = 64;
    c->tuning.batch_blocks_per_cu = (uint32_t)c->max_batch_blocks;
    c->tuning.batch_refill_min = 32;
    c->tuning.shadow_refill_min = 32;
    c->tuning.spread_min_lanes = 1;
    // Material-sorted shading: wavefronts of misses, panel hits and each
    // material apart, and the continuation rays emitted material by
    // material, each run in the queue's spatial order.  (It traverses at
    // camera-ray speed only since every queue counter has its own cache
    // line; arrival order, 0, stays selectable: on the signet still at 2^26
    // paths per wave a 256-spp pass takes 535.6 ms against 588.5.)
    c->tuning.sort_by_material = 1;
    c->tuning.ordered_emission = 0;
    c->tuning.cursor_slices = CURSOR_SLICES;
    c->tuning.launch_graph = 0;
    // shading at 2 waves per SIMD, every value in registers: with the loop's
    // loads pipelined it matches the 3-wave instance (which spills ~70
    // registers) on the signet still, 219 vs 223 ms of shading per 256-spp pass
    c->tuning.shade_waves = 2;
    // camera rays in path order: queueing them by image band, one band per XCD
    // slice, made the signet still's first traversal slower (96.4 -> 101.5 ms:
    // the piece fills the middle bands, whose slices the other XCDs then drain)
    c->tuning.raygen_bands = 1;
    *out = c;
    return AG_OK;
}

extern "C" void ag_ctx_destroy(ag_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : c->evpool) (void)hipEventDestroy(e);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int ag_ctx_synchronize(ag_ctx* c) {
    if (!c) return fail(AG_E_INVALID, "null context");
    HIPCHK(hipStreamSynchronize(c->stream));
    return AG_OK;
}

extern "C" int ag_ctx_get_tuning(const ag_ctx* c, ag_tuning* t) {
    if (!c || !t) return fail(AG_E_INVALID, "null argument");
    *t = c->tuning;
    return AG_OK;
}

extern "C" int ag_ctx_graph_stats(const ag_ctx* c, uint64_t* captures, uint64_t* replays) {
    if (!c || !captures || !replays) return fail(AG_E_INVALID, "null argument");
    *captures = c->graph_captures;
    *replays = c->graph_replays;
    return AG_OK;
}

extern "C" int ag_ctx_set_tuning(ag_ctx* c, const ag_tuning* t) {
    if (!c || !t) return fail(AG_E_INVALID, "null argument");
    if (t->paths_per_wave < 1024 || t->paths_per_wave > AG_MAX_PATHS_PER_WAVE)
        return fail(AG_E_INVALID, "paths_per_wave must be 1024..2^29");
    if (t->block_traverse != TRAV_BLOCK) return fail(AG_E_INVALID, "block_traverse must be 256");
    if (t->blocks_per_cu < 1 || (int)t->blocks_per_cu > c->max_trav_blocks)
        return fail(AG_E_INVALID, "blocks_per_cu must be 1.." + std::to_string(c->max_trav_blocks) +
                                      " (resident render traversal blocks per CU)");
    if (t->batch_blocks_per_cu < 1 || (int)t->batch_blocks_per_cu > c->max_batch_blocks)
        return fail(AG_E_INVALID, "batch_blocks_per_cu must be 1.." + std::to_string(c->max_batch_blocks) +
                                      " (resident ray-batch traversal blocks per CU)");
    if (t->sort_by_material > 4) return fail(AG_E_INVALID, "sort_by_material must be 0..4");
    if (t->ordered_emission > 1) return fail(AG_E_INVALID, "ordered_emission must be 0 or 1");
    if (t->cursor_slices < 1 || t->cursor_slices > (uint32_t)CURSOR_SLICES)
        return fail(AG_E_INVALID, "cursor_slices must be 1.." + std::to_string(CURSOR_SLICES));
    if (t->launch_graph > 1) return fail(AG_E_INVALID, "launch_graph must be 0 or 1");
    if (t->shade_waves < 2 || t->shade_waves > 3) return fail(AG_E_INVALID, "shade_waves must be 2 or 3");
    if (t->raygen_bands < 1 || t->raygen_bands > (uint32_t)CURSOR_SLICES)
        return fail(AG_E_INVALID, "raygen_bands must be 1.." + std::to_string(CURSOR_SLICES));
    for (uint32_t r : {t->refill_min, t->batch_refill_min, t->shadow_refill_min, t->spread_min_lanes})
        if (r < 1 || r > 64)
            return fail(AG_E_INVALID,
                        "refill_min, batch_refill_min, shadow_refill_min and spread_min_lanes must be 1..64");
    c->tuning = *t;
    return AG_OK;
}

extern "C" int ag_ctx_stats(const ag_ctx* c, ag_stats* s) {
    if (!c || !s) return fail(AG_E_INVALID, "null argument");
    *s = c->stats;
    return AG_OK;
}

extern "C" int ag_ctx_depth_stats(const ag_ctx* c, int32_t n, uint64_t* ext, uint64_t* sh, double* ms_extend,
                                  double* ms_shade, double* ms_shadow) {
    if (!c) return fail(AG_E_INVALID, "null context");
    if (n < 0 || n > TALLY_DEPTHS) return fail(AG_E_INVALID, "n must be 0..64");
    for (int d = 0; d < n; d++) {
        bool have = c->depth_rays.size() == 2 * (size_t)TALLY_DEPTHS;
        bool timed = c->depth_ms.size() == 3 * (size_t)TALLY_DEPTHS;
        if (ext) ext[d] = have ? c->depth_rays[d] : 0;
        if (sh) sh[d] = have ? c->depth_rays[TALLY_DEPTHS + d] : 0;
        if (ms_extend) ms_extend[d] = timed ? c->depth_ms[d] : 0.0;
        if (ms_shade) ms_shade[d] = timed ? c->depth_ms[TALLY_DEPTHS + d] : 0.0;
        if (ms_shadow) ms_shadow[d] = timed ? c->depth_ms[2 * TALLY_DEPTHS + d] : 0.0;
    }
    return AG_OK;
}

extern "C" int ag_ctx_enable_timing(ag_ctx* c, int32_t on) {
    if (!c) return fail(AG_E_INVALID, "null context");
    c->timing = on != 0;
    return AG_OK;
}

static int ensure_spill(ag_ctx* c) {
    c->grid_trav = c->n_cu * (int)c->tuning.blocks_per_cu;
    c->grid_batch = c->n_cu * (int)c->tuning.batch_blocks_per_cu;
    c->grid_work = c->n_cu * 8;
    size_t threads = (size_t)std::max(c->grid_trav, c->grid_batch) * TRAV_BLOCK;
    if (threads > c->spill_threads || !c->spill_buf.p) {
        int rc = c->spill_buf.ensure(threads * STACK_SPILL * sizeof(uint2));
        if (rc) return rc;
        c->spill_threads = threads;
    }
    return AG_OK;
}

// Counter words per wave, rounded up to whole 128-byte lines; the cursors follow.
static size_t counter_words(uint32_t max_depth) {
    size_t n = (size_t)(max_depth + 2) * CNT_STRIDE;
    return (n + CURSOR_STRIDE - 1) / CURSOR_STRIDE * CURSOR_STRIDE;
}

// Bucket keys a wave needs: the materials', or the shading families' when
// shading by family
static int wave_keys(const ag_ctx* c, int n_keys) {
    return c->tuning.sort_by_material == 4 ? std::max(n_keys, FAM_KEYS) : n_keys;
}

// Slots of each ray, hit and shadow queue beyond the path slots: room for
// the dead tails of shading's reserved chunks (one per shading wavefront of
// each of up to SHADE_LAUNCHES_MAX launches per depth; the shading grid is
// n_cu * 8 blocks of 4 wavefronts), needed only when a launch can shade
// SHADE_CHUNK_MIN hits.
static size_t queue_slack(const ag_ctx* c, uint32_t cap) {
    return cap >= SHADE_CHUNK_MIN ? (size_t)SHADE_LAUNCHES_MAX * c->n_cu * 8 * (256 / 64) * SHADE_CHUNK : 0;
}

static int ensure_wave(ag_ctx* c, uint32_t cap, int n_keys, uint32_t max_depth) {
    int rc = ensure_spill(c);
    if (rc) return rc;
    if (cap != c->cap) {
        if (cap < c->cap) {  // give a large wave's memory back rather than keep it
            for (DevBuf* b : {&c->paths_buf, &c->rays_buf, &c->hit_buf, &c->shadow_buf, &c->bucket_buf}) b->release();
        }
        const size_t per_path = PATH_STATE_BYTES;
        if ((rc = c->paths_buf.ensure((size_t)cap * per_path))) return rc;
        const size_t qcap = (size_t)cap + queue_slack(c, cap);
        if ((rc = c->rays_buf.ensure(qcap * 4 * sizeof(float4)))) return rc;
        if ((rc = c->hit_buf.ensure(qcap * sizeof(float4)))) return rc;
        if ((rc = c->shadow_buf.ensure(qcap * 4 * sizeof(float4)))) return rc;
        char* b = c->paths_buf.as<char>();
        auto take = [&](size_t bytes) {
            char* p = b;
            b += bytes;
            return p;
        };
        c->P.beta = (float4*)take((size_t)cap * 16);
        c->P.L = (float4*)take((size_t)cap * 16);
        c->P.lam = (float4*)take((size_t)cap * 16);
        c->P.pdf = (float4*)take((size_t)cap * 16);
        c->P.galb = (float4*)take((size_t)cap * 16);
        c->P.gnrm = (float4*)take((size_t)cap * 16);
        c->P.rng = (uint64_t*)take((size_t)cap * 8);
        c->P.prev_pdf = (float*)take((size_t)cap * 4);
        c->P.flags = (uint32_t*)take((size_t)cap * 4);
        c->P.pix_seed = (uint32_t*)take((size_t)cap * 4);
        c->P.si = (uint32_t*)take((size_t)cap * 4);
        float4* r = c->rays_buf.as<float4>();
        c->Q.ray_o[0] = r;
        c->Q.ray_d[0] = r + qcap;
        c->Q.ray_o[1] = r + 2 * qcap;
        c->Q.ray_d[1] = r + 3 * qcap;
        c->Q.hit = c->hit_buf.as<float4>();
        float4* sh = c->shadow_buf.as<float4>();
        c->Q.sh_o = sh;
        c->Q.sh_d = sh + qcap;
        c->Q.sh_c = sh + 2 * qcap;
        c->Q.sh_x = sh + 3 * qcap;
        c->Q.qcap = (uint32_t)qcap;
        c->cap = cap;
        c->n_keys_alloc = 0;
    }
    if (n_keys > c->n_keys_alloc) {
        if ((rc = c->bucket_buf.ensure((size_t)cap * n_keys * sizeof(uint32_t)))) return rc;
        c->n_keys_alloc = n_keys;
    }
    size_t n_cnt = counter_words(max_depth);
    if ((rc = c->counter_buf.ensure((n_cnt + (size_t)(max_depth + 2) * CURSORS_PER_DEPTH) * sizeof(uint32_t)))) return rc;
    c->Q.bucket = c->bucket_buf.as<uint32_t>();
    c->Q.counters = c->counter_buf.as<uint32_t>();
    c->Q.cursors = c->Q.counters + n_cnt;
    c->Q.cap = cap;
    c->Q.n_keys = n_keys;
    c->Q.sort = (int32_t)c->tuning.sort_by_material;
    c->Q.ordered = c->tuning.ordered_emission ? 1 : 0;
    c->Q.slices = c->tuning.cursor_slices;
    c->Q.bands = c->tuning.raygen_bands;
    if ((rc = c->seg_buf.ensure(MAX_SEGMENTS * sizeof(DSeg)))) return rc;
    return AG_OK;
}

// ------------------------------------------------------------------ scene
static size_t tables_per_material() { return 2 * AG_GRID_N + 3 * AG_SPEC_N; }

static void fill_dmat(DMat& m, const ag_material& a, const float* dev_base) {
    m.kind = a.kind;
    m.alpha = a.alpha;
    m.eta = a.eta;
    m.tn = dev_base;
    m.tk = dev_base + AG_GRID_N;
    m.has_patina = a.has_patina;
    m.p_strength = a.patina_strength;
    m.p_lo = a.patina_lo;
    m.p_hi = a.patina_hi;
    m.p_alpha = a.patina_alpha;
    m.p_eta = a.patina_eta;
    m.tpatina = dev_base + 2 * AG_GRID_N;
    m.ta = dev_base + 2 * AG_GRID_N + AG_SPEC_N;
    m.tb = dev_base + 2 * AG_GRID_N + 2 * AG_SPEC_N;
    m.opal_cell = a.opal_cell_mm;
    m.opal_strength = a.opal_strength;
}

static void pack_material(const ag_material& a, float* dst) {
    std::memcpy(dst, a.metal_n, sizeof(a.metal_n));
    std::memcpy(dst + AG_GRID_N, a.metal_k, sizeof(a.metal_k));
    std::memcpy(dst + 2 * AG_GRID_N, a.patina_base, sizeof(a.patina_base));
    std::memcpy(dst + 2 * AG_GRID_N + AG_SPEC_N, a.spec_a, sizeof(a.spec_a));
    std::memcpy(dst + 2 * AG_GRID_N + 2 * AG_SPEC_N, a.spec_b, sizeof(a.spec_b));
}

static int validate_desc(const ag_scene_desc* d) {
    if (!d) return fail(AG_E_INVALID, "null descriptor");
    if (std::memcmp(d->energy, library_energy().data(), 1024 * sizeof(float)) != 0)
        return fail(AG_E_INVALID, "energy table differs from ag_energy_table()");
    if (d->n_triangles && (!d->positions || !d->triangles)) return fail(AG_E_INVALID, "missing mesh arrays");
    if (d->n_materials < 1 || d->n_materials > AG_MAX_MATERIALS) return fail(AG_E_INVALID, "n_materials must be 1..32");
    if (d->n_panels < 0 || d->n_panels > AG_MAX_PANELS) return fail(AG_E_INVALID, "n_panels must be 0..16");
    for (uint32_t i = 0; i < 3 * d->n_triangles; i++)
        if (d->triangles[i] >= d->n_vertices) return fail(AG_E_INVALID, "triangle index out of range");
    if (d->n_triangles >= (1u << 27)) return fail(AG_E_INVALID, "too many triangles (max 2^27)");
    if (d->tri_material)
        for (uint32_t i = 0; i < d->n_triangles; i++)
            if ((int)d->tri_material[i] >= d->n_materials) return fail(AG_E_INVALID, "triangle material out of range");
    if (d->floor.present && (d->floor.material < 0 || d->floor.material >= d->n_materials))
        return fail(AG_E_INVALID, "floor material out of range");
    for (int i = 0; i < d->n_materials; i++)
        if (d->materials[i].kind < AG_MAT_METAL || d->materials[i].kind > AG_MAT_OPAL)
            return fail(AG_E_INVALID, "unknown material kind");
    if (d->env.kind == AG_ENV_MAP && (d->env.width <= 0 || d->env.height <= 0 || !d->env.weights || !d->env.marginal ||
                                      !d->env.conditional))
        return fail(AG_E_INVALID, "incomplete environment map");
    return AG_OK;
}

static int upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
    int rc = b.ensure(bytes);
    if (rc) return rc;
    if (bytes) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
    return AG_OK;
}

static int build_geometry(ag_ctx* ctx, const ag_scene_desc* d, std::shared_ptr<Geometry>* out, bool gpu_bvh = false) {
    hipStream_t st = ctx->stream;
    auto g = std::make_shared<Geometry>();
    WideBvh bvh;
    if (!gpu_bvh) {
        build_wide_bvh(d->positions, d->triangles, d->n_triangles, &bvh);
    } else {
        // upload the mesh, then build the same node format on the device
        size_t nv = std::max<uint32_t>(d->n_vertices, 1), nt = std::max<uint32_t>(d->n_triangles, 1);
        int rc;
        if ((rc = g->pos.ensure(3 * nv * 4)) || (rc = g->idx.ensure(3 * nt * 4)) ||
            (rc = g->nodes.ensure(((size_t)d->n_triangles + 1) * BVH_NODE_FLOATS * 4)) ||
            (rc = g->tris.ensure(nt * 48)))
            return rc;
        if (d->n_vertices) HIPCHK(hipMemcpyAsync(g->pos.p, d->positions, 12 * (size_t)d->n_vertices, hipMemcpyHostToDevice, st));
        if (d->n_triangles)
            HIPCHK(hipMemcpyAsync(g->idx.p, d->triangles, 12 * (size_t)d->n_triangles, hipMemcpyHostToDevice, st));
        GpuBvhInfo info;
        if ((rc = gpu_build_wide_bvh(ctx, g->pos.as<float>(), g->idx.as<uint32_t>(), d->n_triangles,
                                     g->nodes.as<float>(), g->tris.as<float>(), &info)))
            return rc;
        // the node buffer was sized for the worst case (n_tris + 1 nodes): keep what the tree uses
        DevBuf exact;
        if ((rc = exact.ensure((size_t)info.n_nodes * BVH_NODE_FLOATS * 4))) return rc;
        HIPCHK(hipMemcpyAsync(exact.p, g->nodes.p, (size_t)info.n_nodes * BVH_NODE_FLOATS * 4, hipMemcpyDeviceToDevice,
                              st));
        HIPCHK(hipStreamSynchronize(st));
        std::swap(g->nodes.p, exact.p);
        std::swap(g->nodes.bytes, exact.bytes);
        bvh.n_nodes = info.n_nodes;
        bvh.n_leaves = info.n_leaves;
        bvh.depth = info.depth;
        bvh.max_stack = info.max_stack;
        for (int a = 0; a < 3; a++) {
            bvh.lo[a] = info.lo[a];
            bvh.hi[a] = info.hi[a];
        }
    }
    if (bvh.max_stack + 1 > (uint32_t)(STACK_LDS + STACK_SPILL))
        return fail(AG_E_INVALID, "BVH too deep for the traversal stack (" + std::to_string(bvh.max_stack) +
                                      " entries needed)");
    g->n_nodes = bvh.n_nodes;
    g->n_leaves = bvh.n_leaves;
    g->depth = bvh.depth;
    g->n_vertices = d->n_vertices;
    g->n_triangles = d->n_triangles;
    for (int a = 0; a < 3; a++) {
        g->lo[a] = bvh.lo[a];
        g->hi[a] = bvh.hi[a];
        // argent-mesh lib.rs:122-134: zeros for no vertices, else min/max over all of them
        g->vlo[a] = d->n_vertices ? FLT_MAX : 0.0f;
        g->vhi[a] = d->n_vertices ? -FLT_MAX : 0.0f;
        for (uint32_t v = 0; v < d->n_vertices; v++) {
            g->vlo[a] = std::fmin(g->vlo[a], d->positions[3 * (size_t)v + a]);
            g->vhi[a] = std::fmax(g->vhi[a], d->positions[3 * (size_t)v + a]);
        }
    }
    int rc;
    size_t nv = std::max<uint32_t>(d->n_vertices, 1), nt = std::max<uint32_t>(d->n_triangles, 1);
    if (!gpu_bvh) {
        if ((rc = upload(g->nodes, bvh.nodes.data(), bvh.nodes.size() * 4, st))) return rc;
        std::vector<float> dummy(12, 0.0f);
        if ((rc = upload(g->tris, bvh.tris.empty() ? dummy.data() : bvh.tris.data(),
                         std::max<size_t>(bvh.tris.size(), 12) * 4, st)))
            return rc;
        std::vector<float> zeros3(3 * nv, 0.0f);
        std::vector<uint32_t> zidx(3 * nt, 0);
        if ((rc = upload(g->pos, d->n_vertices ? d->positions : zeros3.data(), 3 * nv * 4, st))) return rc;
        if ((rc = upload(g->idx, d->n_triangles ? d->triangles : zidx.data(), 3 * nt * 4, st))) return rc;
    }
    std::vector<uint16_t> tm(nt, 0);
    if (d->tri_material && d->n_triangles) std::memcpy(tm.data(), d->tri_material, d->n_triangles * 2);
    for (uint16_t m : tm) g->max_material = std::max(g->max_material, m);
    if ((rc = upload(g->tmat, tm.data(), nt * 2, st))) return rc;
    g->has_nrm = d->normals && d->n_vertices;
    g->has_cavity = d->cavity && d->n_vertices;
    if (g->has_nrm)
        if ((rc = upload(g->nrm, d->normals, 3 * nv * 4, st))) return rc;
    if (g->has_cavity)
        if ((rc = upload(g->cavity, d->cavity, nv * 4, st))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    *out = std::move(g);
    return AG_OK;
}

// Everything but the mesh: materials, panels, environment, floor, white.
static int scene_over(ag_ctx* ctx, std::shared_ptr<Geometry> geo, const ag_scene_desc* d, ag_scene** out) {
    int rc;
    std::unique_ptr<ag_scene> s(new ag_scene());
    s->ctx = ctx;
    s->geo = std::move(geo);
    hipStream_t st = ctx->stream;
    // table blob: the tables shading reads per hit first and contiguous (the
    // materials', the panels' emission, the gradient sky's three spectra: the
    // prefix the shading kernel stages into LDS), then the env basis, the
    // energy table and the CMF
    s->mat_stride = tables_per_material();
    size_t n_tab = (size_t)d->n_materials * s->mat_stride + (size_t)d->n_panels * AG_SPEC_N + 6 * AG_SPEC_N + 1024 +
                   3 * AG_SPEC_N;
    std::vector<float> tab(n_tab, 0.0f);
    if ((rc = s->tab.ensure(n_tab * 4))) return rc;
    const float* dbase = s->tab.as<float>();
    DScene& S = s->hs;
    std::memset(&S, 0, sizeof(S));
    s->mats.assign(d->materials, d->materials + d->n_materials);
    for (int m = 0; m < d->n_materials; m++) {
        pack_material(d->materials[m], &tab[m * s->mat_stride]);
        fill_dmat(S.mat[m], d->materials[m], dbase + m * s->mat_stride);
    }
    size_t off = (size_t)d->n_materials * s->mat_stride;
    S.n_panels = d->n_panels;
    for (int i = 0; i < d->n_panels; i++) {
        const ag_panel& p = d->panels[i];
        DPanel& q = S.panel[i];
        for (int a = 0; a < 3; a++) {
            q.c[a] = p.center[a];
            q.hu[a] = p.hu[a];
            q.hv[a] = p.hv[a];
        }
        V3 hu(p.hu[0], p.hu[1], p.hu[2]), hv(p.hv[0], p.hv[1], p.hv[2]);
        V3 c = cross(hu, hv);
        V3 n = normalize(c);
        q.n[0] = n.x; q.n[1] = n.y; q.n[2] = n.z;
        q.area = 4.0f * length(c);
        q.hu2 = length_sq(hu);
        q.hv2 = length_sq(hv);
        q.camera_visible = p.camera_visible;
        q.feather = p.feather;
        if (p.has_emission) {
            std::memcpy(&tab[off + i * AG_SPEC_N], p.emission, AG_SPEC_N * 4);
            q.emit = dbase + off + i * AG_SPEC_N;
        } else {
            q.emit = nullptr;
        }
    }
    off += (size_t)d->n_panels * AG_SPEC_N;
    std::memcpy(&tab[off], d->env.top, AG_SPEC_N * 4);
    std::memcpy(&tab[off + AG_SPEC_N], d->env.horizon, AG_SPEC_N * 4);
    std::memcpy(&tab[off + 2 * AG_SPEC_N], d->env.bottom, AG_SPEC_N * 4);
    std::memcpy(&tab[off + 3 * AG_SPEC_N], d->env.basis, 3 * AG_SPEC_N * 4);
    S.env_top = dbase + off;
    S.env_hor = dbase + off + AG_SPEC_N;
    S.env_bot = dbase + off + 2 * AG_SPEC_N;
    S.env_basis = dbase + off + 3 * AG_SPEC_N;
    off += 6 * AG_SPEC_N;
    std::memcpy(&tab[off], d->energy, 1024 * 4);
    S.energy = dbase + off;
    off += 1024;
    std::memcpy(&tab[off], ag_data::CMF_1931_2DEG, 3 * AG_SPEC_N * 4);
    S.cmf = dbase + off;
    // the prefix the shading kernel stages into LDS: whole tables only (a
    // material's, a panel's, a sky spectrum), in blob order, while they fit,
    // so that every table is either staged entire or read from the blob
    {
        size_t n = 0;
        bool full = false;
        auto take = [&](size_t len) {
            if (!full && n + len <= SHADE_TAB_FLOATS) n += len;
            else full = true;
        };
        for (int m = 0; m < d->n_materials; m++) take(s->mat_stride);
        for (int i = 0; i < d->n_panels; i++) take(AG_SPEC_N);
        if (d->env.kind != AG_ENV_MAP)
            for (int i = 0; i < 3; i++) take(AG_SPEC_N);
        S.win = dbase;
        S.win_n = (uint32_t)n;
    }
    HIPCHK(hipMemcpyAsync(s->tab.p, tab.data(), n_tab * 4, hipMemcpyHostToDevice, st));
    S.env_kind = d->env.kind;
    S.env_power = d->env.power;
    if (d->env.kind == AG_ENV_MAP) {
        size_t w = (size_t)d->env.width, h = (size_t)d->env.height;
        if ((rc = upload(s->env_w, d->env.weights, w * h * 3 * 4, st))) return rc;
        if ((rc = upload(s->env_m, d->env.marginal, (h + 1) * 4, st))) return rc;
        if ((rc = upload(s->env_c, d->env.conditional, h * (w + 1) * 4, st))) return rc;
        S.env_w = (int)w;
        S.env_h = (int)h;
        S.env_weights = s->env_w.as<float>();
        S.env_marg = s->env_m.as<float>();
        S.env_cond = s->env_c.as<float>();
        const float k = 3.14159265358979323846f / 180.0f;
        float r = d->env.rotation_deg * k, rn = (-d->env.rotation_deg) * k;
        S.env_rs = std::sin(r);
        S.env_rc = std::cos(r);
        S.env_rsn = std::sin(rn);
        S.env_rcn = std::cos(rn);
    }
    int nl, lights[AG_MAX_PANELS];
    float cdf[AG_MAX_PANELS], env_pick;
    light_cdf(d, &nl, lights, cdf, &env_pick);
    S.n_lights = nl;
    for (int i = 0; i < nl; i++) {
        S.lights[i] = lights[i];
        S.light_cdf[i] = cdf[i];
    }
    S.env_pick = env_pick;
    S.floor_present = d->floor.present;
    S.floor_y = d->floor.y;
    S.floor_mat = d->floor.material;
    S.floor_f0 = d->floor.fade_start;
    S.floor_f1 = d->floor.fade_end;
    const Geometry& G = *s->geo;
    S.nodes = G.nodes.as<float4>();
    S.tris = G.tris.as<float4>();
    S.pos = G.pos.as<float>();
    S.nrm = G.has_nrm ? G.nrm.as<float>() : nullptr;
    S.idx = G.idx.as<uint32_t>();
    S.tmat = G.tmat.as<uint16_t>();
    S.cavity = G.has_cavity ? G.cavity.as<float>() : nullptr;
    S.n_tris = G.n_triangles;
    S.n_mat = d->n_materials;
    if ((rc = upload(s->d_scene, &S, sizeof(DScene), st))) return rc;
    HIPCHK(hipStreamSynchronize(st));
    s->turntable_deg = d->turntable_deg;
    s->n_keys = 2 + d->n_materials;
    camera_matrix(d->has_white != 0, d->white_xyz, s->cam_m);
    s->bytes = s->tab.bytes + s->env_w.bytes + s->env_m.bytes + s->env_c.bytes;
    *out = s.release();
    return AG_OK;
}

extern "C" int ag_scene_create(ag_ctx* ctx, const ag_scene_desc* d, ag_scene** out) {
    if (!ctx || !out) return fail(AG_E_INVALID, "null argument");
    int rc = validate_desc(d);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    std::shared_ptr<Geometry> g;
    if ((rc = build_geometry(ctx, d, &g))) return rc;
    return scene_over(ctx, std::move(g), d, out);
}

extern "C" int ag_scene_create_ex(ag_ctx* ctx, const ag_scene_desc* d, uint32_t flags, ag_scene** out) {
    if (!ctx || !out) return fail(AG_E_INVALID, "null argument");
    if (flags & ~(uint32_t)AG_SCENE_GPU_BVH) return fail(AG_E_INVALID, "unknown scene flags");
    int rc = validate_desc(d);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    std::shared_ptr<Geometry> g;
    if ((rc = build_geometry(ctx, d, &g, (flags & AG_SCENE_GPU_BVH) != 0))) return rc;
    return scene_over(ctx, std::move(g), d, out);
}

// Scene::new over a device mesh (ag_mesh): its positions, triangles and
// (when computed) normals are copied on the device and the BVH built there,
// or on the host after a read-back without AG_SCENE_GPU_BVH.
extern "C" int ag_scene_create_from_mesh(ag_ctx* ctx, const ag_mesh* mesh, const ag_scene_desc* d, uint32_t flags,
                                         ag_scene** out) {
    if (!ctx || !mesh || !d || !out) return fail(AG_E_INVALID, "null argument");
    if (mesh->ctx != ctx) return fail(AG_E_INVALID, "mesh belongs to another context");
    if (flags & ~(uint32_t)AG_SCENE_GPU_BVH) return fail(AG_E_INVALID, "unknown scene flags");
    if (mesh->nt >= (1u << 27)) return fail(AG_E_INVALID, "too many triangles (max 2^27)");
    ag_scene_desc look = *d;  // the mesh comes from `mesh`; the arrays in `d` are ignored
    look.positions = look.normals = look.cavity = nullptr;
    look.triangles = nullptr;
    look.tri_material = nullptr;
    look.n_vertices = look.n_triangles = 0;
    int rc = validate_desc(&look);
    if (rc) return rc;
    if (look.n_materials < 1) return fail(AG_E_INVALID, "the mesh needs material slot 0");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    ag_mesh* m = const_cast<ag_mesh*>(mesh);
    auto g = std::make_shared<Geometry>();
    if ((rc = mesh_check_bounds(m, g->vlo, g->vhi))) return rc;
    const uint32_t nv = mesh->nv, nt = mesh->nt;
    const size_t nv1 = std::max<uint32_t>(nv, 1), nt1 = std::max<uint32_t>(nt, 1);
    if ((rc = g->pos.ensure(12 * nv1)) || (rc = g->idx.ensure(12 * nt1)) || (rc = g->tmat.ensure(2 * nt1))) return rc;
    if (nv) HIPCHK(hipMemcpyAsync(g->pos.p, mesh->pos.p, 12 * (size_t)nv, hipMemcpyDeviceToDevice, st));
    if (nt) HIPCHK(hipMemcpyAsync(g->idx.p, mesh->tris.p, 12 * (size_t)nt, hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemsetAsync(g->tmat.p, 0, 2 * nt1, st));  // one material for the whole mesh (slot 0)
    if (mesh->has_nrm && nv) {
        if ((rc = g->nrm.ensure(12 * nv1))) return rc;
        HIPCHK(hipMemcpyAsync(g->nrm.p, mesh->nrm.p, 12 * (size_t)nv, hipMemcpyDeviceToDevice, st));
        g->has_nrm = true;
    }
    WideBvh bvh;
    if (flags & AG_SCENE_GPU_BVH) {
        if ((rc = g->nodes.ensure(((size_t)nt + 1) * BVH_NODE_FLOATS * 4)) || (rc = g->tris.ensure(nt1 * 48))) return rc;
        GpuBvhInfo info;
        if ((rc = gpu_build_wide_bvh(ctx, g->pos.as<float>(), g->idx.as<uint32_t>(), nt, g->nodes.as<float>(),
                                     g->tris.as<float>(), &info)))
            return rc;
        DevBuf exact;
        if ((rc = exact.ensure((size_t)info.n_nodes * BVH_NODE_FLOATS * 4))) return rc;
        HIPCHK(hipMemcpyAsync(exact.p, g->nodes.p, (size_t)info.n_nodes * BVH_NODE_FLOATS * 4,
                              hipMemcpyDeviceToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        std::swap(g->nodes.p, exact.p);
        std::swap(g->nodes.bytes, exact.bytes);
        bvh.n_nodes = info.n_nodes;
        bvh.n_leaves = info.n_leaves;
        bvh.depth = info.depth;
        bvh.max_stack = info.max_stack;
        for (int k = 0; k < 3; k++) {
            bvh.lo[k] = info.lo[k];
            bvh.hi[k] = info.hi[k];
        }
    } else {
        std::vector<float> hp(3 * (size_t)nv);
        std::vector<uint32_t> ht(3 * (size_t)nt);
        if (nv) HIPCHK(hipMemcpyAsync(hp.data(), mesh->pos.p, 12 * (size_t)nv, hipMemcpyDeviceToHost, st));
        if (nt) HIPCHK(hipMemcpyAsync(ht.data(), mesh->tris.p, 12 * (size_t)nt, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        build_wide_bvh(hp.data(), ht.data(), nt, &bvh);
        if ((rc = upload(g->nodes, bvh.nodes.data(), bvh.nodes.size() * 4, st))) return rc;
        std::vector<float> dummy(12, 0.0f);
        if ((rc = upload(g->tris, bvh.tris.empty() ? dummy.data() : bvh.tris.data(),
                         std::max<size_t>(bvh.tris.size(), 12) * 4, st)))
            return rc;
    }
    if (bvh.max_stack + 1 > (uint32_t)(STACK_LDS + STACK_SPILL))
        return fail(AG_E_INVALID, "BVH too deep for the traversal stack (" + std::to_string(bvh.max_stack) +
                                      " entries needed)");
    g->n_nodes = bvh.n_nodes;
    g->n_leaves = bvh.n_leaves;
    g->depth = bvh.depth;
    g->n_vertices = nv;
    g->n_triangles = nt;
    for (int k = 0; k < 3; k++) {
        g->lo[k] = bvh.lo[k];
        g->hi[k] = bvh.hi[k];
    }
    HIPCHK(hipStreamSynchronize(st));
    return scene_over(ctx, std::move(g), d, out);
}

extern "C" int ag_scene_create_shared(ag_ctx* ctx, const ag_scene* base, const ag_scene_desc* d, ag_scene** out) {
    if (!ctx || !base || !out) return fail(AG_E_INVALID, "null argument");
    if (base->ctx != ctx) return fail(AG_E_INVALID, "base scene belongs to another context");
    ag_scene_desc look = *d;  // the mesh comes from `base`; its arrays in `d` are ignored
    look.positions = look.normals = look.cavity = nullptr;
    look.triangles = nullptr;
    look.tri_material = nullptr;
    look.n_vertices = look.n_triangles = 0;
    int rc = validate_desc(&look);
    if (rc) return rc;
    if ((int)base->geo->max_material >= d->n_materials)
        return fail(AG_E_INVALID, "the mesh uses material slots beyond n_materials");
    HIPCHK(hipSetDevice(ctx->device));
    return scene_over(ctx, base->geo, d, out);
}

extern "C" void ag_scene_destroy(ag_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
}

extern "C" int ag_scene_set_turntable(ag_scene* s, float deg) {
    if (!s) return fail(AG_E_INVALID, "null scene");
    s->turntable_deg = deg;
    return AG_OK;
}

extern "C" int ag_scene_set_material(ag_scene* s, int32_t m, const ag_material* mat) {
    if (!s || !mat) return fail(AG_E_INVALID, "null argument");
    if (m < 0 || m >= s->hs.n_mat) return fail(AG_E_INVALID, "material slot out of range");
    if (mat->kind < AG_MAT_METAL || mat->kind > AG_MAT_OPAL) return fail(AG_E_INVALID, "unknown material kind");
    std::vector<float> t(s->mat_stride);
    pack_material(*mat, t.data());
    float* dst = s->tab.as<float>() + m * s->mat_stride;
    hipStream_t st = s->ctx->stream;
    HIPCHK(hipMemcpyAsync(dst, t.data(), t.size() * 4, hipMemcpyHostToDevice, st));
    fill_dmat(s->hs.mat[m], *mat, dst);
    HIPCHK(hipMemcpyAsync(s->d_scene.p, &s->hs, sizeof(DScene), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    s->mats[m] = *mat;
    return AG_OK;
}

extern "C" int ag_scene_bounds(const ag_scene* s, float* lo, float* hi) {
    if (!s) return fail(AG_E_INVALID, "null scene");
    for (int a = 0; a < 3; a++) {
        lo[a] = s->geo->lo[a];
        hi[a] = s->geo->hi[a];
    }
    return AG_OK;
}

extern "C" int ag_scene_stats(const ag_scene* s, uint32_t* nodes, uint32_t* leaves, uint32_t* depth, size_t* bytes) {
    if (!s) return fail(AG_E_INVALID, "null scene");
    if (nodes) *nodes = s->geo->n_nodes;
    if (leaves) *leaves = s->geo->n_leaves;
    if (depth) *depth = s->geo->depth;
    if (bytes) *bytes = s->bytes + s->geo->bytes();
    return AG_OK;
}

// ------------------------------------------------------------------ film
extern "C" int ag_film_create(ag_ctx* ctx, uint32_t w, uint32_t h, ag_film** out) {
    if (!ctx || !out || w == 0 || h == 0) return fail(AG_E_INVALID, "bad film arguments");
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<ag_film> f(new ag_film());
    f->ctx = ctx;
    f->w = w;
    f->h = h;
    size_t n = (size_t)w * h * 9 * sizeof(float);
    int rc;
    if ((rc = f->sums.ensure(n))) return rc;
    if ((rc = f->acc.ensure(n))) return rc;
    HIPCHK(hipMemsetAsync(f->sums.p, 0, n, ctx->stream));
    HIPCHK(hipMemsetAsync(f->acc.p, 0, n, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = f.release();
    return AG_OK;
}

extern "C" void ag_film_destroy(ag_film* f) {
    if (!f) return;
    (void)hipStreamSynchronize(f->ctx->stream);
    delete f;
}

extern "C" int ag_film_clear(ag_film* f) {
    if (!f) return fail(AG_E_INVALID, "null film");
    size_t n = (size_t)f->w * f->h * 9 * sizeof(float);
    HIPCHK(hipMemsetAsync(f->sums.p, 0, n, f->ctx->stream));
    HIPCHK(hipMemsetAsync(f->acc.p, 0, n, f->ctx->stream));
    HIPCHK(hipStreamSynchronize(f->ctx->stream));
    f->samples = 0;
    return AG_OK;
}

extern "C" uint32_t ag_film_samples(const ag_film* f) { return f ? f->samples : 0; }

extern "C" int ag_film_read_sums(const ag_film* f, int32_t which, float* out) {
    if (!f || !out) return fail(AG_E_INVALID, "null argument");
    if (which < AG_FILM_COLOR || which > AG_FILM_NORMAL) return fail(AG_E_INVALID, "bad plane");
    size_t plane = (size_t)f->w * f->h * 3;
    HIPCHK(hipMemcpyAsync(out, f->sums.as<float>() + which * plane, plane * 4, hipMemcpyDeviceToHost, f->ctx->stream));
    HIPCHK(hipStreamSynchronize(f->ctx->stream));
    return AG_OK;
}

// render.rs:35-64
extern "C" int ag_film_read(const ag_film* f, int32_t which, float* out) {
    if (!f || !out) return fail(AG_E_INVALID, "null argument");
    int src = which == AG_FILM_XYZ ? AG_FILM_COLOR : which;
    int rc = ag_film_read_sums(f, src, out);
    if (rc) return rc;
    size_t n = (size_t)f->w * f->h;
    float k = 1.0f / (float)std::max(f->samples, 1u);
    if (which == AG_FILM_NORMAL) {
        for (size_t i = 0; i < n; i++) {
            float v[3] = {out[3 * i] * k, out[3 * i + 1] * k, out[3 * i + 2] * k};
            float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            for (int c = 0; c < 3; c++) out[3 * i + c] = l > 1e-6f ? v[c] / l : 0.0f;
        }
        return AG_OK;
    }
    for (size_t i = 0; i < 3 * n; i++) out[i] *= k;
    if (which == AG_FILM_XYZ) {
        // invert the camera matrix used for the last pass
        double m[9], inv[9];
        for (int i = 0; i < 9; i++) m[i] = f->cam_m[i];
        double det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
        inv[0] = (m[4] * m[8] - m[5] * m[7]) / det;
        inv[1] = (m[2] * m[7] - m[1] * m[8]) / det;
        inv[2] = (m[1] * m[5] - m[2] * m[4]) / det;
        inv[3] = (m[5] * m[6] - m[3] * m[8]) / det;
        inv[4] = (m[0] * m[8] - m[2] * m[6]) / det;
        inv[5] = (m[2] * m[3] - m[0] * m[5]) / det;
        inv[6] = (m[3] * m[7] - m[4] * m[6]) / det;
        inv[7] = (m[1] * m[6] - m[0] * m[7]) / det;
        inv[8] = (m[0] * m[4] - m[1] * m[3]) / det;
        for (size_t i = 0; i < n; i++) {
            double r = out[3 * i], g = out[3 * i + 1], b = out[3 * i + 2];
            for (int c = 0; c < 3; c++) out[3 * i + c] = (float)(inv[3 * c] * r + inv[3 * c + 1] * g + inv[3 * c + 2] * b);
        }
    }
    return AG_OK;
}

// ------------------------------------------------------------------ rendering
namespace {

struct ViewJob {
    const ag_scene* scene = nullptr;
    DView dv;
    ag_film* film;
    uint32_t count;
    uint32_t first;
};

// mat0: the view-material block index that replaces material 0 (-1 none),
// mat0_kind that material's kind
void make_dview(DView& v, const ag_scene* sc, const ag_camera* cam, float turn_deg, int mat0, int mat0_kind,
                const ag_film* film, uint32_t seed, const ag_path_settings* ps) {
    std::memset(&v, 0, sizeof(v));
    v.scene = sc->d_scene.as<DScene>();
    const float* cam_m_in = sc->cam_m;
    for (int a = 0; a < 3; a++) {
        v.pos[a] = cam->pos[a];
        v.fwd[a] = cam->forward[a];
        v.right[a] = cam->right[a];
        v.up[a] = cam->up[a];
    }
    v.half_w = cam->half_w;
    v.half_h = cam->half_h;
    v.lens_radius = cam->lens_radius;
    v.focus_dist = cam->focus_dist;
    v.shift0 = cam->shift[0];
    v.shift1 = cam->shift[1];
    v.turn_deg = turn_deg;
    const float k = 3.14159265358979323846f / 180.0f;
    float r = turn_deg * k, rn = (-turn_deg) * k;
    v.ts = std::sin(r);
    v.tc = std::cos(r);
    v.tsn = std::sin(rn);
    v.tcn = std::cos(rn);
    v.mat0 = mat0;
    v.width = film->w;
    v.height = film->h;
    v.seed_h = hash32_host(seed ^ 0xa511e9b3u);
    v.aspect_sx = 2.0f / (float)film->w;
    v.aspect_sy = 2.0f / (float)film->h;
    for (int i = 0; i < 9; i++) v.cam_m[i] = cam_m_in[i];
    v.acc = film->acc.as<float>();
    v.max_depth = ps->max_depth;
    v.rr_depth = ps->rr_depth;
    v.clamp_indirect = ps->clamp_indirect;
    // each material's shading family (metal, diffuse, other); indices no
    // surface carries stay "other", which the general kernel shades
    v.fam = 0;
    for (int m = 0; m < AG_MAX_MATERIALS; m++) {
        int kind = m < (int)sc->mats.size() ? (int)sc->mats[m].kind : -1;
        if (m == 0 && mat0 >= 0) kind = mat0_kind;
        const uint64_t f = kind == AG_MAT_METAL ? 0u : kind == AG_MAT_DIFFUSE ? 1u : 2u;
        v.fam |= f << (2 * m);
    }
}

}  // namespace

// A pass enqueued by ag_render_pass_async: completion event, its own ray
// tallies, and the host sources of its copies, kept until it is done.
struct ag_event {
    ag_ctx* ctx = nullptr;
    hipEvent_t done = nullptr;
    DevBuf tally;
    std::vector<std::vector<DSeg>> waves;
    std::vector<DView> dv;
    uint64_t samples = 0;
    uint32_t launches = 0;
    std::chrono::steady_clock::time_point t0;
    bool finished = false;
};

static void set_stats(ag_ctx* c, const std::vector<unsigned long long>& tally, uint32_t launches, uint64_t samples,
                      double ms_total) {
    c->stats.ms_total = ms_total;
    c->stats.samples = samples;
    c->stats.extension_rays = tally[0];
    c->stats.shadow_rays = tally[1];
    c->stats.extend_launches = launches;
    c->stats.shadow_launches = launches;
    c->depth_rays.assign(tally.begin() + 2, tally.end());
}

// With `ev` the pass is only enqueued (no kernel timing); statistics come
// with ag_event_wait.
static int run_views(ag_ctx* c, const ag_scene* s, std::vector<ViewJob>& jobs, const DMat* vmats_dev,
                     ag_event* ev = nullptr) {
    // every view carries its own path settings; the bounce loop runs to the deepest
    ag_path_settings launch{0, 0.0f, 0};
    // a ray carries its view in 24 bits of its w word
    if (jobs.size() > (1u << 24)) return fail(AG_E_INVALID, "more than 2^24 views in one render");
    for (auto& j : jobs) {
        if (j.dv.max_depth > 64) return fail(AG_E_INVALID, "max_depth above 64");
        launch.max_depth = std::max(launch.max_depth, j.dv.max_depth);
    }
    if (c->fail_seed_on)
        for (auto& j : jobs)
            if (j.dv.seed_h == c->fail_seed_h) return fail(AG_E_DEVICE, "injected render failure (ag_debug_fail_seed)");
    const ag_path_settings* ps = &launch;
    HIPCHK(hipSetDevice(c->device));
    uint32_t cap = c->tuning.paths_per_wave;
    int n_keys = s->n_keys;
    for (auto& j : jobs) n_keys = std::max(n_keys, j.scene->n_keys);
    int rc = ensure_wave(c, cap, wave_keys(c, n_keys), ps->max_depth);
    if (rc) return rc;
    hipStream_t st = c->stream;
    // upload views
    std::vector<DView> dv(jobs.size());
    for (size_t i = 0; i < jobs.size(); i++) dv[i] = jobs[i].dv;
    if ((rc = c->view_buf.ensure(dv.size() * sizeof(DView)))) return rc;
    HIPCHK(hipMemcpyAsync(c->view_buf.p, dv.data(), dv.size() * sizeof(DView), hipMemcpyHostToDevice, st));
    const DView* views = c->view_buf.as<DView>();
    // segments, view-major, sample-chunk-major inside a view
    std::vector<DSeg> all;
    for (size_t v = 0; v < jobs.size(); v++) {
        uint32_t npix = jobs[v].film->w * jobs[v].film->h;
        uint32_t s0 = jobs[v].first, s1 = jobs[v].first + jobs[v].count;
        if (npix <= cap) {
            uint32_t spw = std::max(1u, cap / npix);
            for (uint32_t a = s0; a < s1; a += spw) all.push_back(DSeg{0, (uint32_t)v, 0, npix, a, std::min(spw, s1 - a)});
        } else {
            for (uint32_t a = s0; a < s1; a++)
                for (uint32_t p = 0; p < npix; p += cap) all.push_back(DSeg{0, (uint32_t)v, p, std::min(cap, npix - p), a, 1});
        }
    }
    // pack segments into waves, splitting sample ranges to fill
    std::vector<std::vector<DSeg>> waves;
    std::vector<DSeg> cur;
    uint32_t used = 0;
    for (size_t i = 0; i < all.size(); i++) {
        DSeg sg = all[i];
        while (sg.ns > 0) {
            uint32_t room = cap - used;
            uint32_t fit = room / sg.npix;
            if (fit == 0 || cur.size() >= (size_t)MAX_SEGMENTS) {
                waves.push_back(cur);
                cur.clear();
                used = 0;
                continue;
            }
            uint32_t take = std::min(fit, sg.ns);
            DSeg part = sg;
            part.begin = used;
            part.ns = take;
            cur.push_back(part);
            used += take * sg.npix;
            sg.s0 += take;
            sg.ns -= take;
        }
    }
    if (!cur.empty()) waves.push_back(cur);
    c->stats = ag_stats{};
    c->ev_used = 0;
    c->ev_marks.clear();
    const size_t n_tally = 2 + 2 * (size_t)TALLY_DEPTHS;
    DevBuf& tally_buf = ev ? ev->tally : c->tally_buf;
    if ((rc = tally_buf.ensure(n_tally * sizeof(unsigned long long)))) return rc;
    auto timed = [&](int stage, int depth, auto&& launch) {
        if (!c->timing || ev) {
            launch();
            return;
        }
        size_t i0 = c->ev_used;
        hipEvent_t a = c->next_event();
        hipEvent_t b = c->next_event();
        (void)hipEventRecord(a, st);
        launch();
        (void)hipEventRecord(b, st);
        c->ev_marks.push_back({stage, depth, i0});
    };
    // Shadow rays of depth d and extension rays of depth d + 1 run as one
    // launch, except with per-stage timing on, where each stage keeps its own
    // launch so that it can be timed (the results are the same either way).
    const bool fuse = !c->timing || ev;
    const uint32_t spread = c->tuning.spread_min_lanes;
    auto t0 = std::chrono::steady_clock::now();
    uint32_t launches = 0;
    const uint32_t n_counter_words =
        (uint32_t)(counter_words(ps->max_depth) + (size_t)(ps->max_depth + 2) * CURSORS_PER_DEPTH);
    // Every launch of one wave (its segments already in seg_buf).  Kernels
    // only: no copies or memsets, so that the sequence can be captured.
    auto enqueue_wave = [&](uint32_t nseg, uint32_t npaths) {
        launch_zero_u32(st, c->Q.counters, n_counter_words);
        timed(3, -1, [&] { launch_raygen(st, c->grid_work, c->seg_buf.as<DSeg>(), (int)nseg, npaths, views, c->P, c->Q); });
        for (uint32_t d = 0; d < ps->max_depth; d++) {
            if (!fuse || d == 0)
                timed(0, (int)d, [&] { launch_extend(st, c->grid_trav, views, c->P, c->Q, (int)d, c->tuning.refill_min, spread, c->spill_buf.as<uint2>()); });
            timed(1, (int)d, [&] {
                launch_shade(st, c->grid_work, views, vmats_dev, ps->max_depth, c->energy_buf.as<float>(), c->P,
                             c->Q, (int)d, (int)c->tuning.shade_waves, jobs[0].scene->hs.win_n);
            });
            if (fuse && d + 1 < ps->max_depth)
                launch_shadow_extend(st, c->grid_trav, views, c->P, c->Q, (int)d, c->tuning.shadow_refill_min,
                                     c->tuning.refill_min, spread, c->spill_buf.as<uint2>());
            else
                timed(2, (int)d, [&] { launch_shadow(st, c->grid_trav, views, c->P, c->Q, (int)d, c->tuning.shadow_refill_min, spread, c->spill_buf.as<uint2>()); });
        }
        timed(4, -1, [&] { launch_accumulate(st, c->grid_work, c->seg_buf.as<DSeg>(), (int)nseg, views, s->hs.cmf, c->P); });
        launch_tally(st, c->Q.counters, ps->max_depth, tally_buf.as<unsigned long long>());
    };
    auto enqueue_commits = [&]() {
        for (auto& j : jobs) {
            size_t n = (size_t)j.film->w * j.film->h * 9;
            launch_film_commit(st, j.film->sums.as<float>(), j.film->acc.as<float>(), n);
        }
    };
    const uint32_t tally_words = (uint32_t)(2 * n_tally);
    if (c->tuning.launch_graph && fuse && !ev && waves.size() == 1) {
        // one wave: replay the captured pass (the views and segments it reads
        // are copied into device memory first, outside the graph)
        auto& w = waves[0];
        uint32_t npaths = 0;
        for (auto& sg : w) npaths = std::max(npaths, sg.begin + sg.npix * sg.ns);
        HIPCHK(hipMemcpyAsync(c->seg_buf.p, w.data(), w.size() * sizeof(DSeg), hipMemcpyHostToDevice, st));
        std::vector<unsigned char> key;
        auto put = [&](const void* p, size_t n) {
            const unsigned char* b = static_cast<const unsigned char*>(p);
            key.insert(key.end(), b, b + n);
        };
        const void* ptrs[] = {views, c->seg_buf.p, vmats_dev, c->energy_buf.p, s->hs.cmf, tally_buf.p, c->spill_buf.p};
        const uint32_t dims[] = {(uint32_t)c->grid_work, (uint32_t)c->grid_trav, ps->max_depth, npaths,
                                 (uint32_t)w.size(), n_counter_words, tally_words};
        put(&c->P, sizeof c->P);
        put(&c->Q, sizeof c->Q);
        put(&c->tuning, sizeof c->tuning);
        put(ptrs, sizeof ptrs);
        put(dims, sizeof dims);
        for (auto& j : jobs) {
            const void* fp[] = {j.film->sums.p, j.film->acc.p};
            const size_t n = (size_t)j.film->w * j.film->h * 9;
            put(fp, sizeof fp);
            put(&n, sizeof n);
        }
        ag_ctx::Graph* g = nullptr;
        for (auto& e : c->graphs)
            if (e.key == key) g = &e;
        if (!g) {
            hipGraph_t graph = nullptr;
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_zero_u32(st, tally_buf.as<uint32_t>(), tally_words);
            enqueue_wave((uint32_t)w.size(), npaths);
            enqueue_commits();
            const hipError_t ce = hipStreamEndCapture(st, &graph);
            if (ce != hipSuccess) return fail(AG_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
            hipGraphExec_t exec = nullptr;
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return fail(AG_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
            constexpr size_t MAX_GRAPHS = 8;
            if (c->graphs.size() >= MAX_GRAPHS) {
                auto old = std::min_element(c->graphs.begin(), c->graphs.end(),
                                            [](const ag_ctx::Graph& a, const ag_ctx::Graph& b) { return a.used < b.used; });
                (void)hipGraphExecDestroy(old->exec);
                c->graphs.erase(old);
            }
            c->graphs.push_back(ag_ctx::Graph{std::move(key), exec, 0});
            g = &c->graphs.back();
            c->graph_captures++;
        } else {
            c->graph_replays++;
        }
        g->used = ++c->graph_clock;
        HIPCHK(hipGraphLaunch(g->exec, st));
        launches += ps->max_depth;
        c->stats.samples += npaths;
    } else {
        launch_zero_u32(st, tally_buf.as<uint32_t>(), tally_words);
        for (auto& w : waves) {
            uint32_t npaths = 0;
            for (auto& sg : w) npaths = std::max(npaths, sg.begin + sg.npix * sg.ns);
            HIPCHK(hipMemcpyAsync(c->seg_buf.p, w.data(), w.size() * sizeof(DSeg), hipMemcpyHostToDevice, st));
            enqueue_wave((uint32_t)w.size(), npaths);
            launches += ps->max_depth;
            c->stats.samples += npaths;
        }
        enqueue_commits();
    }
    if (ev) {
        HIPCHK(hipGetLastError());
        if (!ev->done) HIPCHK(hipEventCreateWithFlags(&ev->done, hipEventDisableTiming));
        HIPCHK(hipEventRecord(ev->done, st));
        ev->waves = std::move(waves);
        ev->dv = std::move(dv);
        ev->samples = c->stats.samples;
        ev->launches = launches;
        ev->t0 = t0;
        for (auto& j : jobs) {
            j.film->samples += j.count;
            std::memcpy(j.film->cam_m, j.dv.cam_m, sizeof(j.dv.cam_m));
        }
        return AG_OK;
    }
    std::vector<unsigned long long> tally(n_tally, 0);
    HIPCHK(hipMemcpyAsync(tally.data(), tally_buf.p, n_tally * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
    auto t1 = std::chrono::steady_clock::now();
    set_stats(c, tally, launches, c->stats.samples, std::chrono::duration<double, std::milli>(t1 - t0).count());
    c->depth_ms.assign(3 * (size_t)TALLY_DEPTHS, 0.0);
    for (auto& m : c->ev_marks) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->evpool[m.ev], c->evpool[m.ev + 1]);
        double* dst = m.stage == 0 ? &c->stats.ms_extend
                    : m.stage == 1 ? &c->stats.ms_shade
                    : m.stage == 2 ? &c->stats.ms_shadow
                    : m.stage == 3 ? &c->stats.ms_raygen : &c->stats.ms_accum;
        *dst += ms;
        if (m.stage <= 2 && m.depth >= 0 && m.depth < TALLY_DEPTHS) c->depth_ms[(size_t)m.stage * TALLY_DEPTHS + m.depth] += ms;
    }
    for (auto& j : jobs) {
        j.film->samples += j.count;
        std::memcpy(j.film->cam_m, j.dv.cam_m, sizeof(j.dv.cam_m));
    }
    return AG_OK;
}

extern "C" int ag_render_pass(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, ag_film* film,
                              const ag_path_settings* ps, uint32_t count, uint32_t seed) {
    if (!ctx || !scene || !cam || !film || !ps) return fail(AG_E_INVALID, "null argument");
    if (film->ctx != ctx || scene->ctx != ctx) return fail(AG_E_INVALID, "objects belong to another context");
    if (count == 0) return AG_OK;
    if ((uint64_t)film->samples + count > 0xFFFFFFFFull) return fail(AG_E_INVALID, "sample index overflow");
    ViewJob j;
    j.scene = scene;
    j.film = film;
    j.count = count;
    j.first = film->samples;
    make_dview(j.dv, scene, cam, scene->turntable_deg, -1, -1, film, seed, ps);
    std::vector<ViewJob> jobs{j};
    return run_views(ctx, scene, jobs, nullptr);
}

extern "C" int ag_render_pass_async(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, ag_film* film,
                                    const ag_path_settings* ps, uint32_t count, uint32_t seed, ag_event** out) {
    if (!ctx || !scene || !cam || !film || !ps || !out) return fail(AG_E_INVALID, "null argument");
    if (film->ctx != ctx || scene->ctx != ctx) return fail(AG_E_INVALID, "objects belong to another context");
    if ((uint64_t)film->samples + count > 0xFFFFFFFFull) return fail(AG_E_INVALID, "sample index overflow");
    std::unique_ptr<ag_event> ev(new ag_event());
    ev->ctx = ctx;
    if (count > 0) {
        ViewJob j;
        j.scene = scene;
        j.film = film;
        j.count = count;
        j.first = film->samples;
        make_dview(j.dv, scene, cam, scene->turntable_deg, -1, -1, film, seed, ps);
        std::vector<ViewJob> jobs{j};
        int rc = run_views(ctx, scene, jobs, nullptr, ev.get());
        if (rc) return rc;
    } else {
        ev->finished = true;
    }
    *out = ev.release();
    return AG_OK;
}

extern "C" int ag_event_query(const ag_event* ev) {
    if (!ev) return fail(AG_E_INVALID, "null event");
    if (ev->finished || !ev->done) return AG_OK;
    hipError_t e = hipEventQuery(ev->done);
    if (e == hipSuccess) return AG_OK;
    if (e == hipErrorNotReady) return AG_PENDING;
    return fail(AG_E_DEVICE, std::string("event query: ") + hipGetErrorString(e));
}

extern "C" int ag_event_wait(ag_event* ev) {
    if (!ev) return fail(AG_E_INVALID, "null event");
    if (ev->finished) return AG_OK;
    ag_ctx* c = ev->ctx;
    HIPCHK(hipSetDevice(c->device));
    const size_t n_tally = 2 + 2 * (size_t)TALLY_DEPTHS;
    std::vector<unsigned long long> tally(n_tally, 0);
    HIPCHK(hipEventSynchronize(ev->done));
    HIPCHK(hipMemcpy(tally.data(), ev->tally.p, n_tally * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    auto t1 = std::chrono::steady_clock::now();
    c->stats = ag_stats{};
    set_stats(c, tally, ev->launches, ev->samples, std::chrono::duration<double, std::milli>(t1 - ev->t0).count());
    c->depth_ms.assign(3 * (size_t)TALLY_DEPTHS, 0.0);
    ev->finished = true;
    ev->waves.clear();
    ev->dv.clear();
    return AG_OK;
}

extern "C" void ag_event_destroy(ag_event* ev) {
    if (!ev) return;
    (void)hipSetDevice(ev->ctx->device);
    if (ev->done) {
        (void)hipEventSynchronize(ev->done);
        (void)hipEventDestroy(ev->done);
    }
    delete ev;
}

static int upload_view_materials(ag_ctx* ctx, const ag_material* materials, int32_t n_materials) {
    int rc;
    size_t stride = tables_per_material();
    std::vector<float> tab(std::max<size_t>(1, n_materials) * stride, 0.0f);
    if ((rc = ctx->vtab_buf.ensure(tab.size() * 4))) return rc;
    std::vector<DMat> dm(std::max(1, n_materials));
    for (int i = 0; i < n_materials; i++) {
        if (materials[i].kind < AG_MAT_METAL || materials[i].kind > AG_MAT_OPAL) return fail(AG_E_INVALID, "unknown material kind");
        if (materials[i].kind == AG_MAT_GEM) return fail(AG_E_INVALID, "slot-0 override cannot be a transparent gem");
        pack_material(materials[i], &tab[i * stride]);
        fill_dmat(dm[i], materials[i], ctx->vtab_buf.as<float>() + i * stride);
    }
    HIPCHK(hipMemcpyAsync(ctx->vtab_buf.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = ctx->vmat_buf.ensure(dm.size() * sizeof(DMat)))) return rc;
    HIPCHK(hipMemcpyAsync(ctx->vmat_buf.p, dm.data(), dm.size() * sizeof(DMat), hipMemcpyHostToDevice, ctx->stream));
    return AG_OK;
}

extern "C" int ag_render_views(ag_ctx* ctx, const ag_scene* scene, const ag_material* materials, int32_t n_materials,
                               const ag_view* views, int32_t n_views, const ag_path_settings* ps) {
    if (!ctx || !scene || !views || !ps || n_views <= 0) return fail(AG_E_INVALID, "bad arguments");
    if (n_materials < 0 || (n_materials > 0 && !materials)) return fail(AG_E_INVALID, "bad materials");
    HIPCHK(hipSetDevice(ctx->device));
    int rc = upload_view_materials(ctx, materials, n_materials);
    if (rc) return rc;
    std::vector<ViewJob> jobs;
    for (int i = 0; i < n_views; i++) {
        const ag_view& v = views[i];
        if (!v.film || v.film->ctx != ctx) return fail(AG_E_INVALID, "view film missing or from another context");
        if (v.material0 >= n_materials) return fail(AG_E_INVALID, "view material index out of range");
        for (int k = 0; k < i; k++)
            if (views[k].film == v.film) return fail(AG_E_INVALID, "two views share one film");
        if (v.count == 0) continue;
        if ((uint64_t)v.film->samples + v.count > 0xFFFFFFFFull) return fail(AG_E_INVALID, "sample index overflow");
        ViewJob j;
        j.film = v.film;
        j.count = v.count;
        j.first = v.film->samples;
        const ag_scene* vs = v.scene ? v.scene : scene;
        if (vs->ctx != ctx) return fail(AG_E_INVALID, "view scene from another context");
        make_dview(j.dv, vs, &v.camera, v.turntable_deg, v.material0, v.material0 >= 0 ? (int)materials[v.material0].kind : -1,
                   v.film, v.seed, ps);
        j.scene = vs;
        jobs.push_back(j);
    }
    if (jobs.empty()) return AG_OK;
    return run_views(ctx, scene, jobs, ctx->vmat_buf.as<DMat>());
}

// ------------------------------------------------------------------ server
struct ag_server {
    ag_ctx* ctx = nullptr;
    // HBM accounting: films of admitted requests, plus the device bytes of
    // every scene look and piece (geometry) they render, each charged once
    // while at least one admitted request uses it, plus the context's wave
    // state for the largest material-key count admitted so far
    size_t budget = 0, films = 0, scenes = 0;
    int max_keys = 2;
    std::map<const ag_scene*, int> scene_refs;
    std::map<const Geometry*, int> geo_refs;
    uint64_t step_samples = 1u << 24;
    ag_path_settings ps{};
    std::chrono::steady_clock::time_point t0;
    uint64_t steps = 0;
    int preview_run = 0;  // consecutive steps that carried only first previews
    struct Req {
        ag_request r;
        ag_request_status st;
        ag_film* film = nullptr;
        uint32_t next_chunk = 0;
        size_t cost = 0;         // film bytes
        uint64_t last_step = 0;  // step that last rendered a chunk of it
        std::string error;       // why it failed (AG_REQ_FAILED)
    };
    std::map<uint64_t, Req> reqs;
    std::vector<uint64_t> queue;   // FIFO of queued ids
    std::vector<uint64_t> active;  // admission order
    uint64_t next_id = 1;
    double now() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

// At most this many steps in a row carry only requests awaiting their first
// preview; the next one renders the requests in progress too (least recently
// rendered first), so a stream of arrivals cannot stall them.
constexpr int SERVER_PREVIEW_RUN_MAX = 4;

// Device bytes of a context's wave state (ensure_wave, ensure_spill) for
// n_keys material keys: path state, ray / hit / shadow queues, buckets,
// counters, segments and the traversal spill stacks.
size_t wave_state_bytes(const ag_ctx* c, int n_keys, uint32_t max_depth) {
    n_keys = wave_keys(c, n_keys);
    const size_t cap = c->tuning.paths_per_wave;
    const size_t per_path = PATH_STATE_BYTES;
    const size_t qcap = cap + queue_slack(c, (uint32_t)cap);
    size_t b = cap * (per_path + (size_t)n_keys * 4) + qcap * (4 * sizeof(float4) + sizeof(float4) + 4 * sizeof(float4));
    b += (counter_words(max_depth) + (size_t)(max_depth + 2) * CURSORS_PER_DEPTH) * sizeof(uint32_t);
    b += MAX_SEGMENTS * sizeof(DSeg);
    const size_t blocks = (size_t)c->n_cu * std::max(c->tuning.blocks_per_cu, c->tuning.batch_blocks_per_cu);
    b += blocks * TRAV_BLOCK * STACK_SPILL * sizeof(uint2);
    return b;
}

size_t film_bytes(uint32_t w, uint32_t h) { return (size_t)w * h * 9 * sizeof(float) * 2; }  // sums + pass accumulators

// Device bytes a request adds to the server's account when admitted.
size_t admit_bytes(const ag_server* s, const ag_server::Req& q) {
    const ag_scene* sc = q.r.scene;
    size_t b = q.cost;
    if (!s->scene_refs.count(sc)) b += sc->bytes;
    if (!s->geo_refs.count(sc->geo.get())) b += sc->geo->bytes();
    return b;
}

size_t server_used(const ag_server* s, int n_keys) {
    return s->films + s->scenes + wave_state_bytes(s->ctx, n_keys, s->ps.max_depth);
}

void server_charge(ag_server* s, ag_server::Req& q) {
    const ag_scene* sc = q.r.scene;
    if (s->scene_refs[sc]++ == 0) s->scenes += sc->bytes;
    if (s->geo_refs[sc->geo.get()]++ == 0) s->scenes += sc->geo->bytes();
    s->films += q.cost;
    s->max_keys = std::max(s->max_keys, sc->n_keys);
}

// Frees the request's film and returns its bytes to the budget.
void server_uncharge(ag_server* s, ag_server::Req& q) {
    if (!q.film) return;
    ag_film_destroy(q.film);
    q.film = nullptr;
    s->films -= q.cost;
    const ag_scene* sc = q.r.scene;
    if (--s->scene_refs[sc] == 0) {
        s->scene_refs.erase(sc);
        s->scenes -= sc->bytes;
    }
    if (--s->geo_refs[sc->geo.get()] == 0) {
        s->geo_refs.erase(sc->geo.get());
        s->scenes -= sc->geo->bytes();
    }
}

void server_fail(ag_server* s, uint64_t k, const std::string& why) {
    auto& q = s->reqs[k];
    server_uncharge(s, q);
    q.st.state = AG_REQ_FAILED;
    q.st.t_done = s->now();
    q.error = why;
    s->active.erase(std::remove(s->active.begin(), s->active.end(), k), s->active.end());
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), k), s->queue.end());
}

// One shared wavefront schedule over the next chunk of each request in `set`.
int server_render(ag_server* s, const std::vector<uint64_t>& set, const ViewJob* extra = nullptr) {
    ag_ctx* ctx = s->ctx;
    std::vector<ag_material> mats;
    std::vector<ViewJob> jobs;
    if (extra) jobs.push_back(*extra);  // an attached live view's pass (its own look and path settings)
    for (uint64_t k : set) {
        auto& q = s->reqs[k];
        ViewJob j;
        j.film = q.film;
        j.count = q.next_chunk;
        j.first = q.film->samples;
        int m0 = -1;
        if (q.r.has_material0) {
            m0 = (int)mats.size();
            mats.push_back(q.r.material0);
        }
        make_dview(j.dv, q.r.scene, &q.r.camera, q.r.turntable_deg, m0, m0 >= 0 ? (int)mats[m0].kind : -1, q.film, q.r.seed,
                   &s->ps);
        j.scene = q.r.scene;
        jobs.push_back(j);
    }
    if (jobs.empty()) return AG_OK;
    int rc = upload_view_materials(ctx, mats.data(), (int32_t)mats.size());
    if (rc) return rc;
    return run_views(ctx, jobs[0].scene, jobs, ctx->vmat_buf.as<DMat>());
}

// A failed pass leaves film sums and sample counts as they were (the commit
// into the sums is its last kernel and the count moves only on success), but
// its accumulators may hold part of the pass: clear them before a retry.
void server_clear_acc(ag_server* s, const std::vector<uint64_t>& set) {
    for (uint64_t k : set) {
        ag_film* f = s->reqs[k].film;
        launch_zero_u32(s->ctx->stream, f->acc.as<uint32_t>(), (uint32_t)((size_t)f->w * f->h * 9));
    }
    (void)hipStreamSynchronize(s->ctx->stream);
}

void server_progress(ag_server* s, const std::vector<uint64_t>& set) {
    double t = s->now();
    for (uint64_t k : set) {
        auto& q = s->reqs[k];
        q.st.samples = q.film->samples;
        q.last_step = s->steps;
        if (q.st.t_first_preview < 0.0) q.st.t_first_preview = t;
        uint32_t rest = q.r.spp - q.st.samples;
        if (rest == 0) {
            q.st.state = AG_REQ_DONE;
            q.st.t_done = t;
            continue;
        }
        uint64_t px = (uint64_t)q.r.width * q.r.height;
        uint32_t cap = (uint32_t)std::max<uint64_t>(1, s->step_samples / px);
        q.next_chunk = std::min(rest, std::max(q.next_chunk, std::min(cap, q.next_chunk * 2)));
    }
}

}  // namespace

extern "C" int ag_server_create(ag_ctx* ctx, size_t budget, uint64_t step_samples, const ag_path_settings* ps,
                                ag_server** out) {
    if (!ctx || !ps || !out || step_samples == 0) return fail(AG_E_INVALID, "bad arguments");
    if (ps->max_depth > 64) return fail(AG_E_INVALID, "max_depth above 64");
    auto* s = new ag_server();
    s->ctx = ctx;
    s->budget = budget;
    s->step_samples = step_samples;
    s->ps = *ps;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return AG_OK;
}

extern "C" void ag_server_destroy(ag_server* s) {
    if (!s) return;
    for (auto& kv : s->reqs)
        if (kv.second.film) ag_film_destroy(kv.second.film);
    delete s;
}

extern "C" int ag_server_submit(ag_server* s, const ag_request* r, uint64_t* id) {
    if (!s || !r || !id) return fail(AG_E_INVALID, "null argument");
    if (!r->scene || r->scene->ctx != s->ctx) return fail(AG_E_INVALID, "request scene missing or from another context");
    if (r->width == 0 || r->height == 0 || r->spp == 0) return fail(AG_E_INVALID, "empty request");
    if ((uint64_t)r->width * r->height > 0xFFFFFFFFull) return fail(AG_E_INVALID, "more than 2^32 - 1 pixels");
    if (r->has_material0 && (r->material0.kind < AG_MAT_METAL || r->material0.kind > AG_MAT_OPAL ||
                             r->material0.kind == AG_MAT_GEM))
        return fail(AG_E_INVALID, "bad material override");
    ag_server::Req q;
    q.r = *r;
    if (q.r.preview_spp == 0) q.r.preview_spp = 2;
    q.st.state = AG_REQ_QUEUED;
    q.st.samples = 0;
    q.st.t_submit = s->now();
    q.st.t_admit = q.st.t_first_preview = q.st.t_done = -1.0;
    q.cost = film_bytes(r->width, r->height);
    q.next_chunk = std::min(q.r.preview_spp, q.r.spp);
    uint64_t k = s->next_id++;
    // a request that could not be admitted even into an idle server
    const size_t alone = q.cost + r->scene->bytes + r->scene->geo->bytes() +
                         wave_state_bytes(s->ctx, std::max(2, r->scene->n_keys), s->ps.max_depth);
    if (alone > s->budget) q.st.state = AG_REQ_REJECTED;
    s->reqs[k] = q;
    if (q.st.state == AG_REQ_QUEUED) s->queue.push_back(k);
    *id = k;
    return AG_OK;
}

// One scheduling round; `extra` (may be null) is an attached live view's pass,
// rendered in the same schedule ahead of the finals, which share what is left
// of the step budget.  *extra_rc: that pass's result; *extra_ms: its share of
// the step's time (by samples).
static int server_step(ag_server* s, const ViewJob* extra, uint32_t* n_active, uint32_t* n_queued, int* extra_rc,
                       double* extra_ms) {
    if (!s) return fail(AG_E_INVALID, "null server");
    ag_ctx* ctx = s->ctx;
    // admit FIFO while the request's film and any scene it brings fit
    while (!s->queue.empty()) {
        const uint64_t k = s->queue.front();
        auto& q = s->reqs[k];
        const int nk = std::max(s->max_keys, q.r.scene->n_keys);
        if (server_used(s, nk) + admit_bytes(s, q) > s->budget) break;
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(ctx->device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
            q.cost > free_b) {
            server_fail(s, k, "film of " + std::to_string(q.cost) + " bytes exceeds the device's free memory (" +
                                  std::to_string(free_b) + " bytes)");
            continue;
        }
        if (ag_film_create(ctx, q.r.width, q.r.height, &q.film)) {
            q.film = nullptr;
            server_fail(s, k, std::string("film creation failed: ") + ag_last_error());
            continue;
        }
        server_charge(s, q);
        q.st.state = AG_REQ_ACTIVE;
        q.st.t_admit = s->now();
        s->active.push_back(k);
        s->queue.erase(s->queue.begin());
    }
    const uint64_t extra_samples =
        extra ? (uint64_t)extra->count * extra->film->w * extra->film->h : 0;
    if (extra_rc) *extra_rc = AG_OK;
    if (extra_ms) *extra_ms = 0.0;
    if (!s->active.empty() || extra) {
        s->steps++;
        // choose: no preview yet first, then least work left, within the step budget
        std::vector<uint64_t> order = s->active;
        auto left = [&](uint64_t k) {
            const auto& q = s->reqs[k];
            return (uint64_t)(q.r.spp - q.st.samples) * q.r.width * q.r.height;
        };
        auto waiting = [&](uint64_t k) { return s->reqs[k].st.t_first_preview < 0.0; };
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
            bool pa = waiting(a), pb = waiting(b);
            if (pa != pb) return pa;
            return left(a) < left(b);
        });
        // A step with requests still waiting for their first preview renders
        // only those, so the previews are published after a short step rather
        // than a full one.  The others sit the step out; each request's chunk
        // sequence, and so its image, does not depend on the steps it skips.
        // After SERVER_PREVIEW_RUN_MAX such steps in a row the requests in
        // progress go first, least recently rendered first.
        bool preview_step = !order.empty() && waiting(order.front());
        if (preview_step && s->preview_run >= SERVER_PREVIEW_RUN_MAX) {
            std::vector<uint64_t> prog;
            for (uint64_t k : order)
                if (!waiting(k)) prog.push_back(k);
            if (!prog.empty()) {
                std::stable_sort(prog.begin(), prog.end(),
                                 [&](uint64_t a, uint64_t b) { return s->reqs[a].last_step < s->reqs[b].last_step; });
                for (uint64_t k : order)
                    if (waiting(k)) prog.push_back(k);
                order = prog;
                preview_step = false;
            }
        }
        s->preview_run = preview_step ? s->preview_run + 1 : 0;
        std::vector<uint64_t> chosen;
        uint64_t room = s->step_samples > extra_samples ? s->step_samples - extra_samples : 0;
        for (uint64_t k : order) {
            const auto& q = s->reqs[k];
            if (preview_step && !waiting(k)) break;
            uint64_t cost = (uint64_t)q.next_chunk * q.r.width * q.r.height;
            // a live pass may fill the step; requests awaiting a first preview still join it
            if ((!chosen.empty() || (extra && !waiting(k))) && cost > room) continue;
            chosen.push_back(k);
            room = cost >= room ? 0 : room - cost;
            if (room == 0) break;
        }
        uint64_t total = extra_samples;  // samples of this step, for the live pass's share of its time
        for (uint64_t k : chosen) total += (uint64_t)s->reqs[k].next_chunk * s->reqs[k].r.width * s->reqs[k].r.height;
        const auto t_step = std::chrono::steady_clock::now();
        if (server_render(s, chosen, extra) == AG_OK) {
            server_progress(s, chosen);
            if (extra_ms) {
                const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_step).count();
                *extra_ms = total ? ms * (double)extra_samples / (double)total : ms;
            }
        } else {
            // one request must not fail the others: retry each alone, fail
            // those that fail alone (with the pass's error), keep the rest
            const std::string first_error = ag_last_error();
            server_clear_acc(s, chosen);
            if (extra) {
                launch_zero_u32(s->ctx->stream, extra->film->acc.as<uint32_t>(),
                                (uint32_t)((size_t)extra->film->w * extra->film->h * 9));
                const auto t1 = std::chrono::steady_clock::now();
                int rc = server_render(s, {}, extra);
                if (extra_rc) *extra_rc = rc;
                if (extra_ms) *extra_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
            }
            if (chosen.size() == 1 && !extra) {
                // the step rendered this request alone: the error is its own
                server_fail(s, chosen[0], first_error);
            } else if (!chosen.empty()) {
                for (uint64_t k : chosen) {
                    std::vector<uint64_t> one{k};
                    if (server_render(s, one) == AG_OK) {
                        server_progress(s, one);
                    } else {
                        std::string why = ag_last_error();
                        server_clear_acc(s, one);
                        server_fail(s, k, why);
                    }
                }
            }
        }
        std::vector<uint64_t> still;
        for (uint64_t k : s->active)
            if (s->reqs[k].st.state == AG_REQ_ACTIVE) still.push_back(k);
        s->active = still;
    }
    if (n_active) *n_active = (uint32_t)s->active.size();
    if (n_queued) *n_queued = (uint32_t)s->queue.size();
    return AG_OK;
}

extern "C" int ag_debug_fail_seed(ag_ctx* ctx, int64_t seed) {
    if (!ctx) return fail(AG_E_INVALID, "null context");
    if (seed > 0xFFFFFFFFll) return fail(AG_E_INVALID, "seed above 2^32 - 1");
    ctx->fail_seed_on = seed >= 0;
    ctx->fail_seed_h = seed >= 0 ? hash32_host((uint32_t)seed ^ 0xa511e9b3u) : 0u;
    return AG_OK;
}

extern "C" int ag_server_step(ag_server* s, uint32_t* n_active, uint32_t* n_queued) {
    return server_step(s, nullptr, n_active, n_queued, nullptr, nullptr);
}

extern "C" int ag_server_status(const ag_server* s, uint64_t id, ag_request_status* out) {
    if (!s || !out) return fail(AG_E_INVALID, "null argument");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(AG_E_NOTFOUND, "no such request");
    *out = it->second.st;
    return AG_OK;
}

extern "C" const char* ag_server_error(const ag_server* s, uint64_t id) {
    if (!s) return "";
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? "" : it->second.error.c_str();
}

extern "C" int ag_server_budget(const ag_server* s, size_t* used, size_t* budget) {
    if (!s) return fail(AG_E_INVALID, "null server");
    if (used) *used = server_used(s, s->max_keys);
    if (budget) *budget = s->budget;
    return AG_OK;
}

extern "C" int ag_server_read(const ag_server* s, uint64_t id, int32_t which, float* out) {
    if (!s || !out) return fail(AG_E_INVALID, "null argument");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(AG_E_NOTFOUND, "no such request");
    if (!it->second.film) return fail(AG_E_STATE, "request has no film (not admitted, failed or released)");
    return ag_film_read(it->second.film, which, out);
}

extern "C" int ag_server_release(ag_server* s, uint64_t id) {
    if (!s) return fail(AG_E_INVALID, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail(AG_E_NOTFOUND, "no such request");
    if (it->second.st.state == AG_REQ_ACTIVE) return fail(AG_E_STATE, "request still rendering");
    server_uncharge(s, it->second);
    s->queue.erase(std::remove(s->queue.begin(), s->queue.end(), id), s->queue.end());
    s->reqs.erase(it);
    return AG_OK;
}

// ------------------------------------------------------------------ ray batches
static int ray_batch(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, ag_hits_soa* hits,
                     uint8_t* occ, bool device_ptrs) {
    if (!ctx || !s || !rays) return fail(AG_E_INVALID, "null argument");
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_spill(ctx);
    if (rc) return rc;
    if ((rc = ctx->cursor_buf.ensure(CURSOR_SET * sizeof(uint32_t)))) return rc;
    hipStream_t st = ctx->stream;
    RaysSoA r{};
    DevBuf in, outb;
    if (device_ptrs) {
        r = RaysSoA{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    } else {
        size_t f = (size_t)n * 4;
        if ((rc = in.ensure(f * 7))) return rc;
        float* b = in.as<float>();
        const float* src[7] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
        for (int k = 0; k < 6; k++) {
            if (!src[k]) return fail(AG_E_INVALID, "missing ray component");
            HIPCHK(hipMemcpyAsync(b + (size_t)k * n, src[k], f, hipMemcpyHostToDevice, st));
        }
        if (rays->tmax) HIPCHK(hipMemcpyAsync(b + 6 * (size_t)n, rays->tmax, f, hipMemcpyHostToDevice, st));
        r = RaysSoA{b, b + n, b + 2 * (size_t)n, b + 3 * (size_t)n, b + 4 * (size_t)n, b + 5 * (size_t)n,
                    rays->tmax ? b + 6 * (size_t)n : nullptr};
    }
    const float4* nodes = s->geo->nodes.as<float4>();
    const float4* tris = s->geo->tris.as<float4>();
    int grid = ctx->grid_batch;
    uint32_t refill = ctx->tuning.batch_refill_min;
    if (hits) {
        HitsSoA h{};
        if (device_ptrs) {
            h = HitsSoA{hits->t, hits->u, hits->v, hits->prim};
        } else {
            if ((rc = outb.ensure((size_t)n * 16))) return rc;
            float* b = outb.as<float>();
            h = HitsSoA{b, b + n, b + 2 * (size_t)n, reinterpret_cast<uint32_t*>(b + 3 * (size_t)n)};
        }
        launch_zero_u32(st, ctx->cursor_buf.as<uint32_t>(), CURSOR_SET);
        HIPCHK(hipEventRecord(ctx->ev[2], st));
        launch_rays_intersect(st, grid, nodes, tris, r, h, n, ctx->cursor_buf.as<uint32_t>(), refill,
                              ctx->tuning.spread_min_lanes, ctx->spill_buf.as<uint2>());
        HIPCHK(hipEventRecord(ctx->ev[3], st));
        HIPCHK(hipGetLastError());
        if (!device_ptrs) {
            HIPCHK(hipMemcpyAsync(hits->t, h.t, (size_t)n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(hits->u, h.u, (size_t)n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(hits->v, h.v, (size_t)n * 4, hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(hits->prim, h.prim, (size_t)n * 4, hipMemcpyDeviceToHost, st));
        }
    } else {
        uint8_t* o = occ;
        if (!device_ptrs) {
            if ((rc = outb.ensure(n))) return rc;
            o = outb.as<uint8_t>();
        }
        launch_zero_u32(st, ctx->cursor_buf.as<uint32_t>(), CURSOR_SET);
        HIPCHK(hipEventRecord(ctx->ev[2], st));
        launch_rays_occluded(st, grid, nodes, tris, r, o, n, ctx->cursor_buf.as<uint32_t>(), refill,
                             ctx->tuning.spread_min_lanes, ctx->spill_buf.as<uint2>());
        HIPCHK(hipEventRecord(ctx->ev[3], st));
        HIPCHK(hipGetLastError());
        if (!device_ptrs) HIPCHK(hipMemcpyAsync(occ, o, n, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->last_query_ms = ms;
    return AG_OK;
}

// The AO rays of cavity.rs:95-128 on the context stream: blocked-ray counts
// per vertex into `counts` (n_vertices words).  Returns after the launch.
static int ao_counts(ag_ctx* ctx, const Geometry& g, const float* nrm, uint32_t n, float distance, DevBuf& counts) {
    const uint32_t nv = g.n_vertices;
    if ((uint64_t)nv * n > 0xFFFFFFFFull) return fail(AG_E_INVALID, "n_vertices * samples exceeds 2^32 - 1 rays");
    int rc = ensure_spill(ctx);
    if (rc) return rc;
    if ((rc = ctx->cursor_buf.ensure(CURSOR_SET * sizeof(uint32_t)))) return rc;
    // cavity.rs:87-90 and :97: eps from the diagonal of the vertex bounds, in f32
    float dx = g.vhi[0] - g.vlo[0], dy = g.vhi[1] - g.vlo[1], dz = g.vhi[2] - g.vlo[2];
    const float diag = std::sqrt((dx * dx + dy * dy) + dz * dz);
    const float eps = std::fmax(2e-4f * diag, 1e-3f);
    if ((rc = counts.ensure((size_t)nv * sizeof(uint32_t)))) return rc;
    hipStream_t st = ctx->stream;
    launch_zero_u32(st, counts.as<uint32_t>(), nv);
    launch_zero_u32(st, ctx->cursor_buf.as<uint32_t>(), CURSOR_SET);
    launch_ao(st, ctx->grid_batch, g.nodes.as<float4>(), g.tris.as<float4>(), g.pos.as<float>(), nrm, n,
              distance, eps, counts.as<uint32_t>(), nv * n, ctx->cursor_buf.as<uint32_t>(), ctx->tuning.batch_refill_min,
              ctx->tuning.spread_min_lanes, ctx->spill_buf.as<uint2>());
    HIPCHK(hipGetLastError());
    return AG_OK;
}

// cavity.rs:82-87 (vertex_normals): the piece's normals, else
// smooth_vertex_normals computed into `tmp`
static int vertex_normals(ag_ctx* ctx, const Geometry& g, DevBuf& tmp, const float** out) {
    if (g.has_nrm) {
        *out = g.nrm.as<float>();
        return AG_OK;
    }
    int rc = tmp.ensure(12 * (size_t)std::max(g.n_vertices, 1u));
    if (rc) return rc;
    if ((rc = smooth_normals_device(ctx, g.pos.as<float>(), g.idx.as<uint32_t>(), g.n_vertices, g.n_triangles,
                                    tmp.as<float>())))
        return rc;
    *out = tmp.as<float>();
    return AG_OK;
}

extern "C" int ag_ambient_occlusion(ag_ctx* ctx, const ag_scene* s, uint32_t samples, float distance, float* out) {
    if (!ctx || !s || !out) return fail(AG_E_INVALID, "null argument");
    const Geometry& g = *s->geo;
    const uint32_t nv = g.n_vertices;
    if (nv == 0) return AG_OK;
    if (g.n_triangles == 0) {  // cavity.rs:54-56 returns zeros before tracing
        std::fill(out, out + nv, 0.0f);
        return AG_OK;
    }
    const uint32_t n = std::max(samples, 1u);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf counts, smooth;
    const float* nrm = nullptr;
    int rc = vertex_normals(ctx, g, smooth, &nrm);
    if (rc) return rc;
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    if ((rc = ao_counts(ctx, g, nrm, n, distance, counts))) return rc;
    HIPCHK(hipEventRecord(ctx->ev[3], st));
    std::vector<uint32_t> h(nv);
    HIPCHK(hipMemcpyAsync(h.data(), counts.p, (size_t)nv * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->last_query_ms = ms;
    for (uint32_t v = 0; v < nv; v++) out[v] = (float)h[v] / (float)n;  // cavity.rs:126
    return AG_OK;
}

// ------------------------------------------------------------------ patina mask
extern "C" int ag_patina_default(ag_patina_settings* out) {
    if (!out) return fail(AG_E_INVALID, "null argument");
    *out = ag_patina_settings{48u, 1.2f, 0.4f, 0.35f, 3u};  // cavity.rs:38-42
    return AG_OK;
}

extern "C" int ag_patina_for_scene(const ag_scene* s, ag_patina_settings* out) {
    if (!s || !out) return fail(AG_E_INVALID, "null argument");
    ag_patina_default(out);
    // shot.rs:133-141: sized by the piece's vertex bounds, in f32
    const Geometry& g = *s->geo;
    const float hx = (g.vhi[0] - g.vlo[0]) * 0.5f, hy = (g.vhi[1] - g.vlo[1]) * 0.5f, hz = (g.vhi[2] - g.vlo[2]) * 0.5f;
    const float len = std::sqrt(hx * hx + hy * hy + hz * hz);
    const float radius = len > 1e-3f ? len : 1e-3f;  // f32::max(len, 1e-3): NaN gives 1e-3
    auto clampf = [](float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); };
    out->ao_distance = clampf(radius * 0.09f, 0.4f, 2.5f);
    out->curvature_radius = clampf(radius * 0.03f, 0.15f, 1.0f);
    return AG_OK;
}

static int patina_inputs(const ag_scene* s, PatinaInputs* in) {
    const Geometry& g = *s->geo;
    *in = PatinaInputs{};
    in->pos = g.pos.as<float>();
    in->nrm = g.has_nrm ? g.nrm.as<float>() : nullptr;
    in->tris = g.idx.as<uint32_t>();
    in->nv = g.n_vertices;
    in->nt = g.n_triangles;
    for (int k = 0; k < 3; k++) in->lo[k] = g.vlo[k];
    in->lambda = 0.5f;
    return AG_OK;
}

extern "C" int ag_patina_mask(ag_ctx* ctx, ag_scene* s, const ag_patina_settings* settings, int32_t store, float* mask,
                              float* ao, float* concavity) {
    if (!ctx || !s || !settings) return fail(AG_E_INVALID, "null argument");
    if (s->ctx != ctx) return fail(AG_E_INVALID, "scene belongs to another context");
    const ag_patina_settings& P = *settings;
    if (!std::isfinite(P.ao_distance) || !std::isfinite(P.curvature_radius) || !std::isfinite(P.concavity_weight))
        return fail(AG_E_INVALID, "patina settings must be finite");
    Geometry& g = *s->geo;
    const uint32_t nv = g.n_vertices;
    if (nv == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf out, counts;
    int rc = out.ensure(3 * 4 * (size_t)nv);
    if (rc) return rc;
    float* d_mask = out.as<float>();
    float* d_ao = d_mask + nv;
    float* d_cav = d_ao + nv;
    const bool with_ao = P.ao_samples > 0 && P.ao_distance > 0.0f;
    HIPCHK(hipEventRecord(ctx->ev[2], st));
    if (g.n_triangles == 0) {  // cavity.rs:54-56
        HIPCHK(hipMemsetAsync(out.p, 0, 3 * 4 * (size_t)nv, st));
    } else {
        DevBuf smooth;
        const float* nrm = nullptr;
        if ((rc = vertex_normals(ctx, g, smooth, &nrm))) return rc;
        PatinaInputs in;
        patina_inputs(s, &in);
        in.nrm = nrm;
        if (with_ao) {
            if ((rc = ao_counts(ctx, g, nrm, P.ao_samples, P.ao_distance, counts))) return rc;
            in.ao_counts = counts.as<uint32_t>();
            in.ao_samples = P.ao_samples;
        }
        in.clamp01 = 1;
        if ((rc = patina_device(ctx, in, P, d_mask, d_ao, d_cav))) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev[3], st));
    if (store) {
        if ((rc = g.cavity.ensure(4 * (size_t)nv))) return rc;
        HIPCHK(hipMemcpyAsync(g.cavity.p, d_mask, 4 * (size_t)nv, hipMemcpyDeviceToDevice, st));
        g.has_cavity = true;
        s->hs.cavity = g.cavity.as<float>();
        HIPCHK(hipMemcpyAsync(s->d_scene.p, &s->hs, sizeof(DScene), hipMemcpyHostToDevice, st));
    }
    float* host[3] = {mask, ao, concavity};
    for (int k = 0; k < 3; k++)
        if (host[k]) HIPCHK(hipMemcpyAsync(host[k], d_mask + k * (size_t)nv, 4 * (size_t)nv, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]) == hipSuccess) ctx->last_query_ms = ms;
    return AG_OK;
}

extern "C" int ag_smooth_values(ag_ctx* ctx, const ag_scene* s, float* values, uint32_t iterations, float lambda) {
    if (!ctx || !s || !values) return fail(AG_E_INVALID, "null argument");
    if (s->ctx != ctx) return fail(AG_E_INVALID, "scene belongs to another context");
    const uint32_t nv = s->geo->n_vertices;
    if (nv == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevBuf buf;
    int rc = buf.ensure(2 * 4 * (size_t)nv);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(buf.p, values, 4 * (size_t)nv, hipMemcpyHostToDevice, st));
    PatinaInputs in;
    patina_inputs(s, &in);
    in.values = buf.as<float>();
    in.lambda = lambda;
    ag_patina_settings P{};
    P.smooth_iterations = iterations;
    if ((rc = patina_device(ctx, in, P, buf.as<float>() + nv, nullptr, nullptr))) return rc;
    HIPCHK(hipMemcpyAsync(values, buf.as<float>() + nv, 4 * (size_t)nv, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return AG_OK;
}

// ------------------------------------------------------------------ preview development
namespace {
constexpr size_t POST_TABLE_FLOATS = 1 << 16;  // grade tables: at most 3 AgX curves of ~9.3 K + 4097
struct PostWork {
    float4* img[4];    // colour, albedo, normal, result
    AtrousWork W;
    float* stage[3];   // (r, g, b) staging for host images
    uint8_t* u8;
    float* tables;
};
int post_work(ag_ctx* c, size_t n, PostWork* P) {
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t f4 = up(n * 16), f1 = up(n * 4), f3 = up(n * 12), b3 = up(n * 3);
    int rc = c->post_buf.ensure(9 * f4 + 6 * f1 + 3 * f3 + b3 + POST_TABLE_FLOATS * 4);
    if (rc) return rc;
    char* p = c->post_buf.as<char>();
    auto take = [&](size_t b) { char* q = p; p += b; return q; };
    for (auto& im : P->img) im = reinterpret_cast<float4*>(take(f4));
    P->W.demod = reinterpret_cast<float4*>(take(f4));
    P->W.ga = reinterpret_cast<float4*>(take(f4));
    P->W.gn = reinterpret_cast<float4*>(take(f4));
    P->W.sig[0] = reinterpret_cast<float4*>(take(f4));
    P->W.sig[1] = reinterpret_cast<float4*>(take(f4));
    P->W.l[0] = reinterpret_cast<float*>(take(f1));
    P->W.l[1] = reinterpret_cast<float*>(take(f1));
    P->W.var[0] = reinterpret_cast<float*>(take(f1));
    P->W.var[1] = reinterpret_cast<float*>(take(f1));
    P->W.t0 = reinterpret_cast<float*>(take(f1));
    P->W.t1 = reinterpret_cast<float*>(take(f1));
    for (auto& st : P->stage) st = reinterpret_cast<float*>(take(f3));
    P->u8 = reinterpret_cast<uint8_t*>(take(b3));
    P->tables = reinterpret_cast<float*>(take(POST_TABLE_FLOATS * 4));
    return AG_OK;
}
// Tables of `g` onto the device; DevGrade ready for launch_develop.
int post_grade(ag_ctx* c, const ag_grade* grade, const PostWork& P, DevGrade* g) {
    std::vector<float> blob;
    int nc = 0;
    size_t clen = 0, elen = 0;
    int rc = build_grade(grade->exposure_ev, grade->tonemap, grade->saturation, &blob, g, &nc, &clen, &elen);
    if (rc) return rc;
    if (blob.size() > POST_TABLE_FLOATS) return fail(AG_E_INVALID, "grade tables larger than the workspace");
    HIPCHK(hipMemcpyAsync(P.tables, blob.data(), blob.size() * 4, hipMemcpyHostToDevice, c->stream));
    for (int k = 0; k < 3; k++) g->curve[k] = nc ? P.tables + (size_t)std::min(k, nc - 1) * clen : P.tables;
    g->encode = P.tables + (size_t)nc * clen;
    return AG_OK;
}
int post_upload(ag_ctx* c, const float* rgb, size_t n, float* stage, float4* img) {
    HIPCHK(hipMemcpyAsync(stage, rgb, n * 12, hipMemcpyHostToDevice, c->stream));
    launch_pack4(c->stream, stage, n, img);
    return AG_OK;
}
int post_film(ag_ctx* c, const ag_film* f, const PostWork& P) {
    launch_film_images(c->stream, f->sums.as<float>(), (size_t)f->w * f->h, 1.0f / (float)std::max(f->samples, 1u),
                       P.img[0], P.img[1], P.img[2]);
    return AG_OK;
}
int post_finish_u8(ag_ctx* c, const PostWork& P, size_t n, uint8_t* out) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, P.u8, n * 3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AG_OK;
}
int post_finish_f(ag_ctx* c, const PostWork& P, size_t n, float* out) {
    launch_unpack4(c->stream, P.img[3], n, P.stage[0]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, P.stage[0], n * 12, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return AG_OK;
}
}  // namespace

extern "C" int ag_develop_image(ag_ctx* ctx, const float* rgb, uint32_t w, uint32_t h, const ag_grade* grade,
                                uint8_t* out) {
    if (!ctx || !grade || ((!rgb || !out) && (size_t)w * h)) return fail(AG_E_INVALID, "null argument");
    const size_t n = (size_t)w * h;
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    PostWork P;
    DevGrade g;
    int rc;
    if ((rc = post_work(ctx, n, &P)) || (rc = post_grade(ctx, grade, P, &g)) ||
        (rc = post_upload(ctx, rgb, n, P.stage[0], P.img[0])))
        return rc;
    launch_develop(ctx->stream, P.img[0], (int)w, (int)h, g, P.u8);
    return post_finish_u8(ctx, P, n, out);
}

extern "C" int ag_denoise_image(ag_ctx* ctx, const float* color, const float* albedo, const float* normal, uint32_t w,
                                uint32_t h, float* out) {
    if (!ctx || ((!color || !out) && (size_t)w * h)) return fail(AG_E_INVALID, "null argument");
    const size_t n = (size_t)w * h;
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    PostWork P;
    int rc;
    if ((rc = post_work(ctx, n, &P)) || (rc = post_upload(ctx, color, n, P.stage[0], P.img[0])) ||
        (albedo && (rc = post_upload(ctx, albedo, n, P.stage[1], P.img[1]))) ||
        (normal && (rc = post_upload(ctx, normal, n, P.stage[2], P.img[2]))))
        return rc;
    launch_atrous(ctx->stream, (int)w, (int)h, P.img[0], albedo ? P.img[1] : nullptr, normal ? P.img[2] : nullptr, P.W,
                  P.img[3]);
    return post_finish_f(ctx, P, n, out);
}

extern "C" int ag_film_denoise(ag_ctx* ctx, const ag_film* film, float* out) {
    if (!ctx || !film || !out) return fail(AG_E_INVALID, "null argument");
    if (film->ctx != ctx) return fail(AG_E_INVALID, "film belongs to another context");
    const size_t n = (size_t)film->w * film->h;
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    PostWork P;
    int rc;
    if ((rc = post_work(ctx, n, &P)) || (rc = post_film(ctx, film, P))) return rc;
    launch_atrous(ctx->stream, (int)film->w, (int)film->h, P.img[0], P.img[1], P.img[2], P.W, P.img[3]);
    return post_finish_f(ctx, P, n, out);
}

extern "C" int ag_film_develop(ag_ctx* ctx, const ag_film* film, const ag_grade* grade, int32_t denoise,
                               uint8_t* out) {
    if (!ctx || !film || !grade || !out) return fail(AG_E_INVALID, "null argument");
    if (film->ctx != ctx) return fail(AG_E_INVALID, "film belongs to another context");
    const size_t n = (size_t)film->w * film->h;
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    PostWork P;
    DevGrade g;
    int rc;
    if ((rc = post_work(ctx, n, &P)) || (rc = post_grade(ctx, grade, P, &g)) || (rc = post_film(ctx, film, P))) return rc;
    if (denoise) launch_atrous(ctx->stream, (int)film->w, (int)film->h, P.img[0], P.img[1], P.img[2], P.W, P.img[3]);
    launch_develop(ctx->stream, denoise ? P.img[3] : P.img[0], (int)film->w, (int)film->h, g, P.u8);
    return post_finish_u8(ctx, P, n, out);
}

// ------------------------------------------------------------------ live progressive loop
// The render worker of the live view (argent serve live.rs:637-830) over the
// device film: a half-resolution first look at 2 spp, then refinement passes
// sized toward 70 ms, frames published at paced intervals, each denoised
// (à-trous) and developed on the device, and a settle test that ends
// refinement once successive denoised frames differ by less than 0.5 %.
// The caller's thread drives it (one ag_live_step per worker iteration) and
// passes its clock, so the pacing is reproducible.
namespace {
constexpr uint32_t LIVE_SEED = 1;                 // live.rs:39
constexpr double MOTION_INTERVAL_MS = 83.0;       // live.rs:41
constexpr double REFINE_INTERVAL_MS = 166.0;      // live.rs:43
constexpr double REFINE_SPP = 32.0;               // live.rs:45
constexpr double REFINE_MAX_MS = 1500.0;          // live.rs:46
constexpr double MOTION_WINDOW_MS = 350.0;        // live.rs:48
constexpr double SETTLE_DELTA = 0.005;            // live.rs:54
constexpr uint32_t SETTLE_MIN_SPP = 64;           // live.rs:55
constexpr double PASS_TARGET_MS = 70.0;           // live.rs:57
constexpr double DENOISE_SHARE = 0.35;            // live.rs:63
constexpr double DENOISE_MAX_MS = 400.0;          // live.rs:65
}  // namespace

struct ag_live {
    ag_ctx* ctx = nullptr;
    const ag_scene* scene = nullptr;
    ag_camera cam{};
    ag_live_params p{};
    ag_film* film = nullptr;
    ag_film* half = nullptr;
    uint64_t gen = 0, dgen = 0;          // requested
    uint64_t cur_gen = 0, cur_dgen = 0;  // rendered
    double changed_at = 0.0;
    bool half_done = false, final_done = true;
    bool has_last_post = false;
    double last_post = 0.0, ms_per_spp = 0.0, gen_started = -1e30;
    uint64_t published_gen = 0;
    bool burst = false;
    bool settled = false;  // refinement stopped early for settled_max
    ag_server* server = nullptr;  // attached: passes ride in this server's steps
    uint32_t settled_max = 0;
    double dn_ms_per_mpx = 0.0;
    // the last published linear image (live.rs:493-502: Linear)
    DevBuf last_img, part;
    bool has_last = false, last_half = false, last_denoised = false;
    uint64_t last_gen = 0;
    uint32_t last_w = 0, last_h = 0, last_spp = 0;
    ~ag_live() {
        if (film) ag_film_destroy(film);
        if (half) ag_film_destroy(half);
    }
};

namespace {
double live_denoise_ms(const ag_live* L, uint32_t pixels) { return L->dn_ms_per_mpx * pixels / 1e6; }
// live.rs:413-419
double live_paced(const ag_live* L, double base, uint32_t pixels) {
    const double ms = live_denoise_ms(L, pixels);
    if (!L->p.denoise || ms > DENOISE_MAX_MS) return base;
    return std::max(base, ms / DENOISE_SHARE);
}
// live.rs:421-424
double live_refine_interval(const ag_live* L, uint32_t spp, uint32_t pixels) {
    const double base = std::min(REFINE_INTERVAL_MS * std::max(std::sqrt(spp / REFINE_SPP), 1.0), REFINE_MAX_MS);
    return live_paced(L, base, pixels);
}
bool live_due(const ag_live* L, double now, double every) { return !L->has_last_post || now - L->last_post >= every; }

// Develops the stored linear image with the current grade into out.
int live_emit(ag_live* L, const PostWork& P, uint8_t* out, ag_live_frame* fr) {
    ag_ctx* c = L->ctx;
    DevGrade g;
    int rc = post_grade(c, &L->p.grade, P, &g);
    if (rc) return rc;
    const size_t n = (size_t)L->last_w * L->last_h;
    launch_develop(c->stream, L->last_img.as<float4>(), (int)L->last_w, (int)L->last_h, g, P.u8);
    if ((rc = post_finish_u8(c, P, n, out))) return rc;
    fr->published = 1;
    fr->generation = L->last_gen;
    fr->width = L->last_w;
    fr->height = L->last_h;
    fr->spp = L->last_spp;
    fr->half = L->last_half;
    fr->denoised = L->last_denoised;
    return AG_OK;
}

// live.rs:429-504: denoise, settle test, develop, publish.  Returns whether
// the frame went out as final (converged or settled) through *final_out.
int live_publish(ag_live* L, ag_film* f, bool is_half, bool converged, double now, uint8_t* out, ag_live_frame* fr,
                 bool* final_out) {
    ag_ctx* c = L->ctx;
    const size_t n = (size_t)f->w * f->h;
    PostWork P;
    int rc;
    if ((rc = post_work(c, n, &P)) || (rc = post_film(c, f, P))) return rc;
    const float4* img = P.img[0];
    bool denoised = false;
    fr->relative_change = -1.0f;
    if (L->p.denoise && (converged || live_denoise_ms(L, (uint32_t)n) <= DENOISE_MAX_MS)) {
        const auto t0 = std::chrono::steady_clock::now();
        launch_atrous(c->stream, (int)f->w, (int)f->h, P.img[0], P.img[1], P.img[2], P.W, P.img[3]);
        HIPCHK(hipStreamSynchronize(c->stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const double r = ms / std::max(n / 1e6, 1e-3);
        L->dn_ms_per_mpx = L->dn_ms_per_mpx == 0.0 ? r : 0.7 * L->dn_ms_per_mpx + 0.3 * r;
        img = P.img[3];
        denoised = true;
        // settled?  compare with the previous full-resolution frame of this set
        if (!converged && !is_half && f->samples >= SETTLE_MIN_SPP && L->has_last && L->last_gen == L->cur_gen &&
            !L->last_half && L->last_denoised && L->last_w == f->w && L->last_h == f->h) {
            int rc2 = L->part.ensure(2 * 512 * sizeof(double) + 2 * sizeof(double));
            if (rc2) return rc2;
            double* sums = L->part.as<double>() + 2 * 512;
            launch_relative_change(c->stream, L->last_img.as<float4>(), img, n, L->part.as<double>(), sums);
            double hs[2];
            HIPCHK(hipMemcpyAsync(hs, sums, sizeof hs, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(hipStreamSynchronize(c->stream));
            const double rel = hs[1] <= 1e-9 ? 0.0 : hs[0] / hs[1];
            fr->relative_change = (float)rel;
            if (rel < SETTLE_DELTA) converged = true;
        }
    }
    if ((rc = L->last_img.ensure(n * sizeof(float4)))) return rc;
    HIPCHK(hipMemcpyAsync(L->last_img.p, img, n * sizeof(float4), hipMemcpyDeviceToDevice, c->stream));
    L->has_last = true;
    L->last_half = is_half;
    L->last_denoised = denoised;
    L->last_gen = L->cur_gen;
    L->last_w = f->w;
    L->last_h = f->h;
    L->last_spp = f->samples;
    if ((rc = live_emit(L, P, out, fr))) return rc;
    fr->converged = converged;
    L->last_post = now;
    L->has_last_post = true;
    *final_out = converged;
    return AG_OK;
}

int live_render(ag_live* L, ag_film* f, uint32_t count, double* ms) {
    if (L->server) {
        // one server step: this pass first, the finals' chunks in the rest of
        // the step budget; the session is charged its samples' share of the time
        if ((uint64_t)f->samples + count > 0xFFFFFFFFull) return fail(AG_E_INVALID, "sample index overflow");
        ViewJob j;
        j.scene = L->scene;
        j.film = f;
        j.count = count;
        j.first = f->samples;
        make_dview(j.dv, L->scene, &L->cam, L->scene->turntable_deg, -1, -1, f, LIVE_SEED, &L->p.ps);
        int pass_rc = AG_OK;
        int rc = server_step(L->server, &j, nullptr, nullptr, &pass_rc, ms);
        return rc ? rc : pass_rc;
    }
    const auto t0 = std::chrono::steady_clock::now();
    int rc = ag_render_pass(L->ctx, L->scene, &L->cam, f, &L->p.ps, count, LIVE_SEED);
    *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return rc;
}
}  // namespace

extern "C" int ag_relative_change(ag_ctx* ctx, const float* a, const float* b, uint32_t w, uint32_t h, double* out) {
    if (!ctx || !out || ((!a || !b) && (size_t)w * h)) return fail(AG_E_INVALID, "null argument");
    const size_t n = (size_t)w * h;
    *out = 0.0;
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    PostWork P;
    int rc;
    if ((rc = post_work(ctx, n, &P)) || (rc = post_upload(ctx, a, n, P.stage[0], P.img[0])) ||
        (rc = post_upload(ctx, b, n, P.stage[1], P.img[1])))
        return rc;
    DevBuf tmp;  // 2 partial sums per block (at most 512 blocks), then the two totals
    if ((rc = tmp.ensure((2 * 512 + 2) * sizeof(double)))) return rc;
    double* part = tmp.as<double>();
    launch_relative_change(ctx->stream, P.img[0], P.img[1], n, part, part + 2 * 512);
    double hs[2];
    HIPCHK(hipMemcpyAsync(hs, part + 2 * 512, sizeof hs, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    *out = hs[1] <= 1e-9 ? 0.0 : hs[0] / hs[1];
    return AG_OK;
}

extern "C" int ag_live_create(ag_ctx* ctx, ag_live** out) {
    if (!ctx || !out) return fail(AG_E_INVALID, "null argument");
    auto* L = new ag_live();
    L->ctx = ctx;
    *out = L;
    return AG_OK;
}

extern "C" void ag_live_destroy(ag_live* L) { delete L; }

extern "C" int ag_live_attach(ag_live* L, ag_server* srv) {
    if (!L) return fail(AG_E_INVALID, "null live session");
    if (srv && srv->ctx != L->ctx) return fail(AG_E_INVALID, "server belongs to another context");
    L->server = srv;
    return AG_OK;
}

extern "C" int ag_live_set(ag_live* L, const ag_scene* scene, const ag_camera* cam, const ag_live_params* p,
                           double now_ms) {
    if (!L || !scene || !cam || !p) return fail(AG_E_INVALID, "null argument");
    if (scene->ctx != L->ctx) return fail(AG_E_INVALID, "scene belongs to another context");
    if (p->width < 2 || p->height < 2 || p->max_spp < 1 || p->max_spp > 65536)
        return fail(AG_E_INVALID, "live view needs width, height >= 2 and max_spp in 1..65536");
    L->scene = scene;
    L->cam = *cam;
    L->p = *p;
    L->gen++;
    L->changed_at = now_ms;
    return AG_OK;
}

extern "C" int ag_live_set_look(ag_live* L, const ag_grade* grade, int32_t denoise, double now_ms) {
    if (!L || !grade) return fail(AG_E_INVALID, "null argument");
    L->p.grade = *grade;
    L->p.denoise = denoise;
    L->dgen++;
    (void)now_ms;
    return AG_OK;
}

// One iteration of live.rs:656-823.
extern "C" int ag_live_step(ag_live* L, double now_ms, ag_live_frame* fr, uint8_t* out) {
    if (!L || !fr || !out) return fail(AG_E_INVALID, "null argument");
    *fr = ag_live_frame{};
    fr->relative_change = -1.0f;
    if (!L->scene) {
        fr->idle = 1;
        return AG_OK;
    }
    HIPCHK(hipSetDevice(L->ctx->device));
    const ag_live_params& p = L->p;
    const bool more = (!L->film || L->film->samples < p.max_spp) && !(L->settled && L->settled_max == p.max_spp);
    if (!(L->gen != L->cur_gen || L->dgen != L->cur_dgen || more || !L->final_done)) {
        fr->idle = 1;
        fr->generation = L->cur_gen;
        fr->spp = L->film ? L->film->samples : 0;
        fr->max_spp = p.max_spp;
        return AG_OK;
    }
    double now = now_ms;
    int rc;
    if (L->gen != L->cur_gen) {
        // a change: restart accumulation (live.rs:683-719)
        const uint32_t w = p.width, h = p.height;
        if (!L->film || L->film->w != w || L->film->h != h) {
            if (L->film) ag_film_destroy(L->film);
            if (L->half) ag_film_destroy(L->half);
            L->film = L->half = nullptr;
            if ((rc = ag_film_create(L->ctx, w, h, &L->film)) ||
                (rc = ag_film_create(L->ctx, std::max(w / 2, 2u), std::max(h / 2, 2u), &L->half)))
                return rc;
        } else {
            if ((rc = ag_film_clear(L->film)) || (rc = ag_film_clear(L->half))) return rc;
        }
        L->cur_gen = L->gen;
        L->cur_dgen = L->dgen;
        L->half_done = false;
        L->final_done = false;
        L->settled = false;
        L->burst = now - L->gen_started < MOTION_WINDOW_MS;
        L->gen_started = now;
    }
    fr->generation = L->cur_gen;
    fr->max_spp = p.max_spp;
    bool fin = false;
    if (L->dgen != L->cur_dgen) {
        // development-only change: re-develop what was published (live.rs:723-740)
        L->cur_dgen = L->dgen;
        if (L->has_last && L->last_gen == L->cur_gen && L->last_denoised == (p.denoise != 0)) {
            PostWork P;
            if ((rc = post_work(L->ctx, (size_t)L->last_w * L->last_h, &P)) || (rc = live_emit(L, P, out, fr))) return rc;
            fr->redeveloped = 1;
            L->last_post = now;
            L->has_last_post = true;
            return AG_OK;
        }
        if (L->film->samples > 0 || L->half->samples > 0) {
            const bool is_half = L->film->samples == 0;
            return live_publish(L, is_half ? L->half : L->film, is_half, L->film->samples >= p.max_spp, now, out, fr,
                                &fin);
        }
    }
    const bool motion = now - L->gen_started < MOTION_WINDOW_MS;
    if (!L->half_done) {
        // first look at a new set: half resolution, 2 spp (live.rs:744-756)
        double ms;
        if ((rc = live_render(L, L->half, 2, &ms))) return rc;
        now += ms;
        fr->pass_spp = 2;
        fr->pass_ms = (float)ms;
        L->half_done = true;
        const double gap = L->burst ? MOTION_INTERVAL_MS : MOTION_INTERVAL_MS / 4;
        if (live_due(L, now, live_paced(L, gap, L->half->w * L->half->h))) {
            if ((rc = live_publish(L, L->half, true, false, now, out, fr, &fin))) return rc;
            L->published_gen = L->cur_gen;
        }
        fr->spp = L->film->samples;
        return AG_OK;
    }
    const uint32_t max = p.max_spp;
    if (L->film->samples < max && !(L->settled && L->settled_max == max)) {
        // refinement pass sized toward PASS_TARGET_MS (live.rs:757-812)
        uint32_t count = 1;
        if (!motion && L->ms_per_spp > 0.0)
            count = (uint32_t)std::min(std::max(std::floor(PASS_TARGET_MS / L->ms_per_spp), 1.0), 32.0);
        count = std::min(count, max - L->film->samples);
        double ms;
        if ((rc = live_render(L, L->film, count, &ms))) return rc;
        now += ms;
        const double per = ms / count;
        L->ms_per_spp = L->ms_per_spp == 0.0 ? per : 0.75 * L->ms_per_spp + 0.25 * per;
        fr->pass_spp = count;
        fr->pass_ms = (float)ms;
        fr->ms_per_spp = (float)L->ms_per_spp;
        fr->spp = L->film->samples;
        const bool converged = L->film->samples >= max;
        const uint32_t pixels = L->film->w * L->film->h;
        const double interval = motion ? live_paced(L, MOTION_INTERVAL_MS, pixels)
                                       : live_refine_interval(L, L->film->samples, pixels);
        if (!converged && motion && L->film->samples < 4) {
            // while in motion the half-resolution look stands in until 4 spp
            if (L->published_gen != L->cur_gen &&
                live_due(L, now, live_paced(L, MOTION_INTERVAL_MS, L->half->w * L->half->h))) {
                if ((rc = live_publish(L, L->half, true, false, now, out, fr, &fin))) return rc;
                L->published_gen = L->cur_gen;
            }
            return AG_OK;
        }
        if (converged || live_due(L, now, interval)) {
            if ((rc = live_publish(L, L->film, false, converged, now, out, fr, &fin))) return rc;
            L->published_gen = L->cur_gen;
            if (fin) {
                L->final_done = true;
                if (!converged) {
                    L->settled = true;
                    L->settled_max = max;
                    fr->settled = 1;
                }
            }
        }
        return AG_OK;
    }
    if (!L->final_done) {
        // max_spp lowered below what the film holds: publish it as final
        if ((rc = live_publish(L, L->film, false, true, now, out, fr, &fin))) return rc;
        L->published_gen = L->cur_gen;
        L->final_done = true;
    }
    fr->spp = L->film->samples;
    return AG_OK;
}

extern "C" int ag_device_alloc(ag_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return fail(AG_E_INVALID, "null argument");
    HIPCHK(hipSetDevice(ctx->device));
    if (hipMalloc(out, std::max<size_t>(bytes, 1)) != hipSuccess) return fail(AG_E_NOMEM, "hipMalloc failed");
    return AG_OK;
}
extern "C" int ag_device_free(ag_ctx* ctx, void* p) {
    if (!ctx) return fail(AG_E_INVALID, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(p));
    return AG_OK;
}
extern "C" int ag_memcpy_to_device(ag_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(AG_E_INVALID, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AG_OK;
}
extern "C" int ag_memcpy_to_host(ag_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx) return fail(AG_E_INVALID, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return AG_OK;
}
extern "C" double ag_last_query_ms(const ag_ctx* ctx) { return ctx ? ctx->last_query_ms : 0.0; }

extern "C" int ag_intersect(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, ag_hits_soa* out) {
    if (!out) return fail(AG_E_INVALID, "null output");
    return ray_batch(ctx, s, rays, n, out, nullptr, false);
}
extern "C" int ag_occluded(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, uint8_t* out) {
    if (!out) return fail(AG_E_INVALID, "null output");
    return ray_batch(ctx, s, rays, n, nullptr, out, false);
}
extern "C" int ag_intersect_device(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, ag_hits_soa* out) {
    if (!out) return fail(AG_E_INVALID, "null output");
    return ray_batch(ctx, s, rays, n, out, nullptr, true);
}
extern "C" int ag_occluded_device(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n, uint8_t* out) {
    if (!out) return fail(AG_E_INVALID, "null output");
    return ray_batch(ctx, s, rays, n, nullptr, out, true);
}

// The closest-hit query of ag_intersect_device with no allocation: the
// scratch is the device's `scratch` array (cursors, then spill stacks for
// NOALLOC_BLOCKS blocks).  Enqueues a fill kernel for the cursors and the
// traversal kernel on the context stream and returns; no hipMalloc, hipFree,
// memset or synchronisation.  Calls that overlap on one device race on the
// array (it holds the last call's cursors and stacks).
extern "C" int ag_intersect_device_noalloc(ag_ctx* ctx, const ag_scene* s, const ag_rays_soa* rays, uint32_t n,
                                           ag_hits_soa* out) {
    if (!ctx || !s || !rays || !out) return fail(AG_E_INVALID, "null argument");
    if (!rays->ox || !rays->oy || !rays->oz || !rays->dx || !rays->dy || !rays->dz)
        return fail(AG_E_INVALID, "missing ray component");
    if (!out->t || !out->u || !out->v || !out->prim) return fail(AG_E_INVALID, "missing hit component");
    if (!ctx->noalloc_scratch) return fail(AG_E_STATE, "context has no scratch address for this device");
    if (n == 0) return AG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const int want = ctx->n_cu * (int)ctx->tuning.batch_blocks_per_cu;
    const int grid = std::max(1, std::min(want, (int)NOALLOC_BLOCKS));
    if ((size_t)grid * TRAV_BLOCK * STACK_SPILL * sizeof(uint2) > NOALLOC_SPILL_BYTES)
        return fail(AG_E_INVALID, "traversal grid exceeds the scratch array");
    unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
    uint32_t* cursors = reinterpret_cast<uint32_t*>(base);
    uint2* spill = reinterpret_cast<uint2*>(base + NOALLOC_CURSOR_BYTES);
    const RaysSoA r{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmax};
    const HitsSoA h{out->t, out->u, out->v, out->prim};
    launch_zero_u32(ctx->stream, cursors, CURSOR_SET);
    launch_rays_intersect(ctx->stream, grid, s->geo->nodes.as<float4>(), s->geo->tris.as<float4>(), r, h, n, cursors,
                          ctx->tuning.batch_refill_min, ctx->tuning.spread_min_lanes, spill);
    HIPCHK(hipGetLastError());
    return AG_OK;
}

extern "C" int ag_bsdf_probe(ag_ctx* ctx, const ag_scene* s, uint32_t n, const int32_t* mat, const float* lam,
                             const float* mask, const float* u, const float* wo, const float* wi, float* out) {
    if (!ctx || !s || !mat || !lam || !mask || !u || !wo || !wi || !out) return fail(AG_E_INVALID, "null argument");
    if (n == 0) return AG_OK;
    for (uint32_t i = 0; i < n; i++)
        if (mat[i] < 0 || mat[i] >= s->hs.n_mat) return fail(AG_E_INVALID, "material index out of range");
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    size_t floats = (size_t)n * (4 + 1 + 4 + 3 + 3);
    DevBuf in, ob, mb;
    int rc;
    if ((rc = in.ensure(floats * 4)) || (rc = ob.ensure((size_t)n * 18 * 4)) || (rc = mb.ensure((size_t)n * 4))) return rc;
    float* b = in.as<float>();
    HIPCHK(hipMemcpyAsync(mb.p, mat, (size_t)n * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b, lam, (size_t)n * 16, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + 4 * (size_t)n, mask, (size_t)n * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + 5 * (size_t)n, u, (size_t)n * 16, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + 9 * (size_t)n, wo, (size_t)n * 12, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(b + 12 * (size_t)n, wi, (size_t)n * 12, hipMemcpyHostToDevice, st));
    ProbeIn pi{mb.as<int32_t>(), b, b + 4 * (size_t)n, b + 5 * (size_t)n, b + 9 * (size_t)n, b + 12 * (size_t)n};
    launch_bsdf_probe(st, s->d_scene.as<DScene>(), pi, ob.as<float>(), n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ob.p, (size_t)n * 18 * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return AG_OK;
}
