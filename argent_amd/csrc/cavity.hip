// The patina mask's mesh steps on the device (argent-mesh cavity.rs):
// concavity (cavity.rs:130-243), the contrast-shaped blend with the ambient
// occlusion term (cavity.rs:64-73) and the Laplacian smoothing with crease
// copies welded (cavity.rs:245-296).  The ambient-occlusion rays themselves
// are the scene's any-hit traversal (k_ao in kernels.hip); this file gets
// their per-vertex counts.
//
// Every floating-point sum runs in the reference's order, so the result does
// not depend on scheduling: a vertex's area sums its triangles in triangle
// order (vertex -> triangle lists from a stable sort), a cell sums its
// vertices in (key, index) order (a stable sort of the cell keys over the
// vertex indices), a vertex visits its neighbour cells in the (dx, dy, dz)
// loop order, a crease class averages its copies in index order and a
// smoothing step sums a class's neighbours in increasing class order.
//
// Layout: cells are structure-of-arrays (keys apart, for the binary searches;
// mass, centroid and normal in f64).  The neighbour lists of a cell are found
// with one lower_bound per (dx, dy) column and a walk along dz, whose keys
// are consecutive; they are counted, scanned and then filled (CSR).  The
// concavity kernel takes vertices in cell order, so the lanes of a wave
// mostly read the same list.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "mesh_internal.hpp"
#include "patina.hpp"

namespace ag {
namespace {

constexpr uint64_t LIM21 = (1ull << 21) - 1;

struct D3 {
    double x, y, z;
};
__device__ inline D3 ld3(const float* p) { return {(double)p[0], (double)p[1], (double)p[2]}; }
__device__ inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ inline D3 dvd(D3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline D3 cross(D3 a, D3 b) { return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
// glam DVec3::normalize_or_zero: times the reciprocal length when finite and positive
__device__ inline D3 normalize_or_zero(D3 a) {
    const double r = 1.0 / sqrt(dot(a, a));
    return (isfinite(r) && r > 0.0) ? mul(a, r) : D3{0.0, 0.0, 0.0};
}

// Rust `f64 as u64` of a value already >= 0 or NaN: saturating, NaN to 0
__device__ inline uint64_t sat_u64(double q) {
    if (!(q > 0.0)) return 0;
    if (q >= 18446744073709551615.0) return ~0ull;
    return (uint64_t)q;
}
__device__ inline double max0(double q) { return q < 0.0 ? 0.0 : (q == q ? q : 0.0); }  // f64::max(q, 0.0)

// f32::clamp: NaN passes through
__device__ inline float rclamp(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

__device__ inline float smoothstep(float e0, float e1, float x) {  // cavity.rs:44-47
    const float t = rclamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

__device__ inline uint32_t canon_bits(float x) { return __float_as_uint(x + 0.0f); }  // weld.rs:151-154

// ---------------------------------------------------------------- vertex areas
__global__ __launch_bounds__(BLK) void k_tri_area(const float* pos, const uint32_t* tris, uint32_t nt, double* ar) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt) return;
    const uint32_t* tr = tris + 3 * (size_t)t;
    const D3 a = ld3(pos + 3 * (size_t)tr[0]), b = ld3(pos + 3 * (size_t)tr[1]), c = ld3(pos + 3 * (size_t)tr[2]);
    const D3 x = cross(sub(b, a), sub(c, a));
    ar[t] = 0.5 * sqrt(dot(x, x)) / 3.0;
}

__global__ __launch_bounds__(BLK) void k_corner_count(const uint32_t* tris, uint32_t n3, uint32_t* cnt) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n3) atomicAdd(&cnt[tris[i]], 1u);
}

__global__ __launch_bounds__(BLK) void k_iota(uint32_t* a, uint32_t n) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) a[i] = i;
}

// A vertex's corners in increasing corner (so triangle) order: one third of
// each incident triangle's area, summed as cavity.rs:143-150 does.
__global__ __launch_bounds__(BLK) void k_vertex_area(const uint32_t* off, const uint32_t* corners, const double* ar,
                                                     uint32_t nv, double* varea) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    double s = 0.0;
    for (uint32_t k = off[v]; k < off[v + 1]; k++) s += ar[corners[k] / 3u];
    varea[v] = s;
}

// ---------------------------------------------------------------- cells
__global__ __launch_bounds__(BLK) void k_cell_key(const float* pos, uint32_t nv, double lx, double ly, double lz,
                                                  double h, unsigned long long* key) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const D3 q = dvd(sub(ld3(pos + 3 * (size_t)v), D3{lx, ly, lz}), h);
    const uint64_t c0 = min(sat_u64(max0(q.x)), LIM21), c1 = min(sat_u64(max0(q.y)), LIM21),
                   c2 = min(sat_u64(max0(q.z)), LIM21);
    key[v] = (c0 << 42) | (c1 << 21) | c2;
}

__global__ __launch_bounds__(BLK) void k_run_heads(const unsigned long long* key, uint32_t n, uint32_t* head) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) head[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// start of each run; start[n_runs] = n
__global__ __launch_bounds__(BLK) void k_run_starts(const uint32_t* head, const uint32_t* rank, uint32_t n,
                                                    const uint32_t* n_runs, uint32_t* start) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n && head[i]) start[rank[i]] = i;
    if (i == 0) start[*n_runs] = n;
}

struct Cells {
    unsigned long long* key;
    double* m;
    double *px, *py, *pz, *nx, *ny, *nz;
    uint32_t* of;  // vertex -> cell
};

// cavity.rs:161-182: area-weighted centroid and normal of a cell
__global__ __launch_bounds__(BLK) void k_cells(const float* pos, const float* nrm, const double* varea,
                                               const unsigned long long* skey, const uint32_t* sidx,
                                               const uint32_t* start, const uint32_t* n_cells, Cells C) {
    const uint32_t c = blockIdx.x * BLK + threadIdx.x;
    if (c >= *n_cells) return;
    double m = 0.0;
    D3 p{0.0, 0.0, 0.0}, n{0.0, 0.0, 0.0};
    for (uint32_t k = start[c]; k < start[c + 1]; k++) {
        const uint32_t v = sidx[k];
        const double w = fmax(varea[v], 1e-12);
        m += w;
        p = add(p, mul(ld3(pos + 3 * (size_t)v), w));
        n = add(n, mul(ld3(nrm + 3 * (size_t)v), w));
        C.of[v] = c;
    }
    p = dvd(p, m);
    n = normalize_or_zero(n);
    C.key[c] = skey[start[c]];
    C.m[c] = m;
    C.px[c] = p.x;
    C.py[c] = p.y;
    C.pz[c] = p.z;
    C.nx[c] = n.x;
    C.ny[c] = n.y;
    C.nz[c] = n.z;
}

__device__ inline uint32_t lower_bound(const unsigned long long* key, uint32_t n, uint64_t k) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (key[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// cavity.rs:184-211: the cells within reach of cell c, in (dx, dy, dz) order,
// whose centroids lie within max_d.  FILL = false counts them.
template <bool FILL>
__global__ __launch_bounds__(BLK) void k_neighbours(Cells C, const uint32_t* n_cells_dev, int reach, double max_d,
                                                    uint32_t* cnt, const uint32_t* off, uint32_t* list) {
    const uint32_t n_cells = *n_cells_dev;
    const uint32_t c = blockIdx.x * BLK + threadIdx.x;
    if (c >= n_cells) return;
    const uint64_t k = C.key[c];
    const int64_t bx = (int64_t)(k >> 42), by = (int64_t)((k >> 21) & LIM21), bz = (int64_t)(k & LIM21);
    const int64_t lim = (int64_t)LIM21;
    const D3 p{C.px[c], C.py[c], C.pz[c]};
    const int64_t z0 = bz - reach < 0 ? 0 : bz - reach, z1 = bz + reach > lim ? lim : bz + reach;
    uint32_t n = 0;
    uint32_t* out = FILL ? list + off[c] : nullptr;
    for (int64_t dx = -reach; dx <= reach; dx++) {
        const int64_t qx = bx + dx;
        if (qx < 0 || qx > lim) continue;
        for (int64_t dy = -reach; dy <= reach; dy++) {
            const int64_t qy = by + dy;
            if (qy < 0 || qy > lim) continue;
            const uint64_t col = ((uint64_t)qx << 42) | ((uint64_t)qy << 21);
            const uint64_t hi = col | (uint64_t)z1;
            for (uint32_t j = lower_bound(C.key, n_cells, col | (uint64_t)z0); j < n_cells && C.key[j] <= hi; j++) {
                const D3 d = sub(D3{C.px[j], C.py[j], C.pz[j]}, p);
                if (sqrt(dot(d, d)) <= max_d) {
                    if (FILL) out[n] = j;
                    n++;
                }
            }
        }
    }
    if (!FILL) cnt[c] = n;
}

// cavity.rs:214-242, vertex v = sidx[i] (cell order)
__global__ __launch_bounds__(BLK) void k_concavity(const float* pos, const float* nrm, const uint32_t* sidx,
                                                   uint32_t nv, Cells C, const uint32_t* off, const uint32_t* list,
                                                   double radius, float* out) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv) return;
    const uint32_t v = sidx[i];
    const D3 p = ld3(pos + 3 * (size_t)v);
    const D3 n = normalize_or_zero(ld3(nrm + 3 * (size_t)v));
    if (n.x == 0.0 && n.y == 0.0 && n.z == 0.0) {
        out[v] = 0.0f;
        return;
    }
    const double r2 = radius * radius;
    const double sg = 0.5 * radius;
    const double inv2s2 = 1.0 / (2.0 * (sg * sg));
    double sw = 0.0, st = 0.0;
    D3 sd{0.0, 0.0, 0.0};
    const uint32_t c = C.of[v];
    for (uint32_t k = off[c]; k < off[c + 1]; k++) {
        const uint32_t j = list[k];
        const D3 d = sub(D3{C.px[j], C.py[j], C.pz[j]}, p);
        const double dist2 = dot(d, d);
        if (dist2 > r2 || dot(D3{C.nx[j], C.ny[j], C.nz[j]}, n) < -0.5) continue;
        const double w = C.m[j] * exp(-dist2 * inv2s2);
        sw += w;
        sd = add(sd, mul(d, w));
        const double dn = dot(d, n);
        st += w * (dist2 - dn * dn);
    }
    if (sw <= 0.0 || st <= 1e-18) {
        out[v] = 0.0f;
        return;
    }
    const double hgt = dot(dvd(sd, sw), n);
    const double dt2 = st / sw;
    const double mean_curv = 2.0 * hgt / dt2;
    const double val = mean_curv * radius;
    out[v] = (float)(val < -4.0 ? -4.0 : (val > 4.0 ? 4.0 : val));  // NaN stays NaN, as f64::clamp
}

// cavity.rs:64-73; ao = count / samples (cavity.rs:126); no AO term: counts null
__global__ __launch_bounds__(BLK) void k_blend(const uint32_t* counts, uint32_t samples, const float* cav, float w,
                                               uint32_t nv, float* ao_out, float* out) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const float a = counts ? (float)counts[v] / (float)samples : 0.0f;
    const float cv = cav ? cav[v] : 0.0f;
    if (ao_out) ao_out[v] = a;
    const float at = smoothstep(0.04f, 0.55f, a);
    const float ct = smoothstep(0.02f, 0.5f, cv);
    out[v] = smoothstep(0.05f, 0.9f, (1.0f - w) * at + w * ct);
}

// ---------------------------------------------------------------- smoothing
__global__ __launch_bounds__(BLK) void k_key_yz(const float* pos, const uint32_t* idx, uint32_t nv,
                                                unsigned long long* key) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv) return;
    const float* p = pos + 3 * (size_t)idx[i];
    key[i] = ((unsigned long long)canon_bits(p[1]) << 32) | canon_bits(p[2]);
}
__global__ __launch_bounds__(BLK) void k_key_x(const float* pos, const uint32_t* idx, uint32_t nv, uint32_t* key) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < nv) key[i] = canon_bits(pos[3 * (size_t)idx[i]]);
}
__device__ inline bool same_pos(const float* pos, uint32_t a, uint32_t b) {
    const float *p = pos + 3 * (size_t)a, *q = pos + 3 * (size_t)b;
    return canon_bits(p[0]) == canon_bits(q[0]) && canon_bits(p[1]) == canon_bits(q[1]) &&
           canon_bits(p[2]) == canon_bits(q[2]);
}
// sorted position i starts a run of equal coordinates?
__global__ __launch_bounds__(BLK) void k_pos_heads(const float* pos, const uint32_t* sidx, uint32_t nv,
                                                   uint32_t* head) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < nv) head[i] = (i == 0 || !same_pos(pos, sidx[i], sidx[i - 1])) ? 1u : 0u;
}
// rep[v] = the run head (the smallest index: the sorts are stable over 0..nv);
// is_rep[v] = 1 for the heads themselves
__global__ __launch_bounds__(BLK) void k_reps(const uint32_t* sidx, const uint32_t* start, const uint32_t* n_runs,
                                              uint32_t* rep, uint32_t* is_rep) {
    const uint32_t r = blockIdx.x * BLK + threadIdx.x;
    if (r >= *n_runs) return;
    const uint32_t h = sidx[start[r]];
    is_rep[h] = 1u;
    for (uint32_t k = start[r]; k < start[r + 1]; k++) rep[sidx[k]] = h;
}
// weld.rs:208-223: class ids in first-occurrence order
__global__ __launch_bounds__(BLK) void k_uid(const uint32_t* rep, const uint32_t* rank, uint32_t nv, uint32_t* uid) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v < nv) uid[v] = rank[rep[v]];
}
// cavity.rs:258-265: a class starts at the f64 mean of its copies (index order)
__global__ __launch_bounds__(BLK) void k_class_mean(const float* values, const uint32_t* sidx, const uint32_t* start,
                                                    const uint32_t* n_runs, const uint32_t* uid, float* x) {
    const uint32_t r = blockIdx.x * BLK + threadIdx.x;
    if (r >= *n_runs) return;
    double s = 0.0;
    for (uint32_t k = start[r]; k < start[r + 1]; k++) s += (double)values[sidx[k]];
    x[uid[sidx[start[r]]]] = (float)(s / (double)(start[r + 1] - start[r]));
}
// cavity.rs:267-273: six directed edges per triangle over the classes
__global__ __launch_bounds__(BLK) void k_edges(const uint32_t* tris, uint32_t nt, const uint32_t* uid,
                                               unsigned long long* e) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt) return;
    const unsigned long long c0 = uid[tris[3 * (size_t)t]], c1 = uid[tris[3 * (size_t)t + 1]],
                             c2 = uid[tris[3 * (size_t)t + 2]];
    unsigned long long* o = e + 6 * (size_t)t;
    o[0] = (c0 << 32) | c1;
    o[1] = (c1 << 32) | c0;
    o[2] = (c1 << 32) | c2;
    o[3] = (c2 << 32) | c1;
    o[4] = (c2 << 32) | c0;
    o[5] = (c0 << 32) | c2;
}
// distinct sorted edges: neighbour list and per-class counts
__global__ __launch_bounds__(BLK) void k_edge_lists(const unsigned long long* e, const uint32_t* head,
                                                    const uint32_t* rank, uint32_t n, uint32_t* nbr, uint32_t* cnt) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n || !head[i]) return;
    nbr[rank[i]] = (uint32_t)e[i];
    atomicAdd(&cnt[(uint32_t)(e[i] >> 32)], 1u);
}
// cavity.rs:278-292: y = x + lambda (mean of the others - x)
__global__ __launch_bounds__(BLK) void k_smooth_step(const uint32_t* off, const uint32_t* nbr, const float* x,
                                                     const uint32_t* nc, float lambda, float* y) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= *nc) return;
    const uint32_t a = off[i], b = off[i + 1];
    const float xi = x[i];
    if (b > a) {
        float s = 0.0f;
        for (uint32_t k = a; k < b; k++) {
            const uint32_t j = nbr[k];
            if (j != i) s += x[j];
        }
        const float m = s / (float)(b - a);
        y[i] = xi + lambda * (m - xi);
    } else {
        y[i] = xi;
    }
}
// cavity.rs:293-295 and :76-78: back to the vertices, clamped to [0, 1]
__global__ __launch_bounds__(BLK) void k_scatter_clamp(const float* x, const uint32_t* uid, uint32_t nv, int clamp01,
                                                       float* out) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const float a = x[uid[v]];
    out[v] = clamp01 ? rclamp(a, 0.0f, 1.0f) : a;
}


// ---------------------------------------------------------------- smooth normals
// util.rs:39-50: the unit face normal in f64, rounded to f32 (zero if degenerate)
__global__ __launch_bounds__(BLK) void k_face_nrm(const float* pos, const uint32_t* tris, uint32_t nt, float* fn) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    const uint32_t* t = tris + 3 * (size_t)f;
    const D3 a = ld3(pos + 3 * (size_t)t[0]), b = ld3(pos + 3 * (size_t)t[1]), c = ld3(pos + 3 * (size_t)t[2]);
    const D3 x = cross(sub(b, a), sub(c, a));
    const double l = sqrt(dot(x, x));
    const bool ok = l > 0.0 && isfinite(l);
    fn[3 * (size_t)f] = ok ? (float)(x.x / l) : 0.0f;
    fn[3 * (size_t)f + 1] = ok ? (float)(x.y / l) : 0.0f;
    fn[3 * (size_t)f + 2] = ok ? (float)(x.z / l) : 0.0f;
}
// normals.rs:226-247: a vertex's corners in corner order, each adding its
// face normal times the corner angle (f64); normalized, or +z
__global__ __launch_bounds__(BLK) void k_smooth_nrm(const float* pos, const uint32_t* tris, const uint32_t* off,
                                                    const uint32_t* corners, const float* fn, uint32_t nv,
                                                    float* out) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    D3 acc{0.0, 0.0, 0.0};
    const D3 p = ld3(pos + 3 * (size_t)v);
    for (uint32_t i = off[v]; i < off[v + 1]; i++) {
        const uint32_t c = corners[i], f = c / 3u, k = c - 3u * f;
        const uint32_t* t = tris + 3 * (size_t)f;
        const D3 ea = sub(ld3(pos + 3 * (size_t)t[(k + 1) % 3]), p);
        const D3 eb = sub(ld3(pos + 3 * (size_t)t[(k + 2) % 3]), p);
        const D3 x = cross(ea, eb);
        const double ang = atan2(sqrt(dot(x, x)), dot(ea, eb));
        acc = add(acc, D3{ang * (double)fn[3 * (size_t)f], ang * (double)fn[3 * (size_t)f + 1],
                          ang * (double)fn[3 * (size_t)f + 2]});
    }
    const double l = sqrt(dot(acc, acc));
    const bool ok = l > 1e-300;
    out[3 * (size_t)v] = ok ? (float)(acc.x / l) : 0.0f;
    out[3 * (size_t)v + 1] = ok ? (float)(acc.y / l) : 0.0f;
    out[3 * (size_t)v + 2] = ok ? (float)(acc.z / l) : 1.0f;
}

// Stable radix sort of (key, value) pairs; keys_out may be scratch.
template <class K>
int sort_pairs(void* w, size_t bytes, const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, uint32_t n,
               unsigned bits, hipStream_t st) {
    MCHK(rocprim::radix_sort_pairs(w, bytes, kin, kout, vin, vout, n, 0, bits, st));
    return AG_OK;
}

}  // namespace

int smooth_normals_device(ag_ctx* ctx, const float* pos, const uint32_t* tris, uint32_t nv, uint32_t nt,
                          float* out) {
    hipStream_t st = ctx_stream(ctx);
    if (nv == 0) return AG_OK;
    if ((uint64_t)3 * nt >= 0xFFFFFFFFull) return fail(AG_E_INVALID, "too many triangles for smooth normals");
    const uint32_t n3 = 3 * nt;
    unsigned vbits = 1;
    while (vbits < 32 && (1ull << vbits) <= nv) vbits++;
    const size_t NV = (size_t)nv + 1, N3 = (size_t)n3 + 1;
    size_t sort_bytes = 0;
    MCHK(rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, n3, 0, vbits, st));
    Arena a;
    const size_t o_sort = a.add(sort_bytes), o_scan = a.add(4 * device_scan_work_words(std::max(NV, N3))),
                 o_tot = a.add(256), o_cnt = a.add(4 * NV), o_off = a.add(4 * NV), o_cidx = a.add(4 * N3),
                 o_corner = a.add(4 * N3), o_ckey = a.add(4 * N3), o_fn = a.add(12 * (size_t)std::max(nt, 1u));
    int rc = a.bind(ctx);
    if (rc) return rc;
    uint32_t *cnt = a.at<uint32_t>(o_cnt), *off = a.at<uint32_t>(o_off), *cidx = a.at<uint32_t>(o_cidx),
             *corner = a.at<uint32_t>(o_corner);
    float* fn = a.at<float>(o_fn);
    MCHK(hipMemsetAsync(cnt, 0, 4 * NV, st));
    if (nt) {
        k_face_nrm<<<grid_of(nt), BLK, 0, st>>>(pos, tris, nt, fn);
        k_corner_count<<<grid_of(n3), BLK, 0, st>>>(tris, n3, cnt);
    }
    device_scan_u32(st, cnt, off, nv + 1, a.at<uint32_t>(o_scan), a.at<uint32_t>(o_tot));
    if (nt) {
        k_iota<<<grid_of(n3), BLK, 0, st>>>(cidx, n3);
        if ((rc = sort_pairs(a.at<void>(o_sort), sort_bytes, tris, a.at<uint32_t>(o_ckey), cidx, corner, n3, vbits,
                             st)))
            return rc;
    }
    k_smooth_nrm<<<grid_of(nv), BLK, 0, st>>>(pos, tris, off, corner, fn, nv, out);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    return AG_OK;
}

int patina_device(ag_ctx* ctx, const PatinaInputs& in, const ag_patina_settings& s, float* mask, float* ao_out,
                  float* cav_out) {
    hipStream_t st = ctx_stream(ctx);
    const uint32_t nv = in.nv, nt = in.nt;
    if (nv == 0) return AG_OK;
    if ((uint64_t)6 * nt >= 0xFFFFFFFFull) return fail(AG_E_INVALID, "too many triangles for the patina mask");
    const float w = std::fmin(std::fmax(s.concavity_weight, 0.0f), 1.0f);
    const bool with_cav = !in.values && w > 0.0f && s.curvature_radius > 0.0f && nt > 0;
    const double radius = (double)s.curvature_radius, h = radius / 2.0;
    const int reach = with_cav ? (int)std::ceil(radius / h) + 1 : 0;
    const uint32_t n3 = 3 * nt, ne = 6 * nt;
    unsigned vbits = 1;
    while (vbits < 32 && (1ull << vbits) <= nv) vbits++;
    const size_t NV = (size_t)nv + 1, N3 = (size_t)n3 + 1, NE = (size_t)ne + 1;
    size_t sort_bytes = 0, b = 0;
    MCHK(rocprim::radix_sort_pairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                   (uint32_t*)nullptr, n3, 0, vbits, st));
    sort_bytes = std::max(sort_bytes, b);
    MCHK(rocprim::radix_sort_pairs(nullptr, b, (const unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                   (const uint32_t*)nullptr, (uint32_t*)nullptr, nv, 0, 64, st));
    sort_bytes = std::max(sort_bytes, b);
    MCHK(rocprim::radix_sort_pairs(nullptr, b, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
                                   (uint32_t*)nullptr, nv, 0, 32, st));
    sort_bytes = std::max(sort_bytes, b);
    MCHK(rocprim::radix_sort_keys(nullptr, b, (const unsigned long long*)nullptr, (unsigned long long*)nullptr, ne, 0,
                                  32 + vbits, st));
    sort_bytes = std::max(sort_bytes, b);
    Arena a;
    const size_t o_sort = a.add(sort_bytes), o_scan = a.add(4 * device_scan_work_words(std::max(std::max(NE, N3), NV))),
                 o_tot = a.add(256),
                 // per vertex
                 o_k64a = a.add(8 * NV), o_k64b = a.add(8 * NV), o_iota = a.add(4 * NV), o_sidx = a.add(4 * NV),
                 o_head = a.add(4 * NV), o_rank = a.add(4 * NV), o_start = a.add(4 * NV), o_cnt = a.add(4 * NV),
                 o_off = a.add(4 * NV), o_varea = a.add(8 * NV), o_cells = a.add(8 * 8 * NV), o_of = a.add(4 * NV),
                 o_cav = a.add(4 * NV), o_vals = a.add(4 * NV), o_x = a.add(4 * NV), o_y = a.add(4 * NV),
                 o_rep = a.add(4 * NV), o_isrep = a.add(4 * NV), o_uid = a.add(4 * NV), o_u32k = a.add(4 * NV),
                 // per triangle corner / edge
                 o_area = a.add(8 * (size_t)std::max(nt, 1u)), o_cidx = a.add(4 * N3), o_corner = a.add(4 * N3),
                 o_ckey = a.add(4 * N3), o_e = a.add(8 * NE), o_es = a.add(8 * NE), o_ehead = a.add(4 * NE),
                 o_erank = a.add(4 * NE), o_nbr = a.add(4 * NE);
    int rc = a.bind(ctx);
    if (rc) return rc;
    void* sw = a.at<void>(o_sort);
    uint32_t* scanw = a.at<uint32_t>(o_scan);
    uint32_t* tot = a.at<uint32_t>(o_tot);  // [0] cells, [1] list entries, [2] runs, [3] classes, [4] edges, [5] scratch
    unsigned long long *k64a = a.at<unsigned long long>(o_k64a), *k64b = a.at<unsigned long long>(o_k64b);
    uint32_t *iota = a.at<uint32_t>(o_iota), *sidx = a.at<uint32_t>(o_sidx), *head = a.at<uint32_t>(o_head),
             *rank = a.at<uint32_t>(o_rank), *start = a.at<uint32_t>(o_start), *cnt = a.at<uint32_t>(o_cnt),
             *off = a.at<uint32_t>(o_off);
    float* cav = with_cav ? a.at<float>(o_cav) : nullptr;
    float *vals = a.at<float>(o_vals), *x = a.at<float>(o_x), *y = a.at<float>(o_y);
    k_iota<<<grid_of(nv), BLK, 0, st>>>(iota, nv);

    // ---- concavity (cavity.rs:130-243)
    Buf list;
    if (with_cav) {
        double* ar = a.at<double>(o_area);
        double* varea = a.at<double>(o_varea);
        uint32_t *cidx = a.at<uint32_t>(o_cidx), *corner = a.at<uint32_t>(o_corner);
        k_tri_area<<<grid_of(nt), BLK, 0, st>>>(in.pos, in.tris, nt, ar);
        // vertex -> corners in corner order (a stable sort of the corner vertices over 0..3nt)
        MCHK(hipMemsetAsync(cnt, 0, 4 * NV, st));
        k_corner_count<<<grid_of(n3), BLK, 0, st>>>(in.tris, n3, cnt);
        device_scan_u32(st, cnt, off, nv + 1, scanw, tot + 5);
        k_iota<<<grid_of(n3), BLK, 0, st>>>(cidx, n3);
        if ((rc = sort_pairs(sw, sort_bytes, in.tris, a.at<uint32_t>(o_ckey), cidx, corner, n3, vbits, st))) return rc;
        k_vertex_area<<<grid_of(nv), BLK, 0, st>>>(off, corner, ar, nv, varea);
        // cells: vertices sorted by (key, index), runs of one key
        k_cell_key<<<grid_of(nv), BLK, 0, st>>>(in.pos, nv, (double)in.lo[0], (double)in.lo[1], (double)in.lo[2], h,
                                                k64a);
        if ((rc = sort_pairs(sw, sort_bytes, k64a, k64b, iota, sidx, nv, 64, st))) return rc;
        k_run_heads<<<grid_of(nv), BLK, 0, st>>>(k64b, nv, head);
        device_scan_u32(st, head, rank, nv, scanw, tot + 0);
        k_run_starts<<<grid_of(nv), BLK, 0, st>>>(head, rank, nv, tot + 0, start);
        double* cb = a.at<double>(o_cells);
        Cells C{a.at<unsigned long long>(o_k64a), cb, cb + NV, cb + 2 * NV, cb + 3 * NV, cb + 4 * NV, cb + 5 * NV,
                cb + 6 * NV, a.at<uint32_t>(o_of)};
        k_cells<<<grid_of(nv), BLK, 0, st>>>(in.pos, in.nrm, varea, k64b, sidx, start, tot + 0, C);
        // neighbour lists: count, scan, fill
        const double max_d = radius + 2.0 * h;
        k_neighbours<false><<<grid_of(nv), BLK, 0, st>>>(C, tot + 0, reach, max_d, cnt, nullptr, nullptr);
        MCHK(hipGetLastError());
        uint32_t n_cells = 0;
        MCHK(hipMemcpyAsync(&n_cells, tot + 0, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
        uint32_t n_list = 0;
        MCHK(hipMemsetAsync(cnt + n_cells, 0, 4, st));  // past the counts: the scan's total adds it
        device_scan_u32(st, cnt, off, n_cells + 1, scanw, tot + 1);
        MCHK(hipMemcpyAsync(&n_list, tot + 1, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
        if ((rc = list.alloc(4 * (size_t)std::max(n_list, 1u)))) return rc;
        k_neighbours<true><<<grid_of(nv), BLK, 0, st>>>(C, tot + 0, reach, max_d, nullptr, off, list.as<uint32_t>());
        k_concavity<<<grid_of(nv), BLK, 0, st>>>(in.pos, in.nrm, sidx, nv, C, off, list.as<uint32_t>(), radius, cav);
        MCHK(hipGetLastError());
        if (cav_out) MCHK(hipMemcpyAsync(cav_out, cav, 4 * (size_t)nv, hipMemcpyDeviceToDevice, st));
    } else if (cav_out) {
        MCHK(hipMemsetAsync(cav_out, 0, 4 * (size_t)nv, st));
    }

    // ---- blend (cavity.rs:64-73), or the caller's values to smooth
    if (in.values) MCHK(hipMemcpyAsync(vals, in.values, 4 * (size_t)nv, hipMemcpyDeviceToDevice, st));
    else k_blend<<<grid_of(nv), BLK, 0, st>>>(in.ao_counts, std::max(in.ao_samples, 1u), cav, w, nv, ao_out, vals);

    // ---- smoothing (cavity.rs:245-296): classes of exactly equal positions
    // (weld.rs:158-228 at eps 0), by a stable sort on (y, z) then on x
    k_key_yz<<<grid_of(nv), BLK, 0, st>>>(in.pos, iota, nv, k64a);
    if ((rc = sort_pairs(sw, sort_bytes, k64a, k64b, iota, rank, nv, 64, st))) return rc;
    uint32_t* kx = a.at<uint32_t>(o_u32k);
    k_key_x<<<grid_of(nv), BLK, 0, st>>>(in.pos, rank, nv, kx);
    if ((rc = sort_pairs(sw, sort_bytes, kx, head, rank, sidx, nv, 32, st))) return rc;
    k_pos_heads<<<grid_of(nv), BLK, 0, st>>>(in.pos, sidx, nv, head);
    device_scan_u32(st, head, rank, nv, scanw, tot + 2);
    k_run_starts<<<grid_of(nv), BLK, 0, st>>>(head, rank, nv, tot + 2, start);
    uint32_t *rep = a.at<uint32_t>(o_rep), *isrep = a.at<uint32_t>(o_isrep), *uid = a.at<uint32_t>(o_uid);
    MCHK(hipMemsetAsync(isrep, 0, 4 * NV, st));
    k_reps<<<grid_of(nv), BLK, 0, st>>>(sidx, start, tot + 2, rep, isrep);
    device_scan_u32(st, isrep, cnt, nv, scanw, tot + 3);
    k_uid<<<grid_of(nv), BLK, 0, st>>>(rep, cnt, nv, uid);
    k_class_mean<<<grid_of(nv), BLK, 0, st>>>(vals, sidx, start, tot + 2, uid, x);
    if (s.smooth_iterations > 0 && nt > 0) {
        unsigned long long *e = a.at<unsigned long long>(o_e), *es = a.at<unsigned long long>(o_es);
        uint32_t *eh = a.at<uint32_t>(o_ehead), *er = a.at<uint32_t>(o_erank), *nbr = a.at<uint32_t>(o_nbr);
        k_edges<<<grid_of(nt), BLK, 0, st>>>(in.tris, nt, uid, e);
        MCHK(rocprim::radix_sort_keys(sw, sort_bytes, e, es, ne, 0, 32 + vbits, st));
        k_run_heads<<<grid_of(ne), BLK, 0, st>>>(es, ne, eh);
        device_scan_u32(st, eh, er, ne, scanw, tot + 4);
        MCHK(hipMemsetAsync(cnt, 0, 4 * NV, st));
        k_edge_lists<<<grid_of(ne), BLK, 0, st>>>(es, eh, er, ne, nbr, cnt);
        device_scan_u32(st, cnt, off, nv + 1, scanw, tot + 5);
        for (uint32_t it = 0; it < s.smooth_iterations; it++) {
            k_smooth_step<<<grid_of(nv), BLK, 0, st>>>(off, nbr, x, tot + 3, in.lambda, y);
            std::swap(x, y);
        }
    }
    k_scatter_clamp<<<grid_of(nv), BLK, 0, st>>>(x, uid, nv, in.clamp01, mask);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));  // `list` is freed on return
    return AG_OK;
}

}  // namespace ag
