// Crease-aware vertex normals on the device (argent-mesh normals.rs:47-220,
// one crease angle): for every vertex the incident faces are grouped into
// smooth fans (faces sharing an edge at the vertex, wound consistently, with
// normals within the crease angle), each fan becomes one output vertex with
// the corner-angle-weighted mean of its face normals, and fans with nearly
// the same normal merge again.  Before that, inconsistently wound faces are
// flipped to agree with their edge-connected component (majority wins) and
// closed inside-out shells are flipped unless another shell encloses them.
//
// Layout: one thread per vertex for the fan clustering, with its scratch
// (edge entries, angles, union-find, sums) at the vertex's offsets in the
// vertex -> face lists, so any fan size fits.  Winding repair propagates a
// (component, parity) label over manifold edges instead of a depth-first
// walk; shells reuse the lock-free smallest-root union-find; per-shell sums
// use one atomic per wave when a wave's vertices share a shell.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "mesh_internal.hpp"

namespace ag {
namespace {

constexpr uint32_t MAX_FAN = 60000;  // normals.rs:303
constexpr uint16_t NONE16 = 0xFFFF;

struct D3 {
    double x, y, z;
};
__device__ inline D3 d3(const float* p) { return {(double)p[0], (double)p[1], (double)p[2]}; }
__device__ inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
__device__ inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline double len(D3 a) { return sqrt(dot(a, a)); }

// normals.rs:243-246
__device__ inline void normalize_or_z(D3 a, float* o) {
    const double l = len(a);
    if (l > 1e-300) {
        o[0] = (float)(a.x / l);
        o[1] = (float)(a.y / l);
        o[2] = (float)(a.z / l);
    } else {
        o[0] = 0.0f;
        o[1] = 0.0f;
        o[2] = 1.0f;
    }
}

// util.rs:39-50, with a count of zero-area faces (normals.rs:138)
__global__ __launch_bounds__(BLK) void k_face_normals(const float* pos, const uint32_t* tris, uint32_t nt, float* fn,
                                                      uint32_t* n_degenerate) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    bool degen = false;
    if (f < nt) {
        const D3 a = d3(pos + 3 * (size_t)tris[3 * (size_t)f]), b = d3(pos + 3 * (size_t)tris[3 * (size_t)f + 1]),
                 c = d3(pos + 3 * (size_t)tris[3 * (size_t)f + 2]);
        const D3 n = cross(sub(b, a), sub(c, a));
        const double l = len(n);
        float o[3] = {0.0f, 0.0f, 0.0f};
        if (l > 0.0 && isfinite(l)) {
            o[0] = (float)(n.x / l);
            o[1] = (float)(n.y / l);
            o[2] = (float)(n.z / l);
            degen = !((float)(0.5 * l) > 0.0f);
        } else {
            degen = true;
        }
        for (int k = 0; k < 3; k++) fn[3 * (size_t)f + k] = o[k];
    }
    const unsigned long long m = __ballot(degen);
    if (n_degenerate && (threadIdx.x & 63) == 0 && m) atomicAdd(n_degenerate, (uint32_t)__popcll(m));
}

// ---------------------------------------------------------------- vertex -> faces
__global__ __launch_bounds__(BLK) void k_vf_count(const uint32_t* tris, uint32_t nt, uint32_t* cnt) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    for (int k = 0; k < 3; k++) atomicAdd(&cnt[tris[3 * (size_t)f + k]], 1u);
}
// Face index of every incidence (3 f + k -> f): sorted stably by vertex with
// the triangles as keys, they give each vertex's faces in increasing order
// (util.rs:96-128).
__global__ __launch_bounds__(BLK) void k_incidence_faces(uint32_t nt, uint32_t* vals) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < 3 * nt) vals[i] = i / 3;
}

__global__ __launch_bounds__(BLK) void k_double_offsets(const uint32_t* off, uint32_t nv, uint32_t* off2) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v <= nv) off2[v] = 2 * off[v];
}

// ---------------------------------------------------------------- fan clustering
struct Cluster {
    const float* pos;
    const uint32_t* tris;
    const float* fn;
    const uint32_t* off;
    const uint32_t* faces;
    unsigned long long* ent;  // 2 per incidence: other vertex << 32 | face slot << 1 | forward
    double* angle;            // 1 per incidence
    uint16_t* par;            // 1 per incidence: fan union-find, then cluster union-find
    uint16_t* label;          // 1 per incidence: the fan of each face slot (output)
    uint16_t* remap;          // 1 per incidence
    double* sums;             // 4 per incidence: fan normal sum xyz, weight
    float* fan;               // 3 per incidence: fan normals of vertex v at off[v] + c (output)
    uint32_t* nfan;           // per vertex (output)
    uint32_t* vbound;         // boundary edges per vertex (output)
    unsigned long long* counts;  // boundary, non-manifold, inconsistent edges
    double cos_join, cos_merge;
};

__device__ inline uint16_t find16(uint16_t* p, uint16_t x) {
    while (p[x] != x) {
        const uint16_t g = p[p[x]];
        p[x] = g;
        x = g;
    }
    return x;
}
__device__ inline void union16(uint16_t* p, uint16_t a, uint16_t b) {
    const uint16_t ra = find16(p, a), rb = find16(p, b);
    if (ra == rb) return;
    if (ra < rb) p[rb] = ra;
    else p[ra] = rb;
}

__device__ void accumulate(const Cluster& C, uint32_t base, uint32_t k, uint32_t ncl) {
    double* s = C.sums + 4 * (size_t)base;
    for (uint32_t c = 0; c < 4 * ncl; c++) s[c] = 0.0;
    for (uint32_t j = 0; j < k; j++) {
        const uint32_t l = C.label[base + j];
        const D3 nf = d3(C.fn + 3 * (size_t)C.faces[base + j]);
        const double a = C.angle[base + j];
        s[4 * l] += a * nf.x;
        s[4 * l + 1] += a * nf.y;
        s[4 * l + 2] += a * nf.z;
        s[4 * l + 3] += a * len(nf);
    }
}

// Corner angles and the unsorted edge entries of every vertex's fan
// (normals.rs:391-405).
__global__ __launch_bounds__(BLK) void k_fan_prep(Cluster C, uint32_t nv, unsigned long long* ent_in) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const uint32_t base = C.off[v], k = C.off[v + 1] - base;
    const D3 pv = d3(C.pos + 3 * (size_t)v);
    for (uint32_t j = 0; j < k; j++) {
        const uint32_t* t = C.tris + 3 * (size_t)C.faces[base + j];
        const int slot = t[0] == v ? 0 : (t[1] == v ? 1 : 2);
        const uint32_t a = t[(slot + 1) % 3], b = t[(slot + 2) % 3];
        const D3 ea = sub(d3(C.pos + 3 * (size_t)a), pv), eb = sub(d3(C.pos + 3 * (size_t)b), pv);
        C.angle[base + j] = atan2(len(cross(ea, eb)), dot(ea, eb));
        const unsigned long long jj = (unsigned long long)(j & 0xFFFF) << 1;  // fans above MAX_FAN never read them
        ent_in[2 * (size_t)base + 2 * j] = (unsigned long long)a << 32 | jj | 1ull;
        ent_in[2 * (size_t)base + 2 * j + 1] = (unsigned long long)b << 32 | jj;
    }
}

// normals.rs:375-542 (cluster_vertex), with the fan's edge entries already
// sorted per vertex.
__global__ __launch_bounds__(BLK) void k_cluster(Cluster C, uint32_t nv) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    unsigned long long nb = 0, nnm = 0, ninc = 0;
    if (v < nv) {
        const uint32_t base = C.off[v], k = C.off[v + 1] - base;
        C.vbound[v] = 0;
        if (k == 0) {
            C.nfan[v] = 0;
        } else {
            if (k > MAX_FAN) {
                // a pathological fan: one smooth normal, no splitting
                D3 sum = {0.0, 0.0, 0.0};
                for (uint32_t j = 0; j < k; j++) {
                    const D3 n = d3(C.fn + 3 * (size_t)C.faces[base + j]);
                    const double a = C.angle[base + j];
                    sum.x += a * n.x;
                    sum.y += a * n.y;
                    sum.z += a * n.z;
                }
                for (uint32_t j = 0; j < k; j++) C.label[base + j] = 0;
                normalize_or_z(sum, C.fan + 3 * (size_t)base);
                C.nfan[v] = 1;
            } else {
                const unsigned long long* ent = C.ent + 2 * (size_t)base;
                uint16_t* par = C.par + base;
                for (uint32_t j = 0; j < k; j++) par[j] = (uint16_t)j;
                for (uint32_t i = 0; i < 2 * k;) {
                    const uint32_t w = (uint32_t)(ent[i] >> 32);
                    uint32_t e = i + 1;
                    while (e < 2 * k && (uint32_t)(ent[e] >> 32) == w) e++;
                    const uint32_t m = e - i;
                    if (v < w) {
                        if (m == 1) {
                            nb++;
                            C.vbound[v]++;
                        } else if (m == 2) {
                            if ((ent[i] & 1) == (ent[i + 1] & 1)) ninc++;
                        } else {
                            nnm++;
                        }
                    }
                    if (m >= 2)
                        for (uint32_t x = i; x < e; x++)
                            for (uint32_t y = x + 1; y < e; y++) {
                                const uint16_t jx = (uint16_t)((ent[x] >> 1) & 0xFFFF),
                                               jy = (uint16_t)((ent[y] >> 1) & 0xFFFF);
                                if ((ent[x] & 1) == (ent[y] & 1) || jx == jy) continue;
                                const D3 nx = d3(C.fn + 3 * (size_t)C.faces[base + jx]);
                                const D3 ny = d3(C.fn + 3 * (size_t)C.faces[base + jy]);
                                if (dot(nx, ny) >= C.cos_join) union16(par, jx, jy);
                            }
                    i = e;
                }
                // fan labels in face order
                uint16_t* rm = C.remap + base;
                for (uint32_t j = 0; j < k; j++) rm[j] = NONE16;
                uint32_t ncl = 0;
                for (uint32_t j = 0; j < k; j++) {
                    const uint16_t r = find16(par, (uint16_t)j);
                    if (rm[r] == NONE16) rm[r] = (uint16_t)ncl++;
                    C.label[base + j] = rm[r];
                }
                accumulate(C, base, k, ncl);
                if (ncl > 1) {
                    // merge fans with (nearly) the same normal, and fans with no
                    // usable normal into the heaviest (normals.rs:489-535)
                    const double* s = C.sums + 4 * (size_t)base;
                    for (uint32_t c = 0; c < ncl; c++) par[c] = (uint16_t)c;
                    auto has = [&](uint32_t c) { return len(D3{s[4 * c], s[4 * c + 1], s[4 * c + 2]}) > 1e-12; };
                    auto unit = [&](uint32_t c) {
                        const D3 a = {s[4 * c], s[4 * c + 1], s[4 * c + 2]};
                        const double l = len(a);
                        return D3{a.x / l, a.y / l, a.z / l};
                    };
                    int heaviest = -1;
                    for (uint32_t c = 0; c < ncl; c++)
                        if (has(c) && (heaviest < 0 || !(s[4 * c + 3] < s[4 * heaviest + 3]))) heaviest = (int)c;
                    for (uint32_t a = 0; a < ncl; a++) {
                        if (!has(a)) {
                            if (heaviest >= 0) union16(par, (uint16_t)a, (uint16_t)heaviest);
                        } else {
                            const D3 na = unit(a);
                            for (uint32_t b = a + 1; b < ncl; b++)
                                if (has(b) && dot(na, unit(b)) >= C.cos_merge) union16(par, (uint16_t)a, (uint16_t)b);
                        }
                    }
                    for (uint32_t c = 0; c < ncl; c++) rm[c] = NONE16;
                    uint32_t mcount = 0;
                    for (uint32_t c = 0; c < ncl; c++) {
                        const uint16_t r = find16(par, (uint16_t)c);
                        if (rm[r] == NONE16) rm[r] = (uint16_t)mcount++;
                    }
                    if (mcount < ncl) {
                        for (uint32_t j = 0; j < k; j++) C.label[base + j] = rm[find16(par, C.label[base + j])];
                        ncl = mcount;
                        accumulate(C, base, k, ncl);
                    }
                }
                const double* s = C.sums + 4 * (size_t)base;
                for (uint32_t c = 0; c < ncl; c++)
                    normalize_or_z(D3{s[4 * c], s[4 * c + 1], s[4 * c + 2]}, C.fan + 3 * (size_t)(base + c));
                C.nfan[v] = ncl;
            }
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        nb += __shfl_xor(nb, d, 64);
        nnm += __shfl_xor(nnm, d, 64);
        ninc += __shfl_xor(ninc, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nb) atomicAdd(&C.counts[0], nb);
        if (nnm) atomicAdd(&C.counts[1], nnm);
        if (ninc) atomicAdd(&C.counts[2], ninc);
    }
}

// ---------------------------------------------------------------- winding repair
// normals.rs:548-610.  Over each manifold edge (exactly one other face holds
// both ends) the depth-first walk sets flip[g] = flip[f] xor (g winds the
// edge the same way as f).  Here: adj[3 f + e] = g | same << 31 (or NONE),
// then a label w = component << 1 | parity, lowered by atomicMin from the
// neighbours and by jumping to the label's own label until nothing changes:
// the component is its smallest face (the walk's seed) and the parity is
// the walk's flip, for every orientable component.
__global__ __launch_bounds__(BLK) void k_face_adj(const uint32_t* tris, uint32_t nt, const uint32_t* off,
                                                  const uint32_t* faces, uint32_t* adj) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    const uint32_t* t = tris + 3 * (size_t)f;
    for (int e = 0; e < 3; e++) {
        const uint32_t a = t[e], b = t[(e + 1) % 3];
        uint32_t count = 0, nbf = NONE;
        for (uint32_t i = off[a]; i < off[a + 1]; i++) {
            const uint32_t g = faces[i];
            const uint32_t* tg = tris + 3 * (size_t)g;
            if (g != f && (tg[0] == b || tg[1] == b || tg[2] == b)) {
                count++;
                nbf = g;
            }
        }
        uint32_t out = NONE;
        if (count == 1) {
            const uint32_t* tg = tris + 3 * (size_t)nbf;
            bool same = false;
            for (int k = 0; k < 3; k++) same = same || (tg[k] == a && tg[(k + 1) % 3] == b);
            out = nbf | (same ? 0x80000000u : 0u);
        }
        adj[3 * (size_t)f + e] = out;
    }
}

__global__ __launch_bounds__(BLK) void k_parity_init(uint32_t nt, uint32_t* w) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f < nt) w[f] = f << 1;
}

__global__ __launch_bounds__(BLK) void k_parity_relax(const uint32_t* adj, uint32_t nt, uint32_t* w, uint32_t* changed) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    const uint32_t cur = uf_load(&w[f]);
    uint32_t best = cur;
    for (int e = 0; e < 3; e++) {
        const uint32_t a = adj[3 * (size_t)f + e];
        if (a == NONE) continue;
        const uint32_t g = a & 0x7fffffffu, same = a >> 31;
        best = min(best, uf_load(&w[g]) ^ same);
    }
    const uint32_t L = best >> 1;
    if (L != f) {
        const uint32_t wl = uf_load(&w[L]);
        best = min(best, (wl & ~1u) | ((wl ^ best) & 1u));
    }
    if (best < cur) {
        atomicMin(&w[f], best);
        *changed = 1u;
    }
}

// One atomic per wave when every active lane adds to the same slot.
__device__ inline void add_grouped(uint32_t* dst, uint32_t key, uint32_t val, bool active) {
    const unsigned long long act = __ballot(active);
    if (!act) return;
    const int leader = __ffsll((long long)act) - 1;
    const uint32_t k0 = __shfl(key, leader, 64);
    if (__ballot(active && key != k0) == 0) {
        uint32_t s = active ? val : 0u;
        for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
        if ((int)(threadIdx.x & 63) == leader && s) atomicAdd(&dst[k0], s);
    } else if (active && val) {
        atomicAdd(&dst[key], val);
    }
}

__global__ __launch_bounds__(BLK) void k_parity_count(const uint32_t* w, uint32_t nt, uint32_t* size,
                                                      uint32_t* nflip) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    const bool act = f < nt;
    const uint32_t x = act ? w[f] : 0u;
    add_grouped(size, x >> 1, 1u, act);
    add_grouped(nflip, x >> 1, x & 1u, act);
}

// Applies the flips (majority per component) and counts them.
__global__ __launch_bounds__(BLK) void k_parity_apply(const uint32_t* w, uint32_t nt, const uint32_t* size,
                                                      const uint32_t* nflip, uint32_t* tris, uint32_t* total) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    const uint32_t L = w[f] >> 1;
    const bool invert = 2ull * nflip[L] > (unsigned long long)size[L];
    if (((w[f] & 1u) != 0) != invert) {
        const uint32_t t1 = tris[3 * (size_t)f + 1];
        tris[3 * (size_t)f + 1] = tris[3 * (size_t)f + 2];
        tris[3 * (size_t)f + 2] = t1;
    }
    if (L == f) atomicAdd(total, invert ? size[L] - nflip[L] : nflip[L]);
}

// ---------------------------------------------------------------- shells
__global__ __launch_bounds__(BLK) void k_iota_u32(uint32_t* a, uint32_t n) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) a[i] = i;
}

__global__ __launch_bounds__(BLK) void k_shell_unite(const uint32_t* tris, uint32_t nt, uint32_t* parent,
                                                     uint32_t* used) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    const uint32_t a = tris[3 * (size_t)f], b = tris[3 * (size_t)f + 1], c = tris[3 * (size_t)f + 2];
    used[a] = 1u;
    used[b] = 1u;
    used[c] = 1u;
    uf_unite(parent, a, b);
    uf_unite(parent, a, c);
}

__global__ __launch_bounds__(BLK) void k_shell_roots(uint32_t* parent, const uint32_t* used, uint32_t nv,
                                                     uint32_t* is_root) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    uint32_t r = v;
    while (parent[r] != r) r = parent[r];
    parent[v] = r;
    is_root[v] = (used[v] && r == v) ? 1u : 0u;
}

struct ShellStats {
    uint32_t* ref;        // the shell's smallest vertex
    uint32_t* lohi;       // 6 per shell, ordered integers
    uint32_t* faces;
    double* vol;
    unsigned long long* boundary;
};

__device__ inline uint32_t ordf(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// The wave's lanes all on one key?  (leader lane, that key)
__device__ inline bool wave_uniform(uint32_t key, bool active, int* leader, uint32_t* k0) {
    const unsigned long long act = __ballot(active);
    if (!act) return false;
    *leader = __ffsll((long long)act) - 1;
    *k0 = __shfl(key, *leader, 64);
    return __ballot(active && key != *k0) == 0;
}

// comp[v] = shell id (first-occurrence order = order of the smallest vertex),
// bounds and boundary edges, one atomic per wave when its vertices share a shell.
__global__ __launch_bounds__(BLK) void k_shell_vertices(const float* pos, const uint32_t* parent, const uint32_t* used,
                                                        const uint32_t* sid, const uint32_t* vbound, uint32_t nv,
                                                        uint32_t* comp, ShellStats S) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    const bool act = v < nv && used[v];
    uint32_t c = 0, lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u}, vb = 0;
    if (v < nv) {
        if (act) {
            c = sid[parent[v]];
            comp[v] = c;
            if (parent[v] == v) S.ref[c] = v;
            for (int k = 0; k < 3; k++) lo[k] = hi[k] = ordf(pos[3 * (size_t)v + k]);
            vb = vbound[v];
        } else {
            comp[v] = NONE;
        }
    }
    int leader;
    uint32_t c0;
    if (wave_uniform(c, act, &leader, &c0)) {
        for (int d = 32; d >= 1; d >>= 1) {
            for (int k = 0; k < 3; k++) {
                lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], d, 64));
                hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], d, 64));
            }
            vb += (uint32_t)__shfl_xor((int)vb, d, 64);
        }
        if ((int)(threadIdx.x & 63) == leader) {
            for (int k = 0; k < 3; k++) {
                atomicMin(&S.lohi[6 * (size_t)c0 + k], lo[k]);
                atomicMax(&S.lohi[6 * (size_t)c0 + 3 + k], hi[k]);
            }
            if (vb) atomicAdd(&S.boundary[c0], (unsigned long long)vb);
        }
    } else if (act) {
        for (int k = 0; k < 3; k++) {
            atomicMin(&S.lohi[6 * (size_t)c + k], lo[k]);
            atomicMax(&S.lohi[6 * (size_t)c + 3 + k], hi[k]);
        }
        if (vb) atomicAdd(&S.boundary[c], (unsigned long long)vb);
    }
}

// normals.rs:653-661: faces and signed volume about the shell's reference vertex.
__global__ __launch_bounds__(BLK) void k_shell_faces(const float* pos, const uint32_t* tris, uint32_t nt,
                                                     const uint32_t* comp, ShellStats S) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    const bool act = f < nt;
    uint32_t c = 0;
    double vol = 0.0;
    if (act) {
        const uint32_t* t = tris + 3 * (size_t)f;
        c = comp[t[0]];
        const D3 o = d3(pos + 3 * (size_t)S.ref[c]);
        const D3 a = sub(d3(pos + 3 * (size_t)t[0]), o), b = sub(d3(pos + 3 * (size_t)t[1]), o),
                 d = sub(d3(pos + 3 * (size_t)t[2]), o);
        vol = dot(a, cross(b, d)) / 6.0;
    }
    int leader;
    uint32_t c0;
    if (wave_uniform(c, act, &leader, &c0)) {
        uint32_t n = act ? 1u : 0u;
        for (int d = 32; d >= 1; d >>= 1) {
            vol += __shfl_xor(vol, d, 64);
            n += (uint32_t)__shfl_xor((int)n, d, 64);
        }
        if ((int)(threadIdx.x & 63) == leader) {
            atomicAdd(&S.vol[c0], vol);
            atomicAdd(&S.faces[c0], n);
        }
    } else if (act) {
        atomicAdd(&S.vol[c], vol);
        atomicAdd(&S.faces[c], 1u);
    }
}

// normals.rs:761-785: f64 Moller-Trumbore, counted over the faces of the marked shells.
__global__ __launch_bounds__(BLK) void k_ray_crossings(const float* pos, const uint32_t* tris, uint32_t nt,
                                                       const uint32_t* comp, const uint8_t* mark, D3 p, D3 d,
                                                       uint32_t* count) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    bool hit = false;
    if (f < nt) {
        const uint32_t* t = tris + 3 * (size_t)f;
        if (mark[comp[t[0]]]) {
            const D3 a = d3(pos + 3 * (size_t)t[0]), b = d3(pos + 3 * (size_t)t[1]), c = d3(pos + 3 * (size_t)t[2]);
            const D3 e1 = sub(b, a), e2 = sub(c, a);
            const D3 pv = cross(d, e2);
            const double det = dot(e1, pv);
            if (fabs(det) >= 1e-300) {
                const double inv = 1.0 / det;
                const D3 tv = sub(p, a);
                const double u = dot(tv, pv) * inv;
                if (u >= 0.0 && u <= 1.0) {
                    const D3 q = cross(tv, e1);
                    const double w = dot(d, q) * inv;
                    if (!(w < 0.0 || u + w > 1.0)) hit = dot(e2, q) * inv > 0.0;
                }
            }
        }
    }
    const unsigned long long m = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(count, (uint32_t)__popcll(m));
}

__global__ __launch_bounds__(BLK) void k_flip_shells(uint32_t* tris, uint32_t nt, const uint32_t* comp,
                                                     const uint8_t* flip) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt || !flip[comp[tris[3 * (size_t)f]]]) return;
    const uint32_t t1 = tris[3 * (size_t)f + 1];
    tris[3 * (size_t)f + 1] = tris[3 * (size_t)f + 2];
    tris[3 * (size_t)f + 2] = t1;
}

__global__ __launch_bounds__(BLK) void k_negate_fans(uint32_t nv, const uint32_t* comp, const uint8_t* flip,
                                                     const uint32_t* off, const uint32_t* nfan, float* fan) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv || comp[v] == NONE || !flip[comp[v]]) return;
    for (uint32_t c = 0; c < nfan[v]; c++)
        for (int k = 0; k < 3; k++) fan[3 * (size_t)(off[v] + c) + k] = -fan[3 * (size_t)(off[v] + c) + k];
}

// ---------------------------------------------------------------- assembly (normals.rs:181-218)
__global__ __launch_bounds__(BLK) void k_extra(const uint32_t* nfan, uint32_t nv, uint32_t* extra) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v < nv) extra[v] = nfan[v] > 0 ? nfan[v] - 1 : 0;
}

__global__ __launch_bounds__(BLK) void k_asm_vertices(const float* pos, uint32_t nv, const uint32_t* off,
                                                      const uint32_t* nfan, const float* fan, const uint32_t* ebase,
                                                      float* opos, float* onrm) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    const float p[3] = {pos[3 * (size_t)v], pos[3 * (size_t)v + 1], pos[3 * (size_t)v + 2]};
    for (int k = 0; k < 3; k++) opos[3 * (size_t)v + k] = p[k];
    if (nfan[v] > 0) {
        for (int k = 0; k < 3; k++) onrm[3 * (size_t)v + k] = fan[3 * (size_t)off[v] + k];
    } else {
        onrm[3 * (size_t)v] = 0.0f;
        onrm[3 * (size_t)v + 1] = 0.0f;
        onrm[3 * (size_t)v + 2] = 1.0f;
    }
    for (uint32_t c = 1; c < nfan[v]; c++) {
        const size_t o = (size_t)nv + ebase[v] + c - 1;
        for (int k = 0; k < 3; k++) {
            opos[3 * o + k] = p[k];
            onrm[3 * o + k] = fan[3 * (size_t)(off[v] + c) + k];
        }
    }
}

__global__ __launch_bounds__(BLK) void k_asm_tris(uint32_t* tris, uint32_t nt, uint32_t nv, const uint32_t* off,
                                                  const uint32_t* faces, const uint16_t* label, const uint32_t* ebase) {
    const uint32_t f = blockIdx.x * BLK + threadIdx.x;
    if (f >= nt) return;
    for (int k = 0; k < 3; k++) {
        const uint32_t v = tris[3 * (size_t)f + k];
        uint32_t lo = off[v], hi = off[v + 1];  // the face's slot among v's sorted faces
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (faces[mid] < f) lo = mid + 1;
            else hi = mid;
        }
        const uint32_t lab = label[lo];
        tris[3 * (size_t)f + k] = lab == 0 ? v : nv + ebase[v] + lab - 1;
    }
}

__global__ __launch_bounds__(BLK) void k_fill_up(float* nrm, uint32_t nv) {
    const uint32_t v = blockIdx.x * BLK + threadIdx.x;
    if (v >= nv) return;
    nrm[3 * (size_t)v] = 0.0f;
    nrm[3 * (size_t)v + 1] = 0.0f;
    nrm[3 * (size_t)v + 2] = 1.0f;
}

float from_ordf(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// normals.rs:94-99
void thresholds(float deg, double* cj, double* cm) {
    const float d = std::min(std::max(deg, 0.05f), 180.0f);
    const double pi = 3.14159265358979323846;
    *cj = std::cos((double)d * (pi / 180.0));
    const float h = std::min(std::max(0.5f * d, 0.05f), 15.0f);
    *cm = std::cos((double)h * (pi / 180.0));
}

// normals.rs:674-757 on the host, from the shells' device statistics.
int decide_flips(ag_mesh* m, hipStream_t st, const uint32_t* comp_dev, uint32_t ns, const ShellStats& S,
                 uint8_t* mark_dev, uint32_t* count_dev, std::vector<uint8_t>* flip_out, ag_normal_report* rep) {
    std::vector<uint8_t>& flip = *flip_out;
    flip.assign(ns, 0);
    if (ns == 0) return AG_OK;
    std::vector<uint32_t> ref(ns), lohi(6 * (size_t)ns), faces(ns);
    std::vector<double> vol(ns);
    std::vector<unsigned long long> boundary(ns);
    MCHK(hipMemcpyAsync(ref.data(), S.ref, 4 * (size_t)ns, hipMemcpyDeviceToHost, st));
    MCHK(hipMemcpyAsync(lohi.data(), S.lohi, 24 * (size_t)ns, hipMemcpyDeviceToHost, st));
    MCHK(hipMemcpyAsync(faces.data(), S.faces, 4 * (size_t)ns, hipMemcpyDeviceToHost, st));
    MCHK(hipMemcpyAsync(vol.data(), S.vol, 8 * (size_t)ns, hipMemcpyDeviceToHost, st));
    MCHK(hipMemcpyAsync(boundary.data(), S.boundary, 8 * (size_t)ns, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    auto closed = [&](uint32_t c) { return (double)boundary[c] <= 0.01 * 1.5 * (double)faces[c]; };
    std::vector<uint32_t> cand;
    for (uint32_t c = 0; c < ns; c++)
        if (closed(c) && vol[c] < 0.0) cand.push_back(c);
    if (cand.empty()) return AG_OK;
    if (cand.size() > 256) {
        double total = 0.0;
        for (uint32_t c = 0; c < ns; c++)
            if (closed(c)) total += vol[c];
        if (total < 0.0)
            for (uint32_t c : cand) flip[c] = 1;
        else
            rep->cavity_shells += cand.size();
        for (auto f : flip) rep->flipped_shells += f;
        return AG_OK;
    }
    auto lo = [&](uint32_t c, int k) { return from_ordf(lohi[6 * (size_t)c + k]); };
    auto hi = [&](uint32_t c, int k) { return from_ordf(lohi[6 * (size_t)c + 3 + k]); };
    auto contains = [&](uint32_t o, uint32_t i) {
        if (o == i) return false;
        for (int k = 0; k < 3; k++)
            if (!(lo(o, k) <= lo(i, k) && hi(o, k) >= hi(i, k))) return false;
        return true;
    };
    uint64_t budget = 300000000;
    std::vector<uint8_t> mark(ns);
    for (uint32_t c : cand) {
        std::fill(mark.begin(), mark.end(), 0);
        uint64_t cost = 0;
        bool any = false;
        for (uint32_t o = 0; o < ns; o++)
            if (closed(o) && contains(o, c)) {
                mark[o] = 1;
                cost += faces[o];
                any = true;
            }
        if (!any) {
            flip[c] = 1;
            continue;
        }
        if (cost > budget) {
            rep->cavity_shells++;
            continue;
        }
        budget -= cost;
        float pr[3];
        MCHK(hipMemcpyAsync(pr, m->pos.as<float>() + 3 * (size_t)ref[c], 12, hipMemcpyDeviceToHost, st));
        MCHK(hipMemcpyAsync(mark_dev, mark.data(), ns, hipMemcpyHostToDevice, st));
        MCHK(hipMemsetAsync(count_dev, 0, 4, st));
        const D3 p = {(double)pr[0], (double)pr[1], (double)pr[2]};
        const D3 d = {0.5773502691896258, 0.5773519, 0.5773486};  // normals.rs:736 (f64 literals)
        k_ray_crossings<<<grid_of(m->nt), BLK, 0, st>>>(m->pos.as<float>(), m->tris.as<uint32_t>(), m->nt, comp_dev,
                                                         mark_dev, p, d, count_dev);
        MCHK(hipGetLastError());
        uint32_t crossings = 0;
        MCHK(hipMemcpyAsync(&crossings, count_dev, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
        if (crossings % 2 == 1) rep->cavity_shells++;
        else flip[c] = 1;
    }
    for (auto f : flip) rep->flipped_shells += f;
    return AG_OK;
}

// normals.rs:120-220 on a device mesh.
int normals_mesh(ag_mesh* m, float crease_deg, ag_normal_report* rep) {
    hipStream_t st = ctx_stream(m->ctx);
    *rep = ag_normal_report{};
    uint32_t dropped = 0;
    int rc = mesh_drop_collapsed(m, &dropped);
    if (rc) return rc;
    const uint32_t nv = m->nv, nt = m->nt;
    if (nt == 0) {
        if ((rc = m->nrm.alloc(12 * (size_t)std::max(nv, 1u)))) return rc;
        k_fill_up<<<grid_of(nv), BLK, 0, st>>>(m->nrm.as<float>(), nv);
        MCHK(hipGetLastError());
        MCHK(hipStreamSynchronize(st));
        m->has_nrm = true;
        return AG_OK;
    }
    double cj, cm;
    thresholds(crease_deg, &cj, &cm);
    const size_t I = 3 * (size_t)nt;  // vertex-face incidences
    unsigned vbits = 1;
    while (vbits < 32 && (1ull << vbits) <= nv) vbits++;
    size_t sort_bytes = 0, seg_bytes = 0;
    MCHK(rocprim::radix_sort_pairs(nullptr, sort_bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, (uint32_t*)nullptr, (unsigned)I, 0, vbits, st));
    MCHK(rocprim::segmented_radix_sort_keys(nullptr, seg_bytes, (unsigned long long*)nullptr,
                                            (unsigned long long*)nullptr, (unsigned)(2 * I), nv, (uint32_t*)nullptr,
                                            (uint32_t*)nullptr, 0, 64, st));
    sort_bytes = std::max(sort_bytes, seg_bytes);
    Arena a;
    const size_t o_cnt = a.add(4 * ((size_t)nv + 1)), o_off = a.add(4 * ((size_t)nv + 1)),
                 o_cur = a.add(4 * I), o_faces = a.add(4 * I), o_fn = a.add(12 * (size_t)nt),
                 o_ent = a.add(16 * I), o_ent_in = a.add(16 * I), o_keys2 = a.add(4 * I),
                 o_off2 = a.add(4 * ((size_t)nv + 1)), o_sort = a.add(sort_bytes), o_ang = a.add(8 * I), o_par = a.add(2 * I), o_lab = a.add(2 * I),
                 o_rm = a.add(2 * I), o_sums = a.add(32 * I), o_fan = a.add(12 * I), o_nfan = a.add(4 * (size_t)nv),
                 o_vb = a.add(4 * (size_t)nv), o_flag = a.add(256),
                 o_work = a.add(4 * device_scan_work_words(std::max<size_t>((size_t)nv + 1, nt))),
                 o_adj = a.add(12 * (size_t)nt), o_w = a.add(4 * (size_t)nt), o_size = a.add(4 * (size_t)nt),
                 o_nflip = a.add(4 * (size_t)nt), o_parent = a.add(4 * (size_t)nv), o_used = a.add(4 * (size_t)nv),
                 o_root = a.add(4 * (size_t)nv), o_sid = a.add(4 * (size_t)nv), o_comp = a.add(4 * (size_t)nv),
                 o_ref = a.add(4 * (size_t)nv), o_lohi = a.add(24 * (size_t)nv), o_sfaces = a.add(4 * (size_t)nv),
                 o_vol = a.add(8 * (size_t)nv), o_sbound = a.add(8 * (size_t)nv), o_mark = a.add(nv),
                 o_extra = a.add(4 * (size_t)nv), o_ebase = a.add(4 * (size_t)nv);
    if ((rc = a.bind(m->ctx))) return rc;
    uint32_t* flag = a.at<uint32_t>(o_flag);  // [0] degenerate, [1] scan total, [2] changed, [3] repaired, [4] crossings
    auto* counts = reinterpret_cast<unsigned long long*>(flag + 8);  // boundary, non-manifold, inconsistent
    uint32_t* off = a.at<uint32_t>(o_off);
    uint32_t* faces = a.at<uint32_t>(o_faces);
    uint32_t* work = a.at<uint32_t>(o_work);
    float* fn = a.at<float>(o_fn);
    uint32_t* tris = m->tris.as<uint32_t>();
    const float* pos = m->pos.as<float>();
    MCHK(hipMemsetAsync(flag, 0, 256, st));

    // vertex -> faces, each list in increasing face order (util.rs:96-128)
    MCHK(hipMemsetAsync(a.at<uint32_t>(o_cnt), 0, 4 * ((size_t)nv + 1), st));
    k_vf_count<<<grid_of(nt), BLK, 0, st>>>(tris, nt, a.at<uint32_t>(o_cnt));
    device_scan_u32(st, a.at<uint32_t>(o_cnt), off, nv + 1, work, flag + 1);
    k_incidence_faces<<<grid_of(I), BLK, 0, st>>>(nt, a.at<uint32_t>(o_cur));
    MCHK(rocprim::radix_sort_pairs(a.at<void>(o_sort), sort_bytes, tris, a.at<uint32_t>(o_keys2),
                                   a.at<uint32_t>(o_cur), faces, (unsigned)I, 0, vbits, st));
    k_face_normals<<<grid_of(nt), BLK, 0, st>>>(pos, tris, nt, fn, flag);
    Cluster C{pos, tris, fn, off, faces, a.at<unsigned long long>(o_ent), a.at<double>(o_ang), a.at<uint16_t>(o_par),
              a.at<uint16_t>(o_lab), a.at<uint16_t>(o_rm), a.at<double>(o_sums), a.at<float>(o_fan),
              a.at<uint32_t>(o_nfan), a.at<uint32_t>(o_vb), counts, cj, cm};
    auto fans = [&]() -> int {
        k_fan_prep<<<grid_of(nv), BLK, 0, st>>>(C, nv, a.at<unsigned long long>(o_ent_in));
        k_double_offsets<<<grid_of((size_t)nv + 1), BLK, 0, st>>>(off, nv, a.at<uint32_t>(o_off2));
        MCHK(rocprim::segmented_radix_sort_keys(a.at<void>(o_sort), sort_bytes, a.at<unsigned long long>(o_ent_in),
                                                a.at<unsigned long long>(o_ent), (unsigned)(2 * I), nv,
                                                a.at<uint32_t>(o_off2), a.at<uint32_t>(o_off2) + 1, 0, 64, st));
        k_cluster<<<grid_of(nv), BLK, 0, st>>>(C, nv);
        MCHK(hipGetLastError());
        return AG_OK;
    };
    if ((rc = fans())) return rc;
    unsigned long long hc[3];
    uint32_t hdeg = 0;
    MCHK(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, st));
    MCHK(hipMemcpyAsync(&hdeg, flag, 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    rep->degenerate_faces = hdeg;
    if (hc[2] > 0) {
        // winding repair (normals.rs:141-148)
        uint32_t* adj = a.at<uint32_t>(o_adj);
        uint32_t* w = a.at<uint32_t>(o_w);
        k_face_adj<<<grid_of(nt), BLK, 0, st>>>(tris, nt, off, faces, adj);
        k_parity_init<<<grid_of(nt), BLK, 0, st>>>(nt, w);
        for (int it = 0;; it++) {
            if (it > 100000) return fail(AG_E_DEVICE, "winding repair did not converge");
            MCHK(hipMemsetAsync(flag + 2, 0, 4, st));
            k_parity_relax<<<grid_of(nt), BLK, 0, st>>>(adj, nt, w, flag + 2);
            uint32_t changed = 0;
            MCHK(hipMemcpyAsync(&changed, flag + 2, 4, hipMemcpyDeviceToHost, st));
            MCHK(hipStreamSynchronize(st));
            if (!changed) break;
        }
        MCHK(hipMemsetAsync(a.at<uint32_t>(o_size), 0, 4 * (size_t)nt, st));
        MCHK(hipMemsetAsync(a.at<uint32_t>(o_nflip), 0, 4 * (size_t)nt, st));
        k_parity_count<<<grid_of(nt), BLK, 0, st>>>(w, nt, a.at<uint32_t>(o_size), a.at<uint32_t>(o_nflip));
        k_parity_apply<<<grid_of(nt), BLK, 0, st>>>(w, nt, a.at<uint32_t>(o_size), a.at<uint32_t>(o_nflip), tris,
                                                    flag + 3);
        MCHK(hipGetLastError());
        uint32_t flipped = 0;
        MCHK(hipMemcpyAsync(&flipped, flag + 3, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
        if (flipped > 0) {
            rep->repaired_faces = flipped;
            k_face_normals<<<grid_of(nt), BLK, 0, st>>>(pos, tris, nt, fn, nullptr);
            MCHK(hipMemsetAsync(counts, 0, sizeof hc, st));
            if ((rc = fans())) return rc;
            MCHK(hipMemcpyAsync(hc, counts, sizeof hc, hipMemcpyDeviceToHost, st));
            MCHK(hipStreamSynchronize(st));
        }
    }
    rep->boundary_edges = hc[0];
    rep->nonmanifold_edges = hc[1];
    rep->inconsistent_edges = hc[2];

    // shells and inside-out correction (normals.rs:153-179)
    uint32_t* parent = a.at<uint32_t>(o_parent);
    uint32_t* used = a.at<uint32_t>(o_used);
    uint32_t* comp = a.at<uint32_t>(o_comp);
    MCHK(hipMemsetAsync(used, 0, 4 * (size_t)nv, st));
    k_iota_u32<<<grid_of(nv), BLK, 0, st>>>(parent, nv);
    k_shell_unite<<<grid_of(nt), BLK, 0, st>>>(tris, nt, parent, used);
    k_shell_roots<<<grid_of(nv), BLK, 0, st>>>(parent, used, nv, a.at<uint32_t>(o_root));
    device_scan_u32(st, a.at<uint32_t>(o_root), a.at<uint32_t>(o_sid), nv, work, flag + 1);
    uint32_t ns = 0;
    MCHK(hipMemcpyAsync(&ns, flag + 1, 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    rep->shells = ns;
    ShellStats S{a.at<uint32_t>(o_ref), a.at<uint32_t>(o_lohi), a.at<uint32_t>(o_sfaces), a.at<double>(o_vol),
                 a.at<unsigned long long>(o_sbound)};
    if (ns > 0) {
        std::vector<uint32_t> init(6 * (size_t)ns);
        for (size_t c = 0; c < ns; c++)
            for (int k = 0; k < 3; k++) {
                init[6 * c + k] = 0xFFFFFFFFu;
                init[6 * c + 3 + k] = 0u;
            }
        MCHK(hipMemcpyAsync(S.lohi, init.data(), 24 * (size_t)ns, hipMemcpyHostToDevice, st));
        MCHK(hipMemsetAsync(S.faces, 0, 4 * (size_t)ns, st));
        MCHK(hipMemsetAsync(S.vol, 0, 8 * (size_t)ns, st));
        MCHK(hipMemsetAsync(S.boundary, 0, 8 * (size_t)ns, st));
        k_shell_vertices<<<grid_of(nv), BLK, 0, st>>>(pos, parent, used, a.at<uint32_t>(o_sid), a.at<uint32_t>(o_vb),
                                                      nv, comp, S);
        k_shell_faces<<<grid_of(nt), BLK, 0, st>>>(pos, tris, nt, comp, S);
        MCHK(hipGetLastError());
        std::vector<uint8_t> flip;
        if ((rc = decide_flips(m, st, comp, ns, S, a.at<uint8_t>(o_mark), flag + 4, &flip, rep))) return rc;
        bool any = false;
        for (auto f : flip) any = any || f;
        if (any) {
            MCHK(hipMemcpyAsync(a.at<uint8_t>(o_mark), flip.data(), ns, hipMemcpyHostToDevice, st));
            k_flip_shells<<<grid_of(nt), BLK, 0, st>>>(tris, nt, comp, a.at<uint8_t>(o_mark));
            k_negate_fans<<<grid_of(nv), BLK, 0, st>>>(nv, comp, a.at<uint8_t>(o_mark), off, a.at<uint32_t>(o_nfan),
                                                       a.at<float>(o_fan));
            MCHK(hipGetLastError());
        }
    }

    // split vertices (normals.rs:181-218)
    k_extra<<<grid_of(nv), BLK, 0, st>>>(a.at<uint32_t>(o_nfan), nv, a.at<uint32_t>(o_extra));
    device_scan_u32(st, a.at<uint32_t>(o_extra), a.at<uint32_t>(o_ebase), nv, work, flag + 1);
    uint32_t split = 0;
    MCHK(hipMemcpyAsync(&split, flag + 1, 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if ((uint64_t)nv + split > 0x7fffffffull) return fail(AG_E_INVALID, "too many vertices after splitting");
    rep->split_vertices = split;
    Buf npos, nnrm;
    if ((rc = npos.alloc(12 * ((size_t)nv + split))) || (rc = nnrm.alloc(12 * ((size_t)nv + split)))) return rc;
    k_asm_vertices<<<grid_of(nv), BLK, 0, st>>>(pos, nv, off, a.at<uint32_t>(o_nfan), a.at<float>(o_fan),
                                                a.at<uint32_t>(o_ebase), npos.as<float>(), nnrm.as<float>());
    k_asm_tris<<<grid_of(nt), BLK, 0, st>>>(tris, nt, nv, off, faces, a.at<uint16_t>(o_lab), a.at<uint32_t>(o_ebase));
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    std::swap(m->pos.p, npos.p);
    std::swap(m->pos.bytes, npos.bytes);
    std::swap(m->nrm.p, nnrm.p);
    std::swap(m->nrm.bytes, nnrm.bytes);
    m->nv = nv + split;
    m->has_nrm = true;
    return AG_OK;
}

}  // namespace
}  // namespace ag

using namespace ag;

extern "C" int ag_mesh_normals(ag_mesh* m, float crease_angle_deg, ag_normal_report* report) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    ag_normal_report r{};
    const int rc = normals_mesh(m, crease_angle_deg, &r);
    if (rc) return rc;
    if (report) *report = r;
    return AG_OK;
}

extern "C" int ag_mesh_read_normals(const ag_mesh* m, float* normals) {
    if (!m || !normals) return fail(AG_E_INVALID, "null argument");
    if (!m->has_nrm) return fail(AG_E_STATE, "the mesh has no normals (ag_mesh_normals)");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    hipStream_t st = ctx_stream(m->ctx);
    if (m->nv) MCHK(hipMemcpyAsync(normals, m->nrm.p, 12 * (size_t)m->nv, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    return AG_OK;
}
