// Mesh ingest on the device: binary STL records to a triangle soup
// (argent-mesh io/stl.rs:34-62), the invalid-triangle filter (io.rs:206-224),
// vertex welding (weld.rs:38-330) and compaction (weld.rs:66-91), with the
// C entry points ag_mesh_*.
//
// Welding keeps the reference's results exactly while replacing its sorts:
//   exact stage   an open-addressing table keyed by the canonical coordinate
//                 bits; a slot holds a vertex index and every vertex of the
//                 slot's key lowers it with atomicMin, so after one pass the
//                 slot holds the smallest index of the key (the reference's
//                 run head, weld.rs:194-206).  An exclusive scan of the
//                 "is head" flags numbers the classes in first-occurrence
//                 order (weld.rs:210-226).
//   tolerance     the survivors go into a hash of grid cells (cell size and
//                 packing as weld.rs:230-251); each point probes its own cell
//                 and the neighbours it is within eps of (weld.rs:258-300)
//                 and unites with every earlier point within eps, in f64.
//                 The union-find hooks the larger root under the smaller
//                 (lock-free, atomicCAS on roots), so each component's root
//                 is its smallest member whatever the order of unions, the
//                 same root as util.rs:58-87 gives.  Scanned root flags give
//                 the first-occurrence numbering (weld.rs:307-321).
//   triangles     remapped; those with a repeated index dropped by a scan
//                 (weld.rs:104-112, util.rs:132-164).
// Every step is a pure function of its input: no result depends on thread
// timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mesh_internal.hpp"

namespace ag {
namespace {

constexpr unsigned long long CELL_EMPTY = ~0ull;
constexpr int SCAN_ITEMS = 8;
constexpr uint32_t SCAN_TILE = BLK * SCAN_ITEMS;

// ------------------------------------------------------------------ scan
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[BLK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int i = 0; i < BLK / 64; i++) {
        uint32_t s = wsum[i];
        off += i < w ? s : 0u;
        tot += s;
    }
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(BLK) void k_scan_tiles(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* sums) {
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = base + k < n ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_exclusive_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(BLK) void k_scan_add(uint32_t* out, uint32_t n, const uint32_t* offs) {
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    const uint32_t a = offs[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
        if (base + k < n) out[base + k] += a;
}

__global__ void k_copy_word(const uint32_t* src, uint32_t* dst) { *dst = *src; }

size_t tiles_of(size_t n) { return std::max<size_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE); }

// Words of work space scan() needs for n elements.
size_t scan_work_words(size_t n) {
    size_t w = 0;
    for (size_t t = tiles_of(n);; t = tiles_of(t)) {
        w += 2 * t;
        if (t == 1) break;
    }
    return w + 1;
}

// Exclusive scan of n words (in != out); *total (device) gets the sum.
void scan(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* work, uint32_t* total) {
    const uint32_t t = (uint32_t)tiles_of(n);
    uint32_t* sums = work;
    uint32_t* offs = work + t;
    k_scan_tiles<<<t, BLK, 0, st>>>(in, out, n, sums);
    if (t == 1) {
        k_copy_word<<<1, 1, 0, st>>>(sums, total);
        return;
    }
    scan(st, sums, offs, t, work + 2 * t, total);
    k_scan_add<<<t, BLK, 0, st>>>(out, n, offs);
}

// ------------------------------------------------------------------ STL
// One thread per 50-byte record: the 36 coordinate bytes start at 84 + 50 t
// + 12, 4-byte aligned for even t and 2 bytes past it for odd t, so the
// thread reads the 10 aligned words that cover them and shifts.
__global__ __launch_bounds__(BLK) void k_stl_parse(const uint8_t* bytes, uint32_t n, float scale, int do_scale,
                                                   float* pos, uint32_t* tris, uint32_t* keep, uint32_t* n_bad) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    bool bad = false;
    if (t < n) {
        const size_t start = 84 + 50 * (size_t)t + 12;
        const uint32_t* w = reinterpret_cast<const uint32_t*>(bytes + (start & ~(size_t)3));
        const uint32_t sh = (uint32_t)(start & 3) * 8;  // 0 or 16
        uint32_t word[10];
#pragma unroll
        for (int k = 0; k < 10; k++) word[k] = (k < 9 || sh) ? w[k] : 0u;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const uint32_t b = sh ? (word[k] >> 16) | (word[k + 1] << 16) : word[k];
            float f = __uint_as_float(b);
            bad |= (b & 0x7f800000u) == 0x7f800000u;  // io.rs:208-209: non-finite
            if (do_scale) f *= scale;                 // io.rs:105-108, after the filter
            pos[9 * (size_t)t + k] = f;
        }
#pragma unroll
        for (int k = 0; k < 3; k++) tris[3 * (size_t)t + k] = 3 * t + k;
        keep[t] = bad ? 0u : 1u;
    }
    const unsigned long long m = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(n_bad, (uint32_t)__popcll(m));
}

// Compacts triangles by their scanned keep flags.
__global__ __launch_bounds__(BLK) void k_compact_tris(const uint32_t* tris, const uint32_t* keep, const uint32_t* idx,
                                                      uint32_t nt, uint32_t* out) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt || !keep[t]) return;
    const uint32_t o = idx[t];
#pragma unroll
    for (int k = 0; k < 3; k++) out[3 * (size_t)o + k] = tris[3 * (size_t)t + k];
}

// ------------------------------------------------------------------ weld
__device__ inline uint32_t canon_bits(float f) {
    // weld.rs:150-154, `(x + 0.0).to_bits()` on x86-64: -0 becomes +0, a
    // signalling NaN is quieted, everything else (denormals too) is unchanged.
    uint32_t b = __float_as_uint(f);
    if (b == 0x80000000u) return 0u;
    if ((b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu)) return b | 0x00400000u;
    return b;
}

__device__ inline uint32_t mix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ inline uint32_t key_hash(uint32_t x, uint32_t y, uint32_t z) {
    return mix32(x ^ mix32(y ^ mix32(z + 0x9e3779b9u)));
}
__device__ inline uint32_t cell_hash(unsigned long long k) {
    return mix32((uint32_t)k ^ mix32((uint32_t)(k >> 32) + 0x9e3779b9u));
}

__global__ __launch_bounds__(BLK) void k_mark_used(const uint32_t* tris, uint32_t nt, uint32_t nv, uint32_t* used,
                                                   uint32_t* oob) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt) return;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const uint32_t v = tris[3 * (size_t)t + k];
        if (v < nv) used[v] = 1u;
        else *oob = 1u;
    }
}

// Ordered-integer image of a float: larger float, larger integer (-0 < +0).
__device__ inline uint32_t ord(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// weld.rs:126-148: min and max of the referenced positions per axis,
// starting from +-f32::MAX; NaN coordinates are skipped as f32::min/max do.
// The blocks' partial results are folded by k_bounds_final (atomics from
// every wave on six shared words would serialize).
constexpr unsigned BOUNDS_BLOCKS = 1024;

__device__ inline void block_reduce(uint32_t lo[3], uint32_t hi[3], uint32_t cnt, uint32_t* out7) {
    __shared__ uint32_t red[BLK / 64][7];
    for (int d = 32; d >= 1; d >>= 1) {
        for (int k = 0; k < 3; k++) {
            lo[k] = min(lo[k], (uint32_t)__shfl_xor((int)lo[k], d, 64));
            hi[k] = max(hi[k], (uint32_t)__shfl_xor((int)hi[k], d, 64));
        }
        cnt += (uint32_t)__shfl_xor((int)cnt, d, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        for (int k = 0; k < 3; k++) {
            red[w][k] = lo[k];
            red[w][3 + k] = hi[k];
        }
        red[w][6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x < 7) {
        const int j = threadIdx.x;
        uint32_t v = red[0][j];
        for (int q = 1; q < BLK / 64; q++) v = j < 3 ? min(v, red[q][j]) : (j < 6 ? max(v, red[q][j]) : v + red[q][j]);
        out7[j] = v;
    }
}

// Each block writes its six ordered-integer extremes and its count of
// referenced vertices to part[7 * block].
__global__ __launch_bounds__(BLK) void k_bounds(const float* pos, const uint32_t* used, uint32_t nv, uint32_t* part) {
    uint32_t lo[3], hi[3], cnt = 0;
    for (int k = 0; k < 3; k++) {
        lo[k] = 0xFFFFFFFFu;
        hi[k] = 0u;
    }
    for (size_t i = blockIdx.x * (size_t)BLK + threadIdx.x; i < nv; i += (size_t)gridDim.x * BLK) {
        if (!used[i]) continue;
        cnt++;
        for (int k = 0; k < 3; k++) {
            const float f = pos[3 * i + k];
            if (f != f) continue;
            lo[k] = min(lo[k], ord(f));
            hi[k] = max(hi[k], ord(f));
        }
    }
    block_reduce(lo, hi, cnt, part + 7 * blockIdx.x);
}

struct Init6 {
    uint32_t v[6];  // ordered images of +f32::MAX (x3) and -f32::MAX (x3)
};

// Folds nb block results with the start values: out = lo[3], hi[3], count.
__global__ __launch_bounds__(BLK) void k_bounds_final(const uint32_t* part, uint32_t nb, Init6 init, uint32_t* out) {
    uint32_t lo[3], hi[3], cnt = 0;
    for (int k = 0; k < 3; k++) {
        lo[k] = 0xFFFFFFFFu;
        hi[k] = 0u;
    }
    for (uint32_t b = threadIdx.x; b < nb; b += BLK) {
        for (int k = 0; k < 3; k++) {
            lo[k] = min(lo[k], part[7 * b + k]);
            hi[k] = max(hi[k], part[7 * b + 3 + k]);
        }
        cnt += part[7 * b + 6];
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) {
            lo[k] = min(lo[k], init.v[k]);
            hi[k] = max(hi[k], init.v[3 + k]);
        }
    block_reduce(lo, hi, cnt, out);
}

__global__ __launch_bounds__(BLK) void k_weld_insert(const float* pos, const uint32_t* used, uint32_t nv,
                                                     uint32_t* table, uint32_t mask) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv || !used[i]) return;
    const uint32_t x = canon_bits(pos[3 * (size_t)i]), y = canon_bits(pos[3 * (size_t)i + 1]),
                   z = canon_bits(pos[3 * (size_t)i + 2]);
    uint32_t s = key_hash(x, y, z) & mask;
    for (;;) {
        uint32_t cur = __hip_atomic_load(&table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (cur == NONE) {
            cur = atomicCAS(&table[s], NONE, i);
            if (cur == NONE) return;
        }
        // A slot's key never changes: every index stored in it has that key,
        // and its value only decreases, so a smaller value seen once is final
        // enough to skip the atomic.
        if (canon_bits(pos[3 * (size_t)cur]) == x && canon_bits(pos[3 * (size_t)cur + 1]) == y &&
            canon_bits(pos[3 * (size_t)cur + 2]) == z) {
            if (i < cur) atomicMin(&table[s], i);
            return;
        }
        s = (s + 1) & mask;
    }
}

// rep = the smallest index with the same key; head = (rep == i).
__global__ __launch_bounds__(BLK) void k_weld_find(const float* pos, const uint32_t* used, uint32_t nv,
                                                   const uint32_t* table, uint32_t mask, uint32_t* rep,
                                                   uint32_t* head) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv) return;
    if (!used[i]) {
        rep[i] = NONE;
        head[i] = 0u;
        return;
    }
    const uint32_t x = canon_bits(pos[3 * (size_t)i]), y = canon_bits(pos[3 * (size_t)i + 1]),
                   z = canon_bits(pos[3 * (size_t)i + 2]);
    uint32_t s = key_hash(x, y, z) & mask;
    for (;;) {
        const uint32_t cur = table[s];
        if (cur == NONE) {  // unreachable: i's key sits before the first empty slot of its probe run
            rep[i] = i;
            head[i] = 1u;
            return;
        }
        if (canon_bits(pos[3 * (size_t)cur]) == x && canon_bits(pos[3 * (size_t)cur + 1]) == y &&
            canon_bits(pos[3 * (size_t)cur + 2]) == z) {
            rep[i] = cur;
            head[i] = cur == i ? 1u : 0u;
            return;
        }
        s = (s + 1) & mask;
    }
}

// uid[i] = class number of i (first-occurrence order); uniq[uid] = the head's position.
__global__ __launch_bounds__(BLK) void k_weld_number(const float* pos, const uint32_t* rep, const uint32_t* hidx,
                                                     uint32_t nv, uint32_t* uid, float* uniq) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv) return;
    const uint32_t r = rep[i];
    if (r == NONE) {
        uid[i] = NONE;
        return;
    }
    const uint32_t u = hidx[r];
    uid[i] = u;
    if (r == i)
        for (int k = 0; k < 3; k++) uniq[3 * (size_t)u + k] = pos[3 * (size_t)i + k];
}

struct Grid {
    double lo[3], h, eps, eps2;
};

// Rust's `f64 as i64`: saturating, NaN -> 0.
__device__ inline long long rust_i64(double x) {
    if (x != x) return 0;
    if (x >= 9223372036854775807.0) return 0x7fffffffffffffffll;
    if (x <= -9223372036854775808.0) return (long long)0x8000000000000000ull;
    return (long long)x;
}
__device__ inline unsigned long long cell_key(const float* p, const Grid& g) {
    const long long cx = rust_i64(floor(((double)p[0] - g.lo[0]) / g.h));
    const long long cy = rust_i64(floor(((double)p[1] - g.lo[1]) / g.h));
    const long long cz = rust_i64(floor(((double)p[2] - g.lo[2]) / g.h));
    return ((unsigned long long)cx << 42) | ((unsigned long long)cy << 21) | (unsigned long long)cz;  // weld.rs:247
}

__global__ __launch_bounds__(BLK) void k_cell_insert(const float* uniq, uint32_t m, Grid g, unsigned long long* keys,
                                                     uint32_t* cnt, uint32_t* slot_of, uint32_t mask) {
    const uint32_t u = blockIdx.x * BLK + threadIdx.x;
    if (u >= m) return;
    const unsigned long long key = cell_key(uniq + 3 * (size_t)u, g);
    uint32_t s = cell_hash(key) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&keys[s], CELL_EMPTY, key);
        if (prev == CELL_EMPTY || prev == key) {
            atomicAdd(&cnt[s], 1u);
            slot_of[u] = s;
            return;
        }
        s = (s + 1) & mask;
    }
}

// members[start[s] .. start[s + 1]) = the points of slot s (any order).
__global__ __launch_bounds__(BLK) void k_cell_scatter(const uint32_t* slot_of, uint32_t m, const uint32_t* start,
                                                      uint32_t* cnt, uint32_t* members) {
    const uint32_t u = blockIdx.x * BLK + threadIdx.x;
    if (u >= m) return;
    const uint32_t s = slot_of[u];
    members[start[s] + atomicSub(&cnt[s], 1u) - 1u] = u;
}

__global__ __launch_bounds__(BLK) void k_fill_ones(uint32_t* a, uint32_t n) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) a[i] = 1u;
}

// out[i] = vertex i * stride, for i < n (i * stride < the vertex count)
__global__ __launch_bounds__(BLK) void k_gather_stride(const float* pos, uint32_t stride, uint32_t n, float* out) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= n) return;
    const size_t s = (size_t)i * stride * 3, d = (size_t)i * 3;
    out[d] = pos[s];
    out[d + 1] = pos[s + 1];
    out[d + 2] = pos[s + 2];
}

__global__ __launch_bounds__(BLK) void k_iota(uint32_t* a, uint32_t n) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i < n) a[i] = i;
}

// Cell coordinate on one axis, and the neighbour it must also probe:
// -1 / +1 when the point lies within eps of the low / high wall, else 0
// (weld.rs:263-273).
__device__ inline long long axis_cell(float p, double lo, const Grid& g, int* extra) {
    const long long c = rust_i64(floor(((double)p - lo) / g.h));
    const double f = (double)p - (lo + (double)c * g.h);
    *extra = 0;
    if (f < g.eps && c > 0) *extra = -1;
    else if (g.h - f < g.eps && c < (1 << 21) - 1) *extra = 1;
    return c;
}

// weld.rs:252-300: calls f(j, extra[j]) for every earlier point j < i within
// eps of point i, in its cell and in the neighbour cells i lies within eps
// of.  Members are read 8 at a time with their loads issued together (one
// cell can hold hundreds of nearly coincident points).  Scalars only (no
// indexed arrays left after unrolling), so nothing spills to scratch.
template <class F>
__device__ void for_each_near(const float* uniq, uint32_t i, const Grid& g, const unsigned long long* keys,
                              const uint32_t* start, const uint32_t* members, uint32_t mask, const uint32_t* extra,
                              F&& f) {
    const float px = uniq[3 * (size_t)i], py = uniq[3 * (size_t)i + 1], pz = uniq[3 * (size_t)i + 2];
    int ex, ey, ez;
    const long long cx = axis_cell(px, g.lo[0], g, &ex), cy = axis_cell(py, g.lo[1], g, &ey),
                    cz = axis_cell(pz, g.lo[2], g, &ez);
    const int ncell = (ex ? 2 : 1) * (ey ? 2 : 1) * (ez ? 2 : 1);
    for (int q = 0; q < ncell; q++) {
        // q's bits pick the neighbour on each axis that has one
        int bit = 0;
        long long x = cx, y = cy, z = cz;
        if (ex) x += ((q >> bit++) & 1) ? ex : 0;
        if (ey) y += ((q >> bit++) & 1) ? ey : 0;
        if (ez) z += ((q >> bit++) & 1) ? ez : 0;
        const unsigned long long key =
            ((unsigned long long)x << 42) | ((unsigned long long)y << 21) | (unsigned long long)z;
        uint32_t s = cell_hash(key) & mask;
        for (;;) {
            const unsigned long long k2 = keys[s];
            if (k2 == CELL_EMPTY) break;
            if (k2 == key) {
                const uint32_t end = start[s + 1];
                for (uint32_t e = start[s]; e < end; e += 8) {
                    uint32_t jj[8], xv[8];
                    float qx[8], qy[8], qz[8];
#pragma unroll
                    for (int k = 0; k < 8; k++) jj[k] = e + k < end ? members[e + k] : i;
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        const uint32_t j = jj[k] < i ? jj[k] : 0u;
                        qx[k] = uniq[3 * (size_t)j];
                        qy[k] = uniq[3 * (size_t)j + 1];
                        qz[k] = uniq[3 * (size_t)j + 2];
                        xv[k] = extra ? extra[j] : 0u;
                    }
#pragma unroll
                    for (int k = 0; k < 8; k++) {
                        if (jj[k] >= i) continue;
                        const double dx = (double)px - (double)qx[k];
                        const double dy = (double)py - (double)qy[k];
                        const double dz = (double)pz - (double)qz[k];
                        if (dx * dx + dy * dy + dz * dz <= g.eps2) f(jj[k], xv[k]);
                    }
                }
                break;
            }
            s = (s + 1) & mask;
        }
    }
}

// low[i] = the smallest j <= i within eps of i (i when there is none).
__global__ __launch_bounds__(BLK) void k_tol_lowest(const float* uniq, uint32_t m, Grid g,
                                                    const unsigned long long* keys, const uint32_t* start,
                                                    const uint32_t* members, uint32_t mask, uint32_t* low) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    uint32_t lo = i;
    for_each_near(uniq, i, g, keys, start, members, mask, nullptr, [&](uint32_t j, uint32_t) { lo = min(lo, j); });
    low[i] = lo;
}

// Unites every pair (i, j < i) within eps.  Each point is first joined to
// low[i]; a pair with low[i] == low[j] is then already connected through
// that point and needs no union-find work, which makes a cluster of k
// mutually close points cost k unions instead of k^2 / 2.
__global__ __launch_bounds__(BLK) void k_tol_unite(const float* uniq, uint32_t m, Grid g,
                                                   const unsigned long long* keys, const uint32_t* start,
                                                   const uint32_t* members, uint32_t mask, const uint32_t* low,
                                                   uint32_t* parent) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= m) return;
    const uint32_t li = low[i];
    if (li == i) return;  // no earlier point within eps
    uint32_t ri = uf_unite(parent, i, li);  // an ancestor of i from here on
    for_each_near(uniq, i, g, keys, start, members, mask, low, [&](uint32_t j, uint32_t lj) {
        if (lj == li) return;
        if (uf_load(&parent[j]) == ri) return;  // j already hangs under i's root
        ri = uf_unite(parent, ri, j);
    });
}

// root[u] = the smallest member of u's component; flag = (root == u).
__global__ __launch_bounds__(BLK) void k_uf_roots(uint32_t* parent, uint32_t m, uint32_t* flag) {
    const uint32_t u = blockIdx.x * BLK + threadIdx.x;
    if (u >= m) return;
    uint32_t r = u;
    while (parent[r] != r) r = parent[r];
    flag[u] = r == u ? 1u : 0u;
    parent[u] = r;  // every write stores a root, so concurrent readers still reach one
}

__global__ __launch_bounds__(BLK) void k_tol_number(const float* uniq, const uint32_t* root, const uint32_t* cidx,
                                                    uint32_t m, float* out) {
    const uint32_t u = blockIdx.x * BLK + threadIdx.x;
    if (u >= m || root[u] != u) return;
    const uint32_t c = cidx[u];
    for (int k = 0; k < 3; k++) out[3 * (size_t)c + k] = uniq[3 * (size_t)u + k];
}

__global__ __launch_bounds__(BLK) void k_tol_remap(uint32_t* uid, uint32_t nv, const uint32_t* root,
                                                   const uint32_t* cidx) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv || uid[i] == NONE) return;
    uid[i] = cidx[root[uid[i]]];
}

// out = the triangle remapped (remap may be null: a copy), keep = no
// repeated index (weld.rs:104-112); keep may be null.  out may equal tris.
__global__ __launch_bounds__(BLK) void k_tri_remap(const uint32_t* tris, uint32_t nt, const uint32_t* remap,
                                                   uint32_t* out, uint32_t* keep) {
    const uint32_t t = blockIdx.x * BLK + threadIdx.x;
    if (t >= nt) return;
    uint32_t v[3];
    for (int k = 0; k < 3; k++) v[k] = remap ? remap[tris[3 * (size_t)t + k]] : tris[3 * (size_t)t + k];
    for (int k = 0; k < 3; k++) out[3 * (size_t)t + k] = v[k];
    if (keep) keep[t] = (v[0] != v[1] && v[1] != v[2] && v[0] != v[2]) ? 1u : 0u;
}

// compact (weld.rs:75-84): positions of used vertices in order.
__global__ __launch_bounds__(BLK) void k_gather_used(const float* pos, const uint32_t* used, const uint32_t* uidx,
                                                     uint32_t nv, float* out, uint32_t* remap) {
    const uint32_t i = blockIdx.x * BLK + threadIdx.x;
    if (i >= nv) return;
    if (!used[i]) {
        remap[i] = NONE;
        return;
    }
    const uint32_t u = uidx[i];
    remap[i] = u;
    for (int k = 0; k < 3; k++) out[3 * (size_t)u + k] = pos[3 * (size_t)i + k];
}

// ------------------------------------------------------------------ host side
uint32_t pow2_at_least(size_t n) {
    uint32_t s = 1024;
    while (s < n) s <<= 1;
    return s;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s) {
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, st);
    }
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    float stop_ms() {
        float ms = 0.0f;
        (void)hipEventRecord(b, st);
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    }
};

}  // namespace
}  // namespace ag

namespace ag {
namespace {

// Workspace pieces of a triangle compaction over nt triangles.
struct TriWork {
    size_t keep, idx, work, tot, tmp;
};
TriWork plan_tris(Arena& a, uint32_t nt) {
    TriWork w;
    w.keep = a.add(4 * (size_t)nt);
    w.idx = a.add(4 * (size_t)nt);
    w.work = a.add(4 * scan_work_words(nt));
    w.tot = a.add(4);
    w.tmp = a.add(12 * (size_t)nt);
    return w;
}

// Keeps the triangles of tmp whose keep flag is 1, in order, in m->tris
// (util.rs:132-164).  Synchronizes.
int compact_from_tmp(hipStream_t st, ag_mesh* m, const Arena& a, const TriWork& w, uint32_t* removed) {
    const uint32_t nt = m->nt;
    scan(st, a.at<uint32_t>(w.keep), a.at<uint32_t>(w.idx), nt, a.at<uint32_t>(w.work), a.at<uint32_t>(w.tot));
    k_compact_tris<<<grid_of(nt), BLK, 0, st>>>(a.at<uint32_t>(w.tmp), a.at<uint32_t>(w.keep), a.at<uint32_t>(w.idx),
                                                nt, m->tris.as<uint32_t>());
    MCHK(hipGetLastError());
    uint32_t kept = 0;
    MCHK(hipMemcpyAsync(&kept, a.at<uint32_t>(w.tot), 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    *removed = nt - kept;
    m->nt = kept;
    return AG_OK;
}

// Remaps the triangles (remap may be null) and drops those with a repeated
// index (weld.rs:104-112).  Synchronizes.
int remap_drop_collapsed(hipStream_t st, ag_mesh* m, const Arena& a, const TriWork& w, const uint32_t* remap,
                         uint32_t* dropped) {
    k_tri_remap<<<grid_of(m->nt), BLK, 0, st>>>(m->tris.as<uint32_t>(), m->nt, remap, a.at<uint32_t>(w.tmp),
                                                 a.at<uint32_t>(w.keep));
    MCHK(hipGetLastError());
    return compact_from_tmp(st, m, a, w, dropped);
}

// used flags of referenced vertices; flag_dev[0] = 1 if an index is out of range.
int mark_used(hipStream_t st, const ag_mesh* m, uint32_t* used, uint32_t* flag_dev) {
    MCHK(hipMemsetAsync(used, 0, 4 * (size_t)m->nv, st));
    MCHK(hipMemsetAsync(flag_dev, 0, 4, st));
    k_mark_used<<<grid_of(m->nt), BLK, 0, st>>>(m->tris.as<uint32_t>(), m->nt, m->nv, used, flag_dev);
    MCHK(hipGetLastError());
    return AG_OK;
}

float from_ord(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Rust's f64::max / f32::max: a NaN operand yields the other one.
template <class T>
T rust_max(T a, T b) {
    if (b != b) return a;
    if (a != a) return b;
    return b > a ? b : a;
}

// weld.rs:38-62 on a device mesh.  Syncs: after the exact stage, after the
// tolerance stage, after the triangle compaction.
int weld_mesh(ag_mesh* m, float rel_tolerance, ag_weld_report* rep) {
    hipStream_t st = ctx_stream(m->ctx);
    const uint32_t nv = m->nv, nt = m->nt;
    const uint32_t tsize = pow2_at_least(2 * (size_t)nv);  // also bounds the cell table (classes <= nv)
    const unsigned nb = std::min<unsigned>(grid_of(nv), BOUNDS_BLOCKS);
    Arena a;
    const size_t o_used = a.add(4 * (size_t)nv), o_flag = a.add(256), o_part = a.add(28 * (size_t)nb),
                 o_table = a.add(4 * (size_t)tsize), o_rep = a.add(4 * (size_t)nv), o_head = a.add(4 * (size_t)nv),
                 o_hidx = a.add(4 * (size_t)nv), o_work = a.add(4 * scan_work_words((size_t)tsize + 1)),
                 o_uniq = a.add(12 * (size_t)nv), o_keys = a.add(8 * (size_t)tsize), o_cnt = a.add(4 * (size_t)tsize),
                 o_start = a.add(4 * ((size_t)tsize + 1)), o_slot = a.add(4 * (size_t)nv),
                 o_mem = a.add(4 * (size_t)nv), o_par = a.add(4 * (size_t)nv), o_cidx = a.add(4 * (size_t)nv);
    const TriWork tw = plan_tris(a, nt);
    int rc = a.bind(m->ctx);
    if (rc) return rc;
    uint32_t* used = a.at<uint32_t>(o_used);
    uint32_t* flag = a.at<uint32_t>(o_flag);  // [0] out of range, [1..6] lo/hi, [7] used count, [8] classes, [9] final
    uint32_t* work = a.at<uint32_t>(o_work);
    const float* pos = m->pos.as<float>();
    if ((rc = mark_used(st, m, used, flag))) return rc;

    // bounds and count of the referenced vertices (weld.rs:40)
    Init6 init;
    const float fmax_ = 3.40282347e38f;
    uint32_t bmax;
    std::memcpy(&bmax, &fmax_, 4);
    for (int k = 0; k < 3; k++) {
        init.v[k] = bmax | 0x80000000u;         // ord(+MAX)
        init.v[3 + k] = ~(bmax | 0x80000000u);  // ord(-MAX)
    }
    k_bounds<<<nb, BLK, 0, st>>>(pos, used, nv, a.at<uint32_t>(o_part));
    k_bounds_final<<<1, BLK, 0, st>>>(a.at<uint32_t>(o_part), nb, init, flag + 1);

    // exact stage
    uint32_t* table = a.at<uint32_t>(o_table);
    uint32_t* hidx = a.at<uint32_t>(o_hidx);
    uint32_t* uid = a.at<uint32_t>(o_head);  // the head flags, then (after the scan) class numbers
    float* uniq = a.at<float>(o_uniq);
    MCHK(hipMemsetAsync(table, 0xFF, 4 * (size_t)tsize, st));
    k_weld_insert<<<grid_of(nv), BLK, 0, st>>>(pos, used, nv, table, tsize - 1);
    k_weld_find<<<grid_of(nv), BLK, 0, st>>>(pos, used, nv, table, tsize - 1, a.at<uint32_t>(o_rep), uid);
    scan(st, uid, hidx, nv, work, flag + 8);
    k_weld_number<<<grid_of(nv), BLK, 0, st>>>(pos, a.at<uint32_t>(o_rep), hidx, nv, uid, uniq);
    MCHK(hipGetLastError());
    uint32_t hv[9];
    MCHK(hipMemcpyAsync(hv, flag, sizeof hv, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if (hv[0]) return fail(AG_E_INVALID, "triangle vertex index out of range");
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = from_ord(hv[1 + k]);
        hi[k] = from_ord(hv[4 + k]);
    }
    if (lo[0] > hi[0])
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = 0.0f;
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float diag = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float eps = (float)((double)rust_max(rel_tolerance, 0.0f) * (double)diag);
    const uint32_t used_count = hv[7], mu = hv[8];

    uint32_t mfinal = mu;
    float* out_pos = m->pos.as<float>();  // capacity nv >= every class count; the soup is no longer read
    if (eps > 0.0f && mu > 1) {
        // tolerance stage; the bounds of the classes equal those of the
        // referenced vertices (weld.rs:230-237 recomputes them over uniq)
        double ext = 0.0;
        for (int k = 0; k < 3; k++) ext = rust_max(ext, (double)(hi[k] - lo[k]));
        Grid g;
        for (int k = 0; k < 3; k++) g.lo[k] = (double)lo[k];
        g.h = rust_max(rust_max(8.0 * (double)eps, ext / (double)(1ull << 20)), 1e-30);
        g.eps = (double)eps;
        g.eps2 = g.eps * g.eps;
        const uint32_t csize = pow2_at_least(2 * (size_t)mu);
        auto* keys = a.at<unsigned long long>(o_keys);
        uint32_t* cnt = a.at<uint32_t>(o_cnt);
        uint32_t* start = a.at<uint32_t>(o_start);
        uint32_t* parent = a.at<uint32_t>(o_par);
        uint32_t* cidx = a.at<uint32_t>(o_cidx);
        uint32_t* slot = a.at<uint32_t>(o_slot);
        MCHK(hipMemsetAsync(keys, 0xFF, 8 * (size_t)csize, st));
        MCHK(hipMemsetAsync(cnt, 0, 4 * (size_t)csize, st));
        k_cell_insert<<<grid_of(mu), BLK, 0, st>>>(uniq, mu, g, keys, cnt, slot, csize - 1);
        scan(st, cnt, start, csize, work, start + csize);
        k_cell_scatter<<<grid_of(mu), BLK, 0, st>>>(slot, mu, start, cnt, a.at<uint32_t>(o_mem));
        k_iota<<<grid_of(mu), BLK, 0, st>>>(parent, mu);
        k_tol_lowest<<<grid_of(mu), BLK, 0, st>>>(uniq, mu, g, keys, start, a.at<uint32_t>(o_mem), csize - 1, cidx);
        k_tol_unite<<<grid_of(mu), BLK, 0, st>>>(uniq, mu, g, keys, start, a.at<uint32_t>(o_mem), csize - 1, cidx,
                                                 parent);
        k_uf_roots<<<grid_of(mu), BLK, 0, st>>>(parent, mu, slot);  // slot numbers are done with: root flags
        scan(st, slot, cidx, mu, work, flag + 9);
        k_tol_number<<<grid_of(mu), BLK, 0, st>>>(uniq, parent, cidx, mu, out_pos);
        k_tol_remap<<<grid_of(nv), BLK, 0, st>>>(uid, nv, parent, cidx);
        MCHK(hipGetLastError());
        MCHK(hipMemcpyAsync(&mfinal, flag + 9, 4, hipMemcpyDeviceToHost, st));
    } else {
        MCHK(hipMemcpyAsync(out_pos, uniq, 12 * (size_t)mu, hipMemcpyDeviceToDevice, st));
    }
    uint32_t dropped = 0;
    if ((rc = remap_drop_collapsed(st, m, a, tw, uid, &dropped))) return rc;  // synchronizes
    m->nv = mfinal;
    if (rep) {
        rep->input_vertices = nv;
        rep->output_vertices = mfinal;
        rep->exact_merges = used_count - mu;
        rep->tolerance_merges = mu - mfinal;
        rep->dropped_triangles = dropped;
        rep->tolerance_mm = eps;
    }
    return AG_OK;
}

// weld.rs:66-91 on a device mesh.
int compact_mesh(ag_mesh* m, uint64_t* dropped_out) {
    hipStream_t st = ctx_stream(m->ctx);
    const uint32_t nv = m->nv, nt = m->nt;
    Arena a;
    const size_t o_used = a.add(4 * (size_t)nv), o_flag = a.add(256), o_idx = a.add(4 * (size_t)nv),
                 o_work = a.add(4 * scan_work_words(nv)), o_remap = a.add(4 * (size_t)nv),
                 o_pos = a.add(12 * (size_t)nv);
    const TriWork tw = plan_tris(a, nt);
    int rc = a.bind(m->ctx);
    if (rc) return rc;
    uint32_t* used = a.at<uint32_t>(o_used);
    uint32_t* flag = a.at<uint32_t>(o_flag);
    // out-of-range indices first, so an error leaves the mesh unchanged
    if ((rc = mark_used(st, m, used, flag))) return rc;
    uint32_t oob = 0;
    MCHK(hipMemcpyAsync(&oob, flag, 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if (oob) return fail(AG_E_INVALID, "triangle vertex index out of range");
    uint32_t dropped = 0;
    if ((rc = remap_drop_collapsed(st, m, a, tw, nullptr, &dropped))) return rc;
    if (dropped_out) *dropped_out = dropped;
    if ((rc = mark_used(st, m, used, flag))) return rc;
    scan(st, used, a.at<uint32_t>(o_idx), nv, a.at<uint32_t>(o_work), flag + 1);
    uint32_t nused = 0;
    MCHK(hipMemcpyAsync(&nused, flag + 1, 4, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if (nused == nv) return AG_OK;
    k_gather_used<<<grid_of(nv), BLK, 0, st>>>(m->pos.as<float>(), used, a.at<uint32_t>(o_idx), nv,
                                               a.at<float>(o_pos), a.at<uint32_t>(o_remap));
    MCHK(hipMemcpyAsync(m->pos.p, a.at<float>(o_pos), 12 * (size_t)nused, hipMemcpyDeviceToDevice, st));
    k_tri_remap<<<grid_of(m->nt), BLK, 0, st>>>(m->tris.as<uint32_t>(), m->nt, a.at<uint32_t>(o_remap),
                                                 m->tris.as<uint32_t>(), nullptr);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    m->nv = nused;
    return AG_OK;
}

constexpr uint32_t MAX_VERTICES = 0x7fffffffu;

}  // namespace
}  // namespace ag

namespace ag {
void device_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* work, uint32_t* total) {
    scan(st, in, out, n, work, total);
}
size_t device_scan_work_words(size_t n) { return scan_work_words(n); }

// Every triangle index below nv?  And the min / max of all vertex positions
// (argent-mesh lib.rs:122-134, Mesh::bounds: zeros without vertices).
int mesh_check_bounds(ag_mesh* m, float lo[3], float hi[3]) {
    hipStream_t st = ctx_stream(m->ctx);
    const uint32_t nv = m->nv;
    const unsigned nb = std::min<unsigned>(grid_of(nv), BOUNDS_BLOCKS);
    Arena a;
    const size_t o_ones = a.add(4 * (size_t)std::max(nv, 1u)), o_flag = a.add(256), o_part = a.add(28 * (size_t)nb);
    int rc = a.bind(m->ctx);
    if (rc) return rc;
    uint32_t* flag = a.at<uint32_t>(o_flag);
    if ((rc = mark_used(st, m, a.at<uint32_t>(o_ones), flag))) return rc;  // out-of-range indices -> flag[0]
    k_fill_ones<<<grid_of(nv), BLK, 0, st>>>(a.at<uint32_t>(o_ones), nv);
    Init6 init;
    const float fmax_ = 3.40282347e38f;
    uint32_t bmax;
    std::memcpy(&bmax, &fmax_, 4);
    for (int k = 0; k < 3; k++) {
        init.v[k] = bmax | 0x80000000u;
        init.v[3 + k] = ~(bmax | 0x80000000u);
    }
    if (nv) {
        k_bounds<<<nb, BLK, 0, st>>>(m->pos.as<float>(), a.at<uint32_t>(o_ones), nv, a.at<uint32_t>(o_part));
        k_bounds_final<<<1, BLK, 0, st>>>(a.at<uint32_t>(o_part), nb, init, flag + 1);
    }
    MCHK(hipGetLastError());
    uint32_t hv[7] = {0, 0, 0, 0, 0, 0, 0};
    MCHK(hipMemcpyAsync(hv, flag, sizeof hv, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    if (hv[0]) return fail(AG_E_INVALID, "triangle vertex index out of range");
    for (int k = 0; k < 3; k++) {
        lo[k] = nv ? from_ord(hv[1 + k]) : 0.0f;
        hi[k] = nv ? from_ord(hv[4 + k]) : 0.0f;
    }
    return AG_OK;
}

extern "C" int ag_mesh_frame_sample(const ag_mesh* m, uint32_t stride, float* hull, uint32_t cap, uint32_t* n_hull,
                                    float* lo, float* hi) {
    if (!m || !n_hull || !lo || !hi) return fail(AG_E_INVALID, "null argument");
    if (stride < 1) return fail(AG_E_INVALID, "stride must be at least 1");
    if (cap && !hull) return fail(AG_E_INVALID, "null hull buffer with cap > 0");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    hipStream_t st = ctx_stream(m->ctx);
    const uint32_t nv = m->nv;
    const uint32_t n = (uint32_t)std::min<uint64_t>(cap, ((uint64_t)nv + stride - 1) / stride);
    const unsigned nb = std::min<unsigned>(grid_of(nv), BOUNDS_BLOCKS);
    Arena a;
    const size_t o_ones = a.add(4 * (size_t)std::max(nv, 1u)), o_flag = a.add(256), o_part = a.add(28 * (size_t)nb),
                 o_hull = a.add(12 * (size_t)std::max(n, 1u));
    int rc = a.bind(m->ctx);
    if (rc) return rc;
    uint32_t* flag = a.at<uint32_t>(o_flag);
    Init6 init;
    const float fmax_ = 3.40282347e38f;
    uint32_t bmax;
    std::memcpy(&bmax, &fmax_, 4);
    for (int k = 0; k < 3; k++) {
        init.v[k] = bmax | 0x80000000u;
        init.v[3 + k] = ~(bmax | 0x80000000u);
    }
    if (nv) {
        k_fill_ones<<<grid_of(nv), BLK, 0, st>>>(a.at<uint32_t>(o_ones), nv);
        k_bounds<<<nb, BLK, 0, st>>>(m->pos.as<float>(), a.at<uint32_t>(o_ones), nv, a.at<uint32_t>(o_part));
        k_bounds_final<<<1, BLK, 0, st>>>(a.at<uint32_t>(o_part), nb, init, flag + 1);
    }
    if (n) k_gather_stride<<<grid_of(n), BLK, 0, st>>>(m->pos.as<float>(), stride, n, a.at<float>(o_hull));
    MCHK(hipGetLastError());
    uint32_t hv[6] = {0, 0, 0, 0, 0, 0};
    if (nv) MCHK(hipMemcpyAsync(hv, flag + 1, sizeof hv, hipMemcpyDeviceToHost, st));
    if (n) MCHK(hipMemcpyAsync(hull, a.at<float>(o_hull), 12 * (size_t)n, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    for (int k = 0; k < 3; k++) {
        lo[k] = nv ? from_ord(hv[k]) : 0.0f;
        hi[k] = nv ? from_ord(hv[3 + k]) : 0.0f;
    }
    *n_hull = n;
    return AG_OK;
}

int mesh_drop_collapsed(ag_mesh* m, uint32_t* dropped) {
    *dropped = 0;
    if (m->nt == 0) return AG_OK;
    Arena a;
    const TriWork tw = plan_tris(a, m->nt);
    int rc = a.bind(m->ctx);
    if (rc) return rc;
    return remap_drop_collapsed(ctx_stream(m->ctx), m, a, tw, nullptr, dropped);
}
}  // namespace ag

using namespace ag;

extern "C" int ag_mesh_create(ag_ctx* ctx, const float* positions, uint32_t n_vertices, const uint32_t* triangles,
                              uint32_t n_triangles, ag_mesh** out) {
    if (!ctx || !out || (n_vertices && !positions) || (n_triangles && !triangles))
        return fail(AG_E_INVALID, "null argument");
    if (n_vertices > MAX_VERTICES || (uint64_t)n_triangles * 3 >= 0xFFFFFFFFull)
        return fail(AG_E_INVALID, "mesh too large");
    MCHK(hipSetDevice(ctx_device(ctx)));
    auto* m = new ag_mesh();
    m->ctx = ctx;
    int rc = m->pos.alloc(12 * (size_t)n_vertices);
    if (!rc) rc = m->tris.alloc(12 * (size_t)n_triangles);
    if (rc) {
        delete m;
        return rc;
    }
    hipStream_t st = ctx_stream(ctx);
    if (hipMemcpyAsync(m->pos.p, positions, 12 * (size_t)n_vertices, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(m->tris.p, triangles, 12 * (size_t)n_triangles, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        delete m;
        return fail(AG_E_DEVICE, "mesh upload failed");
    }
    m->nv = n_vertices;
    m->nt = n_triangles;
    *out = m;
    return AG_OK;
}

extern "C" void ag_mesh_destroy(ag_mesh* m) { delete m; }

extern "C" int ag_mesh_info(const ag_mesh* m, uint32_t* n_vertices, uint32_t* n_triangles) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    if (n_vertices) *n_vertices = m->nv;
    if (n_triangles) *n_triangles = m->nt;
    return AG_OK;
}

extern "C" int ag_mesh_read(const ag_mesh* m, float* positions, uint32_t* triangles) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    hipStream_t st = ctx_stream(m->ctx);
    if (positions && m->nv) MCHK(hipMemcpyAsync(positions, m->pos.p, 12 * (size_t)m->nv, hipMemcpyDeviceToHost, st));
    if (triangles && m->nt) MCHK(hipMemcpyAsync(triangles, m->tris.p, 12 * (size_t)m->nt, hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));
    return AG_OK;
}

extern "C" int ag_mesh_device_arrays(const ag_mesh* m, const float** positions, const uint32_t** triangles) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    if (positions) *positions = m->pos.as<float>();
    if (triangles) *triangles = m->tris.as<uint32_t>();
    return AG_OK;
}

extern "C" int ag_mesh_weld(ag_mesh* m, float rel_tolerance, ag_weld_report* report) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    m->has_nrm = false;  // weld.rs:56: welding clears the normals
    if (m->nt == 0 || m->nv == 0) {
        if (report) {
            *report = ag_weld_report{};
            report->input_vertices = m->nv;
        }
        // nothing is referenced: every vertex goes
        m->nv = 0;
        return AG_OK;
    }
    return weld_mesh(m, rel_tolerance, report);
}

extern "C" int ag_mesh_compact(ag_mesh* m, uint64_t* dropped_triangles) {
    if (!m) return fail(AG_E_INVALID, "null mesh");
    MCHK(hipSetDevice(ctx_device(m->ctx)));
    m->has_nrm = false;
    if (m->nt == 0) {
        if (dropped_triangles) *dropped_triangles = 0;
        m->nv = 0;
        return AG_OK;
    }
    return compact_mesh(m, dropped_triangles);
}

extern "C" int ag_mesh_load_stl(ag_ctx* ctx, const void* bytes, size_t len, int32_t bytes_on_device,
                                const ag_load_options* opts, ag_mesh** out, ag_load_report* report) {
    if (!ctx || !bytes || !opts || !out) return fail(AG_E_INVALID, "null argument");
    if (len < 84) return fail(AG_E_INVALID, "binary STL shorter than its 84-byte header");
    if (bytes_on_device && (reinterpret_cast<uintptr_t>(bytes) & 3))
        return fail(AG_E_INVALID, "device STL bytes must be 4-byte aligned");
    MCHK(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    ag_load_report rep{};
    rep.file_bytes = len;
    const uint8_t* dev = static_cast<const uint8_t*>(bytes);
    uint32_t declared = 0;
    if (bytes_on_device) {
        MCHK(hipMemcpyAsync(&declared, dev + 80, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
    } else {
        std::memcpy(&declared, dev + 80, 4);
    }
    // io/stl.rs:38-50: the declared count, or the records the file holds
    const uint64_t available = (len - 84) / 50;
    const uint64_t n = declared <= available ? declared : available;
    rep.declared_triangles = declared;
    rep.trailing_bytes = declared <= available ? len - 84 - 50 * (uint64_t)declared : 0;
    rep.raw_triangles = n;
    rep.raw_vertices = 3 * n;
    if (n == 0) return fail(AG_E_INVALID, "STL file contains no triangles");
    if (3 * n > MAX_VERTICES) return fail(AG_E_INVALID, "STL file too large");

    Arena a;
    const size_t o_bytes = bytes_on_device ? 0 : a.add(len), o_nbad = a.add(4);
    const TriWork tw = plan_tris(a, (uint32_t)n);
    int rc = a.bind(ctx);
    if (rc) return rc;
    if (!bytes_on_device) {
        const auto t0 = std::chrono::steady_clock::now();
        MCHK(hipMemcpyAsync(a.at<uint8_t>(o_bytes), bytes, len, hipMemcpyHostToDevice, st));
        MCHK(hipStreamSynchronize(st));
        rep.upload_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
        dev = a.at<uint8_t>(o_bytes);
    }
    auto* m = new ag_mesh();
    m->ctx = ctx;
    std::unique_ptr<ag_mesh> hold(m);
    rc = m->pos.alloc(36 * n);
    if (!rc) rc = m->tris.alloc(12 * n);
    if (rc) return rc;
    m->nv = (uint32_t)(3 * n);
    m->nt = (uint32_t)n;
    const float s = opts->unit_scale;
    const int do_scale = s != 1.0f && std::isfinite(s) && s > 0.0f;
    {
        Timer tm(st);
        uint32_t* nbad = a.at<uint32_t>(o_nbad);
        MCHK(hipMemsetAsync(nbad, 0, 4, st));
        k_stl_parse<<<grid_of(n), BLK, 0, st>>>(dev, (uint32_t)n, s, do_scale, m->pos.as<float>(),
                                                m->tris.as<uint32_t>(), a.at<uint32_t>(tw.keep), nbad);
        MCHK(hipGetLastError());
        uint32_t bad = 0;
        MCHK(hipMemcpyAsync(&bad, nbad, 4, hipMemcpyDeviceToHost, st));
        MCHK(hipStreamSynchronize(st));
        if (bad) {  // io.rs:206-224
            k_tri_remap<<<grid_of(n), BLK, 0, st>>>(m->tris.as<uint32_t>(), m->nt, nullptr, a.at<uint32_t>(tw.tmp),
                                                    nullptr);
            uint32_t removed = 0;
            if ((rc = compact_from_tmp(st, m, a, tw, &removed))) return rc;
            rep.invalid_triangles = removed;
        }
        rep.parse_ms = tm.stop_ms();
    }
    if (m->nt == 0) return fail(AG_E_INVALID, "STL file has only degenerate triangles");
    {
        Timer tm(st);
        if (opts->weld) {
            rep.welded = 1;
            rc = weld_mesh(m, opts->weld_tolerance, &rep.weld);
        } else {
            rc = compact_mesh(m, nullptr);
        }
        if (rc) return rc;
        rep.weld_ms = tm.stop_ms();
    }
    if (m->nt == 0) return fail(AG_E_INVALID, "STL file has only degenerate triangles");
    if (report) *report = rep;
    *out = hold.release();
    return AG_OK;
}
