// BVH build on the device, for scenes whose time to first picture matters
// more than the last few percent of traversal speed (a new mesh in serving).
// Same output format as the host build (host_bvh.cpp): compressed 8-wide
// nodes of 80 bytes with conservatively quantized child boxes, triangles of
// 48 bytes grouped per node, so the traversal kernels take either tree.
//
//   1. triangle boxes and centroid bounds;
//   2. 30-bit Morton codes of the centroids, sorted with their triangle
//      index (rocPRIM radix sort; ties keep index order);
//   3. the binary radix tree over the sorted codes (Karras 2012, "Maximizing
//      parallelism in the construction of BVHs, octrees and k-d trees"):
//      internal node i covers a contiguous range of sorted triangles, found
//      from the longest common prefixes of its neighbours' keys (key = code,
//      then sorted position, so all keys differ);
//   4. node boxes as unions over each node's sorted range (block prefix /
//      suffix unions and a sparse table over blocks: no inter-thread
//      signalling, whose agent-scope fences write back L2 on this part);
//   5. collapse to 8-wide nodes level by level with the host build's policy
//      (open the child of largest surface area until 8; a subtree of at most
//      4 triangles is a leaf), node and triangle slots handed out by scans so
//      the layout does not depend on thread timing, and quantization of the
//      child boxes exactly as host_bvh.cpp does it.
// Closest-hit results do not depend on the tree; traversal cost does (an
// LBVH is typically 10-30 % more expensive to traverse than binned SAH).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <string>
#include <vector>

#include "host_internal.hpp"

namespace ag {
namespace {

constexpr int BB = 256;
constexpr uint32_t LEAF_MAX = BVH_MAX_LEAF;  // triangles per leaf

#define BCHK(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return fail(AG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

unsigned blocks(size_t n) { return (unsigned)std::max<size_t>(1, (n + BB - 1) / BB); }

__device__ inline uint32_t ordf(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float unordf(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u); }

// ---------------------------------------------------------------- 1. boxes
// box[6 * t] = lo xyz, hi xyz (fmin / fmax of the corners, as host_bvh.cpp);
// part[6 * block] = the block's centroid min / max as ordered integers.
__global__ __launch_bounds__(BB) void k_tri_boxes(const float* pos, const uint32_t* idx, uint32_t n, float* box,
                                                  uint32_t* part) {
    __shared__ uint32_t red[BB / 64][6];
    const uint32_t t = blockIdx.x * BB + threadIdx.x;
    uint32_t cl[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, ch[3] = {0u, 0u, 0u};
    if (t < n) {
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = INFINITY;
            hi[a] = -INFINITY;
        }
        for (int k = 0; k < 3; k++) {
            const float* p = pos + 3 * (size_t)idx[3 * (size_t)t + k];
            for (int a = 0; a < 3; a++) {
                lo[a] = fminf(lo[a], p[a]);
                hi[a] = fmaxf(hi[a], p[a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            box[6 * (size_t)t + a] = lo[a];
            box[6 * (size_t)t + 3 + a] = hi[a];
            const uint32_t c = ordf((lo[a] + hi[a]) * 0.5f);
            cl[a] = c;
            ch[a] = c;
        }
    }
    for (int d = 32; d >= 1; d >>= 1)
        for (int a = 0; a < 3; a++) {
            cl[a] = min(cl[a], (uint32_t)__shfl_xor((int)cl[a], d, 64));
            ch[a] = max(ch[a], (uint32_t)__shfl_xor((int)ch[a], d, 64));
        }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) {
            red[w][a] = cl[a];
            red[w][3 + a] = ch[a];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int j = threadIdx.x;
        uint32_t v = red[0][j];
        for (int q = 1; q < BB / 64; q++) v = j < 3 ? min(v, red[q][j]) : max(v, red[q][j]);
        part[6 * blockIdx.x + j] = v;
    }
}

__global__ __launch_bounds__(BB) void k_fold6(const uint32_t* part, uint32_t nb, uint32_t* out) {
    __shared__ uint32_t red[BB][6];
    uint32_t v[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t b = threadIdx.x; b < nb; b += BB)
        for (int j = 0; j < 6; j++) v[j] = j < 3 ? min(v[j], part[6 * b + j]) : max(v[j], part[6 * b + j]);
    for (int j = 0; j < 6; j++) red[threadIdx.x][j] = v[j];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int j = threadIdx.x;
        uint32_t r = red[0][j];
        for (int q = 1; q < BB; q++) r = j < 3 ? min(r, red[q][j]) : max(r, red[q][j]);
        out[j] = r;
    }
}

// ---------------------------------------------------------------- 2. Morton
__device__ inline uint32_t spread10(uint32_t v) {  // 10 bits -> every third of 30
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

__global__ __launch_bounds__(BB) void k_morton(const float* box, uint32_t n, const uint32_t* cb, uint32_t* code,
                                               uint32_t* id) {
    const uint32_t t = blockIdx.x * BB + threadIdx.x;
    if (t >= n) return;
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        const float lo = unordf(cb[a]), hi = unordf(cb[3 + a]);
        const float c = (box[6 * (size_t)t + a] + box[6 * (size_t)t + 3 + a]) * 0.5f;
        const float ext = hi - lo;
        const float u = ext > 0.0f ? (c - lo) / ext : 0.0f;
        q[a] = (uint32_t)fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f);
    }
    code[t] = spread10(q[0]) << 2 | spread10(q[1]) << 1 | spread10(q[2]);
    id[t] = t;
}

// ---------------------------------------------------------------- 3. radix tree
// Binary nodes: internal 0..n-2 (0 is the root), leaf i at n-1+i.
struct BinTree {
    int32_t* left;    // child node ids
    int32_t* right;
    int32_t* parent;  // 2n-1 entries
    uint32_t* first;  // sorted range of each node
    uint32_t* count;
    float* box;       // 6 per node
};

__device__ inline int delta(const uint32_t* code, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = code[i], b = code[j];
    return a != b ? __clz(a ^ b) : 32 + __clz((uint32_t)(i ^ j));
}

__global__ __launch_bounds__(BB) void k_karras(const uint32_t* code, int n, BinTree T) {
    const int i = blockIdx.x * BB + threadIdx.x;
    if (i >= n - 1) return;
    const int d = delta(code, n, i, i + 1) - delta(code, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = delta(code, n, i, i - d);
    int lmax = 2;
    while (delta(code, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (delta(code, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(code, n, i, j);
    int s = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (delta(code, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
    }
    const int g = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const int lc = lo == g ? (n - 1) + g : g;
    const int rc = hi == g + 1 ? (n - 1) + g + 1 : g + 1;
    T.left[i] = lc;
    T.right[i] = rc;
    T.parent[lc] = i;
    T.parent[rc] = i;
    T.first[i] = (uint32_t)lo;
    T.count[i] = (uint32_t)(hi - lo + 1);
}

// ---------------------------------------------------------------- 4. node boxes
// A node's box is the union of the boxes of its sorted range [first, first +
// count), built without inter-thread signalling: leaves first (they sit in
// sorted order at n-1 .. 2n-2), then per 64-triangle block the unions from
// each element to the block's end (suf) and from the block's start (pre),
// then a sparse table over whole blocks (level k: unions of 2^k blocks); a
// range is suf[a] u blocks between u pre[b], or a short loop inside one
// block.  Unions of min / max are exact, so the order does not matter.
constexpr int RB = 64;

__global__ __launch_bounds__(BB) void k_leaf_boxes(const float* tbox, const uint32_t* id, int n, BinTree T) {
    const int i = blockIdx.x * BB + threadIdx.x;
    if (i >= n) return;
    const int node = (n - 1) + i;
    T.first[node] = (uint32_t)i;
    T.count[node] = 1;
    for (int a = 0; a < 6; a++) T.box[6 * (size_t)node + a] = tbox[6 * (size_t)id[i] + a];
}

__device__ inline void join(float* b, const float* c) {
    for (int a = 0; a < 3; a++) {
        b[a] = fminf(b[a], c[a]);
        b[3 + a] = fmaxf(b[3 + a], c[3 + a]);
    }
}

// One wave per 64-element block of the sorted leaf boxes sb.
__global__ __launch_bounds__(BB) void k_block_unions(const float* sb, int n, float* pre, float* suf, float* blk) {
    const int i = blockIdx.x * BB + threadIdx.x;
    const int lane = threadIdx.x & 63;
    float v[6], p[6], q[6];
    for (int a = 0; a < 3; a++) {
        v[a] = INFINITY;
        v[3 + a] = -INFINITY;
    }
    if (i < n)
        for (int a = 0; a < 6; a++) v[a] = sb[6 * (size_t)i + a];
    for (int a = 0; a < 6; a++) p[a] = q[a] = v[a];
    for (int d = 1; d < 64; d <<= 1) {
        float up[6], dn[6];
        for (int a = 0; a < 6; a++) {
            up[a] = __shfl_up(p[a], d, 64);
            dn[a] = __shfl_down(q[a], d, 64);
        }
        if (lane >= d) join(p, up);
        if (lane + d < 64) join(q, dn);
    }
    if (i < n)
        for (int a = 0; a < 6; a++) {
            pre[6 * (size_t)i + a] = p[a];
            suf[6 * (size_t)i + a] = q[a];
        }
    if (lane == 0 && i < n)
        for (int a = 0; a < 6; a++) blk[6 * (size_t)(i / 64) + a] = q[a];
}

// Sparse-table level k from level k - 1: out[j] = in[j] u in[j + half].
__global__ __launch_bounds__(BB) void k_table_level(const float* in, int nb, int half, float* out) {
    const int j = blockIdx.x * BB + threadIdx.x;
    if (j >= nb) return;
    float b[6];
    for (int a = 0; a < 6; a++) b[a] = in[6 * (size_t)j + a];
    if (j + half < nb) join(b, &in[6 * (size_t)(j + half)]);
    for (int a = 0; a < 6; a++) out[6 * (size_t)j + a] = b[a];
}

__global__ __launch_bounds__(BB) void k_node_boxes(const float* sb, const float* pre, const float* suf,
                                                   const float* table, int nb, int n, BinTree T) {
    const int i = blockIdx.x * BB + threadIdx.x;
    if (i >= n - 1) return;
    const int lo = (int)T.first[i], hi = lo + (int)T.count[i] - 1;
    const int ba = lo / RB, bb = hi / RB;
    float b[6];
    if (ba == bb) {
        for (int a = 0; a < 6; a++) b[a] = sb[6 * (size_t)lo + a];
        for (int k = lo + 1; k <= hi; k++) join(b, &sb[6 * (size_t)k]);
    } else {
        for (int a = 0; a < 6; a++) b[a] = suf[6 * (size_t)lo + a];
        join(b, &pre[6 * (size_t)hi]);
        if (bb - ba >= 2) {
            const int l = ba + 1, r = bb - 1, len = r - l + 1;
            const int k = 31 - __clz(len);  // 2^k <= len
            const float* lev = table + (size_t)6 * nb * k;
            join(b, &lev[6 * (size_t)l]);
            join(b, &lev[6 * (size_t)(r - (1 << k) + 1)]);
        }
    }
    for (int a = 0; a < 6; a++) T.box[6 * (size_t)i + a] = b[a];
}

// ---------------------------------------------------------------- 5. collapse
__device__ inline float half_area(const float* b) {
    const float e0 = b[3] - b[0], e1 = b[4] - b[1], e2 = b[5] - b[2];
    return e0 * e1 + e1 * e2 + e2 * e0;
}
__device__ inline bool openable(const BinTree& T, int n, int b) { return b < n - 1 && T.count[b] > LEAF_MAX; }

// Per queued wide node: its up to 8 children (host_bvh.cpp fill(): start from
// the subtree root's two children, or the root itself when it is a leaf, and
// open the largest-area openable child while fewer than 8).
__global__ __launch_bounds__(BB) void k_plan(const int32_t* queue, uint32_t nq, int n, BinTree T, int32_t* kids,
                                             uint32_t* nkids, uint32_t* n_inner, uint32_t* n_tris) {
    const uint32_t q = blockIdx.x * BB + threadIdx.x;
    if (q >= nq) return;
    const int root = queue[q];
    int k[8];
    int nk = 0;
    if (!openable(T, n, root)) {
        k[nk++] = root;
    } else {
        k[nk++] = T.left[root];
        k[nk++] = T.right[root];
    }
    while (nk < 8) {
        int best = -1;
        float ba = -1.0f;
        for (int c = 0; c < nk; c++)
            if (openable(T, n, k[c])) {
                const float a = half_area(&T.box[6 * (size_t)k[c]]);
                if (best < 0 || a > ba) {
                    best = c;
                    ba = a;
                }
            }
        if (best < 0) break;
        const int b = k[best];
        k[best] = k[nk - 1];
        k[nk - 1] = T.left[b];
        k[nk++] = T.right[b];
    }
    uint32_t ni = 0, nt = 0;
    for (int c = 0; c < 8; c++) {
        kids[8 * (size_t)q + c] = c < nk ? k[c] : -1;
        if (c < nk) {
            if (openable(T, n, k[c])) ni++;
            else nt += T.count[k[c]];
        }
    }
    nkids[q] = (uint32_t)nk;
    n_inner[q] = ni;
    n_tris[q] = nt;
}

// The exponent byte of the smallest power-of-two scale s with p + 255 s >= hi
// (host_bvh.cpp axis_exponent).
__device__ uint32_t axis_exponent(float p, float hi) {
    const float ext = hi - p;
    int e = 1;
    if (ext > 0.0f) {
        int ex;
        frexpf(ext / 255.0f, &ex);
        e = min(max(ex + 127, 1), 254);
    }
    for (;; e++) {
        const float s = __uint_as_float((uint32_t)e << 23);
        if (fmaf(255.0f, s, p) >= hi || e >= 254) return (uint32_t)e;
    }
}

// host_bvh.cpp quantize_axis: fmaf(qlo, s, p) <= lo and fmaf(qhi, s, p) >= hi.
__device__ void quantize_axis(float p, float s, float lo, float hi, uint32_t* qlo, uint32_t* qhi) {
    const float fl = floorf((lo - p) / s), fh = ceilf((hi - p) / s);
    int l = (int)fminf(fmaxf(fl, 0.0f), 255.0f), h = (int)fminf(fmaxf(fh, 0.0f), 255.0f);
    while (l > 0 && fmaf((float)l, s, p) > lo) l--;
    while (h < 255 && fmaf((float)h, s, p) < hi) h++;
    *qlo = (uint32_t)l;
    *qhi = (uint32_t)h;
}

struct Emit {
    uint32_t node_base;   // global index of this level's first node
    uint32_t next_base;   // global index of the next level's first node
    uint32_t tri_base;    // first triangle slot of this level
};

// Writes each queued node (layout of host_internal.hpp), its triangles, and
// the next level's queue.
__global__ __launch_bounds__(BB) void k_emit(const int32_t* queue, uint32_t nq, int n, BinTree T,
                                             const int32_t* kids, const uint32_t* nkids, const uint32_t* inner_off,
                                             const uint32_t* tri_off, const uint32_t* stack_in, Emit E,
                                             const float* pos, const uint32_t* idx, const uint32_t* id, float* nodes,
                                             float* tris, int32_t* next, uint32_t* stack_out, uint32_t* max_stack) {
    const uint32_t q = blockIdx.x * BB + threadIdx.x;
    if (q >= nq) return;
    const int nk = (int)nkids[q];
    float nb[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < nk; c++) {
        const float* b = &T.box[6 * (size_t)kids[8 * (size_t)q + c]];
        for (int a = 0; a < 3; a++) {
            nb[a] = fminf(nb[a], b[a]);
            nb[3 + a] = fmaxf(nb[3 + a], b[3 + a]);
        }
    }
    uint32_t e[3];
    for (int a = 0; a < 3; a++) e[a] = axis_exponent(nb[a], nb[3 + a]);
    const uint32_t child_base = E.next_base + inner_off[q];
    const uint32_t tri_base = E.tri_base + tri_off[q];
    uint32_t imask = 0, n_inner = 0, toff = 0;
    uint32_t meta[2] = {0u, 0u};
    uint32_t qw[12];  // 48 bytes of quantized planes: qlo_x, qlo_y, qlo_z, qhi_x, qhi_y, qhi_z (8 each)
    for (int w = 0; w < 12; w++) qw[w] = 0u;
    const uint32_t my_stack = stack_in[q] + (uint32_t)nk - 1u;
    for (int c = 0; c < nk; c++) {
        const int b = kids[8 * (size_t)q + c];
        const float* cb = &T.box[6 * (size_t)b];
        for (int a = 0; a < 3; a++) {
            const float s = __uint_as_float(e[a] << 23);
            uint32_t ql, qh;
            quantize_axis(nb[a], s, cb[a], cb[3 + a], &ql, &qh);
            const int lo_byte = a * 8 + c, hi_byte = (3 + a) * 8 + c;
            qw[lo_byte >> 2] |= ql << (8 * (lo_byte & 3));
            qw[hi_byte >> 2] |= qh << (8 * (hi_byte & 3));
        }
        uint32_t m;
        if (openable(T, n, b)) {
            imask |= 1u << c;
            m = n_inner;
            next[inner_off[q] + n_inner] = b;
            stack_out[inner_off[q] + n_inner] = my_stack;
            n_inner++;
        } else {
            const uint32_t cnt = T.count[b], f = T.first[b];
            m = cnt << 5 | toff;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t tid = id[f + k];
                const float* p0 = pos + 3 * (size_t)idx[3 * (size_t)tid];
                const float* p1 = pos + 3 * (size_t)idx[3 * (size_t)tid + 1];
                const float* p2 = pos + 3 * (size_t)idx[3 * (size_t)tid + 2];
                float* o = tris + 12 * (size_t)(tri_base + toff + k);
                o[0] = p0[0];
                o[1] = p0[1];
                o[2] = p0[2];
                o[3] = p1[0] - p0[0];
                o[4] = p1[1] - p0[1];
                o[5] = p1[2] - p0[2];
                o[6] = p2[0] - p0[0];
                o[7] = p2[1] - p0[1];
                o[8] = p2[2] - p0[2];
                o[9] = __uint_as_float(tid);
                o[10] = 0.0f;
                o[11] = 0.0f;
            }
            toff += cnt;
        }
        meta[c >> 2] |= m << (8 * (c & 3));
    }
    float* nd = nodes + (size_t)BVH_NODE_FLOATS * (E.node_base + q);
    nd[0] = nb[0];
    nd[1] = nb[1];
    nd[2] = nb[2];
    nd[3] = __uint_as_float(e[0] | e[1] << 8 | e[2] << 16 | imask << 24);
    nd[4] = __uint_as_float(child_base);
    nd[5] = __uint_as_float(tri_base);
    nd[6] = __uint_as_float(meta[0]);
    nd[7] = __uint_as_float(meta[1]);
    for (int w = 0; w < 12; w++) nd[8 + w] = __uint_as_float(qw[w]);
    for (int w = 20; w < BVH_NODE_FLOATS; w++) nd[w] = 0.0f;
    atomicMax(max_stack, my_stack);
}

}  // namespace

// Exclusive scan for the collapse (mesh.hip's device scan, reached through
// this declaration so both files share one implementation).
void device_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* work, uint32_t* total);
size_t device_scan_work_words(size_t n);

int gpu_build_wide_bvh(ag_ctx* ctx, const float* pos, const uint32_t* idx, uint32_t n, float* nodes, float* tris,
                       GpuBvhInfo* info) {
    hipStream_t st = ctx_stream(ctx);
    *info = GpuBvhInfo{};
    if (n == 0) {
        // one node without children: every query misses
        std::vector<float> nd(BVH_NODE_FLOATS, 0.0f);
        const uint32_t w = 127u | 127u << 8 | 127u << 16;
        std::memcpy(&nd[3], &w, 4);
        BCHK(hipMemcpyAsync(nodes, nd.data(), nd.size() * 4, hipMemcpyHostToDevice, st));
        BCHK(hipStreamSynchronize(st));
        info->n_nodes = 1;
        info->depth = 1;
        return AG_OK;
    }
    const size_t nn = 2 * (size_t)n - 1;
    const unsigned nb = blocks(n);
    // workspace layout
    size_t used = 0;
    auto add = [&](size_t bytes) {
        const size_t o = used;
        used += (bytes + 255) & ~(size_t)255;
        return o;
    };
    size_t sort_bytes = 0;
    BCHK(rocprim::radix_sort_pairs(nullptr, sort_bytes, (uint32_t*)nullptr, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                   (uint32_t*)nullptr, n, 0, 30, st));
    const size_t o_tbox = add(24 * (size_t)n), o_part = add(24 * (size_t)nb), o_cb = add(64),
                 o_code = add(4 * (size_t)n), o_id = add(4 * (size_t)n), o_code2 = add(4 * (size_t)n),
                 o_id2 = add(4 * (size_t)n), o_sort = add(sort_bytes), o_left = add(4 * (size_t)n),
                 o_right = add(4 * (size_t)n), o_parent = add(4 * nn), o_first = add(4 * nn), o_count = add(4 * nn),
                 o_box = add(24 * nn), o_pre = add(24 * (size_t)n), o_suf = add(24 * (size_t)n),
                 o_table = add(24 * (((size_t)n + RB - 1) / RB) * 32), o_q0 = add(4 * (size_t)n), o_q1 = add(4 * (size_t)n),
                 o_s0 = add(4 * (size_t)n), o_s1 = add(4 * (size_t)n), o_kids = add(32 * (size_t)n),
                 o_nk = add(4 * (size_t)n), o_ni = add(4 * (size_t)n), o_nt = add(4 * (size_t)n),
                 o_io = add(4 * (size_t)n), o_to = add(4 * (size_t)n), o_ni2 = add(4 * (size_t)n), o_work = add(4 * device_scan_work_words(n)),
                 o_tot = add(64);
    void* base = nullptr;
    int rc = ctx_workspace(ctx, used, &base);
    if (rc) return rc;
    char* W = static_cast<char*>(base);
    auto at = [&](size_t o) { return W + o; };
    float* tbox = (float*)at(o_tbox);
    uint32_t* cb = (uint32_t*)at(o_cb);
    uint32_t* tot = (uint32_t*)at(o_tot);  // [0] inner total, [1] tri total, [2] max stack, [3] children
    k_tri_boxes<<<nb, BB, 0, st>>>(pos, idx, n, tbox, (uint32_t*)at(o_part));
    k_fold6<<<1, BB, 0, st>>>((uint32_t*)at(o_part), nb, cb);
    k_morton<<<nb, BB, 0, st>>>(tbox, n, cb, (uint32_t*)at(o_code), (uint32_t*)at(o_id));
    BCHK(hipGetLastError());
    BCHK(rocprim::radix_sort_pairs(at(o_sort), sort_bytes, (uint32_t*)at(o_code), (uint32_t*)at(o_code2),
                                   (uint32_t*)at(o_id), (uint32_t*)at(o_id2), n, 0, 30, st));
    const uint32_t* code = (uint32_t*)at(o_code2);
    const uint32_t* id = (uint32_t*)at(o_id2);
    BinTree T{(int32_t*)at(o_left), (int32_t*)at(o_right), (int32_t*)at(o_parent), (uint32_t*)at(o_first),
              (uint32_t*)at(o_count), (float*)at(o_box)};
    if (n > 1) k_karras<<<blocks(n - 1), BB, 0, st>>>(code, (int)n, T);
    k_leaf_boxes<<<nb, BB, 0, st>>>(tbox, id, (int)n, T);
    if (n > 1) {
        const float* sb = T.box + 6 * (size_t)(n - 1);  // leaf boxes in sorted order
        const int nblk = (int)((n + RB - 1) / RB);
        float* pre = (float*)at(o_pre);
        float* suf = (float*)at(o_suf);
        float* table = (float*)at(o_table);
        k_block_unions<<<blocks(n), BB, 0, st>>>(sb, (int)n, pre, suf, table);
        for (int k = 1; (1 << k) <= nblk; k++)
            k_table_level<<<blocks(nblk), BB, 0, st>>>(table + (size_t)6 * nblk * (k - 1), nblk, 1 << (k - 1),
                                                       table + (size_t)6 * nblk * k);
        k_node_boxes<<<blocks(n - 1), BB, 0, st>>>(sb, pre, suf, table, nblk, (int)n, T);
    }
    BCHK(hipGetLastError());

    // collapse, one level per iteration
    int32_t* q_in = (int32_t*)at(o_q0);
    int32_t* q_out = (int32_t*)at(o_q1);
    uint32_t* s_in = (uint32_t*)at(o_s0);
    uint32_t* s_out = (uint32_t*)at(o_s1);
    const int32_t root = 0;  // internal node 0, or leaf 0 (binary id n - 1 = 0) when n == 1
    BCHK(hipMemcpyAsync(q_in, &root, 4, hipMemcpyHostToDevice, st));
    BCHK(hipMemsetAsync(s_in, 0, 4, st));
    BCHK(hipMemsetAsync(tot + 2, 0, 4, st));
    uint32_t nq = 1, node_base = 0, tri_base = 0, depth = 0;
    uint64_t kids_total = 0;  // every child slot: inner children (all nodes but the root) and leaves
    while (nq > 0) {
        k_plan<<<blocks(nq), BB, 0, st>>>(q_in, nq, (int)n, T, (int32_t*)at(o_kids), (uint32_t*)at(o_nk),
                                          (uint32_t*)at(o_ni), (uint32_t*)at(o_nt));
        device_scan_u32(st, (uint32_t*)at(o_ni), (uint32_t*)at(o_io), nq, (uint32_t*)at(o_work), tot);
        device_scan_u32(st, (uint32_t*)at(o_nt), (uint32_t*)at(o_to), nq, (uint32_t*)at(o_work), tot + 1);
        device_scan_u32(st, (uint32_t*)at(o_nk), (uint32_t*)at(o_ni2), nq, (uint32_t*)at(o_work), tot + 3);
        Emit E{node_base, node_base + nq, tri_base};
        k_emit<<<blocks(nq), BB, 0, st>>>(q_in, nq, (int)n, T, (int32_t*)at(o_kids), (uint32_t*)at(o_nk),
                                          (uint32_t*)at(o_io), (uint32_t*)at(o_to), s_in, E, pos, idx, id, nodes,
                                          tris, q_out, s_out, tot + 2);
        BCHK(hipGetLastError());
        uint32_t h[4];
        BCHK(hipMemcpyAsync(h, tot, 16, hipMemcpyDeviceToHost, st));
        BCHK(hipStreamSynchronize(st));
        kids_total += h[3];
        node_base += nq;
        tri_base += h[1];
        depth++;
        nq = h[0];
        std::swap(q_in, q_out);
        std::swap(s_in, s_out);
    }
    uint32_t h[7];
    BCHK(hipMemcpyAsync(h, tot + 2, 4, hipMemcpyDeviceToHost, st));
    BCHK(hipMemcpyAsync(h + 1, T.box, 24, hipMemcpyDeviceToHost, st));  // root box (internal 0, or leaf 0)
    BCHK(hipStreamSynchronize(st));
    if (tri_base != n) return fail(AG_E_DEVICE, "GPU BVH build placed " + std::to_string(tri_base) + " of " +
                                                   std::to_string(n) + " triangles");
    info->n_nodes = node_base;
    info->n_leaves = (uint32_t)(kids_total - (node_base - 1));
    info->depth = depth;
    info->max_stack = h[0];
    for (int a = 0; a < 3; a++) {
        std::memcpy(&info->lo[a], &h[1 + a], 4);
        std::memcpy(&info->hi[a], &h[4 + a], 4);
    }
    return AG_OK;
}

}  // namespace ag
