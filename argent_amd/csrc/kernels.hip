// Wavefront path tracer and ray queries for gfx950.
//
// One wave of paths (up to `cap` path slots) runs as a fixed sequence of
// kernels per bounce: extend (closest hit over the BVH + panels + floor),
// shade (material-sorted: emission, BSDF, next-event sample, continuation),
// shadow (any-hit for the next-event rays).  Every kernel is a persistent
// grid-stride loop whose trip count it reads from a device counter, so the
// whole sequence launches without host round trips.  Queues are appended
// with one atomic per wavefront (ballot + mbcnt prefix).
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "dev_types.hpp"
#include "kernels.hpp"

namespace ag {

// ------------------------------------------------------------------ wave helpers
// An extension ray's w word: its view (bits 8-31) and depth (bits 0-7; depth
// never exceeds 64).  Traversal and shading take the view from here instead
// of a per-path array, a load scattered over the path slots; a shadow
// record's w word carries the view the same way (bits 1-31, bit 0 its kind).
AGD uint32_t ray_tag(uint32_t view, uint32_t depth) { return view << 8 | depth; }
AGD uint32_t tag_view(uint32_t tag) { return tag >> 8; }

AGD uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Runs f(u) for the active lanes once per distinct value u of `key` among
// them, with u in a wavefront-uniform register: what f reads through u (a
// view, its scene, panels, materials) comes in by scalar loads instead of one
// vector load per lane.  One pass when the lanes agree, the usual case.
template <class F>
AGD void per_uniform(uint32_t key, F&& f) {
    uint64_t todo = __ballot(1);
    while (todo) {
        int leader = __ffsll((unsigned long long)todo) - 1;
        uint32_t u = __builtin_amdgcn_readlane(key, leader);
        bool mine = key == u;
        // inside, key == u on every active lane: take it with readfirstlane so
        // the compiler keeps a scalar (it would otherwise reuse the per-lane key)
        if (mine) f(__builtin_amdgcn_readfirstlane(key));
        todo &= ~__ballot(mine);
    }
}

// Read *p through the constant address space.  The view, scene and material
// blocks are read this way wherever a kernel reads them per lane too: the
// loads are then known not to change while the kernel runs (no reload after
// the kernel's own stores), and a pointer read this way is known to point to
// global memory, so what is read through it takes global loads, not flat ones
// (which wait on the LDS and scalar counters as well and serialise).  With a wavefront-uniform p
// (inside per_uniform) the compiler then issues scalar loads into SGPRs; the
// scene and view blocks are read-only while a kernel runs.
template <class T>
AGD T uload(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(4))) T*)p;
#else
    return *p;  // host pass of this file: never executed
#endif
}

// Field by field through uload: a whole-struct copy through the constant
// space is lowered to flat loads, and the table pointers inside would lose
// their global address space.
// The spectral tables a shading block staged into LDS: a pointer into one of
// the scene blob's staged ranges reads the LDS copy (through the flat
// aperture, as every table pointer already is a flat one), any other the
// original.  W null: no window.
struct TabWin {
    const float* g;  // the staged range [g, g + n) of the scene's table blob
    uint32_t n;
    const float* lds;
};
AGD const float* tw(const TabWin* W, const float* p) {
    if (W) {
        const size_t off = (size_t)(p - W->g);  // wraps for pointers below g
        if (off < W->n) return W->lds + off;
    }
    return p;
}

AGD DMat load_mat(const DMat* p) {
    DMat m;
    m.kind = uload(&p->kind);
    m.alpha = uload(&p->alpha);
    m.eta = uload(&p->eta);
    m.tn = uload(&p->tn);
    m.tk = uload(&p->tk);
    m.has_patina = uload(&p->has_patina);
    m.p_strength = uload(&p->p_strength);
    m.p_lo = uload(&p->p_lo);
    m.p_hi = uload(&p->p_hi);
    m.p_alpha = uload(&p->p_alpha);
    m.p_eta = uload(&p->p_eta);
    m.tpatina = uload(&p->tpatina);
    m.ta = uload(&p->ta);
    m.tb = uload(&p->tb);
    m.opal_cell = uload(&p->opal_cell);
    m.opal_strength = uload(&p->opal_strength);
    return m;
}
AGD DPanel load_panel(const DPanel* p, const TabWin* W = nullptr) {
    DPanel q;
    for (int i = 0; i < 3; i++) {
        q.c[i] = uload(&p->c[i]);
        q.hu[i] = uload(&p->hu[i]);
        q.hv[i] = uload(&p->hv[i]);
        q.n[i] = uload(&p->n[i]);
    }
    q.area = uload(&p->area);
    q.hu2 = uload(&p->hu2);
    q.hv2 = uload(&p->hv2);
    q.emit = tw(W, uload(&p->emit));
    q.camera_visible = uload(&p->camera_visible);
    q.feather = uload(&p->feather);
    return q;
}

// Append for every lane with pred set; returns the slot (valid where pred).
// Must be reached by all lanes of the wavefront.
AGD uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    uint32_t prefix = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    uint32_t cnt = (uint32_t)__popcll(mask);
    uint32_t base = 0;
    if (cnt) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        if ((int)lane_id() == leader) base = atomicAdd(counter, cnt);
        base = __shfl(base, leader);
    }
    return base + prefix;
}

// Append into one of several counters by key; lanes with key < 0 skip.
AGD uint32_t wave_append_keyed(uint32_t* counters, int key) {
    uint32_t slot = 0;
    uint64_t todo = __ballot(key >= 0);
    while (todo) {
        int leader = __ffsll((unsigned long long)todo) - 1;
        int k = __shfl(key, leader);
        bool mine = key == k && ((todo >> lane_id()) & 1ull);
        uint64_t mask = __ballot(mine);
        uint32_t prefix = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        uint32_t base = 0;
        if ((int)lane_id() == leader) base = atomicAdd(&counters[k * CNT_PAD], (uint32_t)__popcll(mask));
        base = __shfl(base, leader);
        if (mine) slot = base + prefix;
        todo &= ~mask;
    }
    return slot;
}

// ------------------------------------------------------------------ BVH traversal
// Bvh::intersect / Bvh::occluded (bvh.rs:159-245) over the compressed 8-wide device BVH.
// Per-lane stack: STACK_LDS entries in LDS, the rest in a per-thread HBM spill
// area.  Box tests are widened by a relative 2^-16 on both ends: enough for
// the box test's own rounding and for the f32 Moller-Trumbore hits that land
// slightly outside their triangle on grazing rays, so the answer does not
// depend on how the tree groups triangles (except for a rare grazing ray
// whose hit lies outside its triangle by more than the widening).  The triangle test is
// Moller-Trumbore evaluated exactly as the reference does (no contraction),
// with ties on t resolved to the lower original triangle index.
// The stack's two homes carry their address spaces in the pointer types, so
// a pop is an LDS read or a global read: with plain pointers the compiler
// merges the two branches into one flat load, which waits on both counters.
struct TravStack {
    AG_LDS uint2* ent;      // LDS [STACK_LDS][blockDim]: (entry, entry distance bits), one 64-bit access each
    AG_GLOBAL uint2* spill; // HBM [STACK_SPILL] for this thread
    uint32_t tid, nthreads;
};

AGD void st_push(const TravStack& s, int& sp, uint32_t e, float t) {
    if (sp < STACK_LDS) {
        s.ent[sp * s.nthreads + s.tid] = make_uint2(e, __float_as_uint(t));
    } else if (sp < STACK_LDS + STACK_SPILL) {
        s.spill[sp - STACK_LDS] = make_uint2(e, __float_as_uint(t));
    } else {
        return;  // full: unreachable, the host rejects trees deeper than the stack
    }
    sp++;
}
AGD void st_pop(const TravStack& s, int& sp, uint32_t& e, float& t) {
    sp--;
    uint2 v;
    if (sp < STACK_LDS)
        v = s.ent[sp * s.nthreads + s.tid];
    else
        v = s.spill[sp - STACK_LDS];
    e = v.x;
    t = __uint_as_float(v.y);
}

constexpr float BOX_LO = 1.0f - 1.52587890625e-5f;  // 1 - 2^-16
constexpr float BOX_HI = 1.0f + 1.52587890625e-5f;

// Node layout (80 B): see the host builder.  Children's boxes are
// p + q * 2^(e-127) per plane (an exact fma: the quantized box contains the
// exact one as evaluated here); slab test as bvh.rs:280-299 on each.
#ifndef AG_NODE_FLOATS
#define AG_NODE_FLOATS 20
#endif
constexpr int NODE_F4 = AG_NODE_FLOATS / 4;  // node stride in float4 (the node itself is 5 float4)

// What a node holds to name child k: the first inner child and first leaf
// triangle, the 8 child bytes (count << 5 | offset) and the inner-child mask.
struct NodeKids {
    uint32_t inner_base, leaf_base, mlo, mhi, imask;
};

// Child k's stack entry: an inner node index, or a leaf (bit 31, count - 1
// in bits 27-30, first triangle below).
AGD uint32_t child_entry(const NodeKids& K, uint32_t k) {
    uint32_t m = ((k < 4u ? K.mlo : K.mhi) >> ((k & 3u) * 8u)) & 0xFFu;
    bool inner = (K.imask >> k) & 1u;
    return inner ? K.inner_base + m : (0x80000000u | (((m >> 5) - 1u) << 27) | (K.leaf_base + (m & 31u)));
}

// Slab tests of a node's 8 children: tm[k] = child k's entry distance, or
// INF_F for a miss or an empty slot.
AGD void slab8_t(const float4* __restrict__ nd, f3 o, f3 inv, float tmax, float tm[8], NodeKids& K) {
    float4 n0 = nd[0];
    uint4 n1 = *reinterpret_cast<const uint4*>(&nd[1]);
    uint4 qa = *reinterpret_cast<const uint4*>(&nd[2]);
    uint4 qb = *reinterpret_cast<const uint4*>(&nd[3]);
    uint4 qc = *reinterpret_cast<const uint4*>(&nd[4]);
    uint32_t w = __float_as_uint(n0.w);
    float sx = __uint_as_float((w & 0xFFu) << 23);
    float sy = __uint_as_float(((w >> 8) & 0xFFu) << 23);
    float sz = __uint_as_float(((w >> 16) & 0xFFu) << 23);
    uint32_t imask = w >> 24;
    K = NodeKids{n1.x, n1.y, n1.z, n1.w, imask};
    if (fabsf(inv.x) < INF_F && fabsf(inv.y) < INF_F && fabsf(inv.z) < INF_F) {
        // finite 1/d: each axis's near and far planes follow the direction's
        // sign, chosen once for the node (min/max of the two products would
        // pick the same values)
        const bool px = inv.x >= 0.0f, py = inv.y >= 0.0f, pz = inv.z >= 0.0f;
        const uint32_t nx0 = px ? qa.x : qb.z, nx1 = px ? qa.y : qb.w, fx0 = px ? qb.z : qa.x, fx1 = px ? qb.w : qa.y;
        const uint32_t ny0 = py ? qa.z : qc.x, ny1 = py ? qa.w : qc.y, fy0 = py ? qc.x : qa.z, fy1 = py ? qc.y : qa.w;
        const uint32_t nz0 = pz ? qb.x : qc.z, nz1 = pz ? qb.y : qc.w, fz0 = pz ? qc.z : qb.x, fz1 = pz ? qc.w : qb.y;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int sh = (k & 3) * 8;
            const bool hiw = k >= 4;
            uint32_t m = ((hiw ? n1.w : n1.z) >> sh) & 0xFFu;
            float tnx = (fmaf((float)(((hiw ? nx1 : nx0) >> sh) & 0xFFu), sx, n0.x) - o.x) * inv.x;
            float tfx = (fmaf((float)(((hiw ? fx1 : fx0) >> sh) & 0xFFu), sx, n0.x) - o.x) * inv.x;
            float tny = (fmaf((float)(((hiw ? ny1 : ny0) >> sh) & 0xFFu), sy, n0.y) - o.y) * inv.y;
            float tfy = (fmaf((float)(((hiw ? fy1 : fy0) >> sh) & 0xFFu), sy, n0.y) - o.y) * inv.y;
            float tnz = (fmaf((float)(((hiw ? nz1 : nz0) >> sh) & 0xFFu), sz, n0.z) - o.z) * inv.z;
            float tfz = (fmaf((float)(((hiw ? fz1 : fz0) >> sh) & 0xFFu), sz, n0.z) - o.z) * inv.z;
            bool inner = (imask >> k) & 1u;
            uint32_t cnt = m >> 5;
            float tmin = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.0f));
            float tmx = fminf(fminf(tfx, tfy), tfz);
            tmin *= BOX_LO;
            tmx = fminf(tmx * BOX_HI, tmax);
            bool hit = (inner || cnt != 0u) && tmin <= tmx;
            tm[k] = hit ? tmin : INF_F;
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int sh = (k & 3) * 8;
        const bool hiw = k >= 4;
        uint32_t m = ((hiw ? n1.w : n1.z) >> sh) & 0xFFu;
        float lx = fmaf((float)(((hiw ? qa.y : qa.x) >> sh) & 0xFFu), sx, n0.x);
        float ly = fmaf((float)(((hiw ? qa.w : qa.z) >> sh) & 0xFFu), sy, n0.y);
        float lz = fmaf((float)(((hiw ? qb.y : qb.x) >> sh) & 0xFFu), sz, n0.z);
        float hx = fmaf((float)(((hiw ? qb.w : qb.z) >> sh) & 0xFFu), sx, n0.x);
        float hy = fmaf((float)(((hiw ? qc.y : qc.x) >> sh) & 0xFFu), sy, n0.y);
        float hz = fmaf((float)(((hiw ? qc.w : qc.z) >> sh) & 0xFFu), sz, n0.z);
        bool inner = (imask >> k) & 1u;
        uint32_t cnt = m >> 5;
        float tx0 = (lx - o.x) * inv.x, tx1 = (hx - o.x) * inv.x;
        float ty0 = (ly - o.y) * inv.y, ty1 = (hy - o.y) * inv.y;
        float tz0 = (lz - o.z) * inv.z, tz1 = (hz - o.z) * inv.z;
        float tmin = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        float tmx = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
        tmin *= BOX_LO;
        tmx = fminf(tmx * BOX_HI, tmax);
        bool hit = (inner || cnt != 0u) && tmin <= tmx;
        tm[k] = hit ? tmin : INF_F;
    }
}

// bvh.rs:391-409, returning t (or -1) and u, v
AGD float tri_test(const float4* tp, f3 o, f3 d, float* uo, float* vo, uint32_t* prim) {
    float4 a = tp[0], b = tp[1], c = tp[2];
    f3 v0{a.x, a.y, a.z}, e1{a.w, b.x, b.y}, e2{b.z, b.w, c.x};
    *prim = __float_as_uint(c.y);
    f3 p = cross(d, e2);
    float det = dot(e1, p);
    if (fabsf(det) < 1e-14f) return -1.0f;
    float inv_det = 1.0f / det;
    f3 s = o - v0;
    float u = dot(s, p) * inv_det;
    if (!(u >= 0.0f && u <= 1.0f)) return -1.0f;
    f3 q = cross(s, e1);
    float v = dot(d, q) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return -1.0f;
    float t = dot(e2, q) * inv_det;
    *uo = u;
    *vo = v;
    return t;
}

// ------------------------------------------------------------------ persistent traversal
// Rays are pulled by wavefronts, not assigned to threads.  A queue of n rays
// is cut into CURSOR_SLICES slices, one per XCD (blocks are dealt to the XCDs
// round-robin, so block b starts on slice b % 8), each with its own cursor;
// a wavefront takes GRAB rays from a slice per atomic and hands them to its
// idle lanes whenever at least refill_min lanes are idle, moving on to the
// next slice when its own is spent.  Every busy lane advances by one inner
// node or one leaf per iteration, and the lanes that finished are handed to
// the job together.  A lane's ray therefore never waits for the slowest ray
// of a static batch of 64, and the atomics stay few and spread over lines.
constexpr uint32_t NO_RAY = 0xFFFFFFFFu;
constexpr uint32_t TRAV_DONE = 0xFFFFFFFEu;  // never an entry: leaf entries stay below 0xA0000000
#ifndef AG_GRAB
#define AG_GRAB 64
#endif
#ifndef AG_TRAV_WAVES_PER_EU
#define AG_TRAV_WAVES_PER_EU 0
#endif
constexpr uint32_t GRAB = AG_GRAB;              // rays taken per cursor atomic

AGD uint32_t lane_prefix(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Queue appends of one wavefront that reserve slots `chunk` at a time (0:
// one reservation per append, exact).  Called by every lane of the wavefront
// at a converged point, so the chunk state stays wavefront-uniform.
struct ChunkAppend {
    uint32_t* counter;
    uint32_t chunk;
    uint32_t base, used;  // the current chunk [base, base + chunk) and the slots taken from it
    bool have;
    AGD uint32_t append(bool pred) {
        const uint64_t m = __ballot(pred);
        const uint32_t k = (uint32_t)__popcll(m), prefix = lane_prefix(m);
        if (k == 0) return 0;
        if (chunk == 0) {
            uint32_t b = 0;
            if (lane_id() == 0) b = atomicAdd(counter, k);
            return __shfl(b, 0) + prefix;
        }
        const uint32_t rem = have ? chunk - used : 0u;
        if (k <= rem) {
            const uint32_t s = base + used + prefix;
            used += k;
            return s;
        }
        // the rest of this chunk, then a new one
        uint32_t nb = 0;
        if (lane_id() == 0) nb = atomicAdd(counter, chunk);
        nb = __shfl(nb, 0);
        const uint32_t s = prefix < rem ? base + used + prefix : nb + (prefix - rem);
        base = nb;
        used = k - rem;
        have = true;
        return s;
    }
    // mark the unused tail of the last chunk dead (w word = DEAD_PATH)
    AGD void close(float4* q, uint32_t qcap) {
        if (chunk == 0 || !have) return;
        for (uint32_t e = used + lane_id(); e < chunk; e += 64)
            if (base + e < qcap) q[base + e] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(DEAD_PATH));
    }
};


// Wavefront-uniform state of the ray supply.
struct RaySupply {
    uint32_t n;
    uint32_t* cursors;
    uint32_t lo, hi;      // grabbed range not yet handed out
    uint32_t slice, tried;
    uint32_t grab;        // rays per cursor atomic (GRAB, or fewer for a spread queue)
    uint32_t slices;      // 1..CURSOR_SLICES
};

AGD uint32_t slice_begin(uint32_t n, uint32_t s, uint32_t slices) { return (uint32_t)(((uint64_t)n * s) / slices); }

// Up to `want` consecutive ray indices from *first; 0 once every slice is
// spent.  Reached by all lanes of the wavefront.
AGD uint32_t supply_take(RaySupply& R, uint32_t want, uint32_t* first) {
    while (R.lo == R.hi) {
        if (R.tried >= R.slices) return 0;
        uint32_t b = slice_begin(R.n, R.slice, R.slices), e = slice_begin(R.n, R.slice + 1, R.slices);
        uint32_t got = e - b;  // spent unless the cursor says otherwise
        if (lane_id() == 0) {
            // Own slice: straight to the atomic.  Another wavefront's slice: a
            // cursor only grows, so a plain read past the slice proves it spent
            // without an atomic on a contended line (every wavefront probes the
            // other slices once the queue runs dry).
            uint32_t* cur = R.cursors + R.slice * CURSOR_STRIDE;
            if (R.tried == 0 || __hip_atomic_load(cur, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < e - b)
                got = atomicAdd(cur, R.grab);
        }
        got = __shfl(got, 0);
        if (got < e - b) {
            R.lo = b + got;
            R.hi = b + min(got + R.grab, e - b);
        } else {
            R.slice = (R.slice + 1) % R.slices;
            R.tried++;
        }
    }
    uint32_t k = min(want, R.hi - R.lo);
    *first = R.lo;
    R.lo += k;
    return k;
}

struct TravLane {
    const float4* nodes;
    const float4* tris;
    f3 o, d, inv;
    float t;        // closest hit: best t so far (starts at t_max); any hit: t_max
    float u, v;
    uint32_t prim;  // closest hit: original triangle or NO_RAY; any hit: 1 once occluded
    uint32_t cur;   // entry to visit next, or TRAV_DONE
    int sp;
    uint32_t aux;   // the job's own (view index)
};

AGD void lane_begin(TravLane& L, const float4* nodes, const float4* tris, f3 o, f3 d, float tmax) {
    L.nodes = nodes;
    L.tris = tris;
    L.o = o;
    L.d = d;
    L.inv = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    L.t = tmax;
    L.u = 0.0f;
    L.v = 0.0f;
    L.prim = NO_RAY;
    L.cur = 0;  // the root is always an inner node
    L.sp = 0;
}

// Bvh::intersect (bvh.rs:159-213), one step: an inner node pushes its hit
// children far to near and continues with the nearest; a leaf tests its
// triangles (ties on t go to the lower original index).  Then the next stack
// entry not beyond the best hit, or done.
AGD void step_closest(TravLane& L, const TravStack& st) {
    if (!(L.cur & 0x80000000u)) {
        float tm[8];
        NodeKids K;
        slab8_t(L.nodes + (size_t)L.cur * NODE_F4, L.o, L.inv, L.t, tm, K);
        // Sort keys: the entry distance (>= 0, so its bits order as integers)
        // with the child slot in the 3 low mantissa bits; a miss sorts last.
        // The distance pushed is the key's, rounded down by at most 7 ulp:
        // a box is only ever visited more readily, and the closest hit
        // (least t, then least triangle id) does not depend on visit order.
        uint32_t key[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            key[k] = tm[k] != INF_F ? ((__float_as_uint(tm[k]) & 0x7FFFFFF8u) | (uint32_t)k) : 0xFFFFFFFFu;
        // ascending: the 19-comparator sorting network for 8 keys
#define KSWAP(i, j)                                             \
    {                                                           \
        uint32_t a = min(key[i], key[j]), b = max(key[i], key[j]); \
        key[i] = a;                                             \
        key[j] = b;                                             \
    }
        KSWAP(0, 2) KSWAP(1, 3) KSWAP(4, 6) KSWAP(5, 7) KSWAP(0, 4) KSWAP(1, 5) KSWAP(2, 6) KSWAP(3, 7)
        KSWAP(0, 1) KSWAP(2, 3) KSWAP(4, 5) KSWAP(6, 7) KSWAP(2, 4) KSWAP(3, 5) KSWAP(1, 4) KSWAP(3, 6)
        KSWAP(1, 2) KSWAP(3, 4) KSWAP(5, 6)
#undef KSWAP
        if (key[0] != 0xFFFFFFFFu) {
#pragma unroll
            for (int k = 7; k >= 1; k--)
                if (key[k] != 0xFFFFFFFFu)
                    st_push(st, L.sp, child_entry(K, key[k] & 7u), __uint_as_float(key[k] & ~7u));
            L.cur = child_entry(K, key[0] & 7u);
            return;
        }
    } else {
        uint32_t first = L.cur & 0x07FFFFFFu;
        uint32_t cnt = ((L.cur >> 27) & 0xFu) + 1u;
        for (uint32_t k = 0; k < cnt; k++) {
            float u, v;
            uint32_t prim;
            float t = tri_test(L.tris + (size_t)(first + k) * 3, L.o, L.d, &u, &v, &prim);
            if (t > 1e-7f && (t < L.t || (t == L.t && L.prim != NO_RAY && prim < L.prim))) {
                L.t = t;
                L.u = u;
                L.v = v;
                L.prim = prim;
            }
        }
    }
    L.cur = TRAV_DONE;
    while (L.sp > 0) {
        uint32_t e;
        float tn;
        st_pop(st, L.sp, e, tn);
        if (!(tn > L.t)) {
            L.cur = e;
            break;
        }
    }
}

// Bvh::occluded (bvh.rs:216-245), one step: no ordering, stop at the first
// triangle hit in (1e-7, t_max).
AGD void step_any(TravLane& L, const TravStack& st) {
    if (!(L.cur & 0x80000000u)) {
        float tm[8];
        NodeKids K;
        slab8_t(L.nodes + (size_t)L.cur * NODE_F4, L.o, L.inv, L.t, tm, K);
        // every child's entry up front: decoding inside the branches below
        // measured slower (7.57 vs 7.87 G rays/s on an incoherent batch)
        uint32_t ch[8];
#pragma unroll
        for (int k = 0; k < 8; k++) ch[k] = child_entry(K, (uint32_t)k);
        bool any = false;
        uint32_t next = 0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            if (tm[k] != INF_F) {
                if (!any) {
                    any = true;
                    next = ch[k];
                } else {
                    st_push(st, L.sp, ch[k], tm[k]);
                }
            }
        }
        if (any) {
            L.cur = next;
            return;
        }
    } else {
        uint32_t first = L.cur & 0x07FFFFFFu;
        uint32_t cnt = ((L.cur >> 27) & 0xFu) + 1u;
        for (uint32_t k = 0; k < cnt; k++) {
            float u, v;
            uint32_t prim;
            float t = tri_test(L.tris + (size_t)(first + k) * 3, L.o, L.d, &u, &v, &prim);
            if (t > 1e-7f && t < L.t) {
                L.prim = 1;
                L.cur = TRAV_DONE;
                return;
            }
        }
    }
    if (L.sp > 0) {
        float tn;
        st_pop(st, L.sp, L.cur, tn);
    } else {
        L.cur = TRAV_DONE;
    }
}

// Job: bool begin(uint32_t i, TravLane& L) sets up ray i (false: nothing to
// trace, finish at once); void finish(uint32_t i, const TravLane& L, bool mine)
// is called by every lane of the wavefront, `mine` where ray i just finished;
// void drain() once by every lane after the last ray.
//
// A queue too short to give every wavefront of the grid 64 rays is spread
// over all of them: each takes `lanes` = ceil(n / wavefronts) rays at a time
// (at least min_lanes, at most 64) in its lowest lanes.  A launch over a few
// thousand rays then lasts about as long as its slowest ray instead of its
// slowest batch of 64 divergent ones, and every CU takes part.  Which rays
// share a wavefront does not change any ray's result.
template <bool ANY, class Job>
AGD void traverse_persistent(Job& J, uint32_t n, uint32_t* cursors, uint32_t refill_min, const TravStack& st,
                             uint32_t min_lanes = 64u, uint32_t slices = CURSOR_SLICES) {
    TravLane L;
    L.cur = TRAV_DONE;
    L.sp = 0;
    uint32_t idx = NO_RAY;
    const uint32_t nwaves = gridDim.x * (blockDim.x / 64u);
    uint32_t lanes = (uint32_t)(((uint64_t)n + nwaves - 1u) / nwaves);
    lanes = min(max(lanes, min_lanes), 64u);
    const uint32_t grab = lanes < 64u ? lanes : GRAB;
    const bool lane_ok = lane_id() < lanes;
    const uint32_t refill = min(refill_min, lanes);
    RaySupply R{n, cursors, 0u, 0u, blockIdx.x % slices, 0u, grab, slices};
    // wavefronts beyond what n rays can occupy (`grab` each) leave at once
    uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;
    // uniform across the wavefront; 64-bit so that n near 2^32 cannot wrap
    bool more = (uint64_t)wave < ((uint64_t)n + grab - 1u) / grab;
    // A lane whose ray has finished traversing stays parked (idx kept, L.cur
    // == TRAV_DONE) until the next refill, when all parked lanes are finished
    // in one call: a job's finish reads memory (a material id, a radiance
    // slot), and one round trip per refill costs far less than one per step
    // at which some lane happened to finish.
    for (;;) {
        bool parked = idx != NO_RAY && L.cur == TRAV_DONE;
        if (more && (uint32_t)__popcll(__ballot(lane_ok && (idx == NO_RAY || parked))) >= refill) {
            if (__ballot(parked) != 0) {
                J.finish(idx, L, parked);
                if (parked) idx = NO_RAY;
            }
            // Fill the idle lanes with the next rays in queue order, taking
            // again while lanes stay idle (a take returns at most what is left
            // of the wavefront's current grab); a ray its job settles without
            // the piece finishes at once and its lane takes the next one, so
            // traversal batches hold only rays that need it.
            for (;;) {
                uint64_t idle = __ballot(lane_ok && idx == NO_RAY);
                uint32_t cnt = (uint32_t)__popcll(idle);
                if (cnt == 0) break;
                uint32_t first = 0;
                uint32_t k = supply_take(R, cnt, &first);
                if (k == 0) {
                    more = false;
                    break;
                }
                bool fin = false;
                if (lane_ok && idx == NO_RAY) {
                    uint32_t r = lane_prefix(idle);
                    if (r < k) {
                        idx = first + r;
                        if (!J.begin(idx, L)) {
                            L.cur = TRAV_DONE;
                            fin = true;
                        }
                    }
                }
                if (__ballot(fin) != 0) {
                    J.finish(idx, L, fin);
                    if (fin) idx = NO_RAY;
                } else if (k == cnt) {
                    break;  // every idle lane has a ray to trace
                }
            }
        }
        if (__ballot(idx != NO_RAY && L.cur != TRAV_DONE) == 0) {
            // nothing traversing: finish the parked lanes; leave once the
            // queue is spent, else go round to refill every lane
            bool last = idx != NO_RAY;
            if (__ballot(last) != 0) {
                J.finish(idx, L, last);
                if (last) idx = NO_RAY;
            }
            if (!more) break;
            continue;
        }
        if (idx != NO_RAY && L.cur != TRAV_DONE) {
            if (ANY)
                step_any(L, st);
            else
                step_closest(L, st);
        }
    }
    J.drain();
}

// Traversal kernels keep AG_TRAV_WAVES_PER_EU wavefronts per SIMD (8: 64
// VGPRs): latency hiding matters more to them than registers.
#if AG_TRAV_WAVES_PER_EU > 0
#define TRAV_WAVES __attribute__((amdgpu_waves_per_eu(AG_TRAV_WAVES_PER_EU, AG_TRAV_WAVES_PER_EU)))
#else
#define TRAV_WAVES
#endif

#define TRAV_STACK_DECL                                                     \
    __shared__ uint2 s_ent[STACK_LDS * TRAV_BLOCK];                         \
    TravStack st;                                                           \
    st.ent = (AG_LDS uint2*)s_ent;                                          \
    st.tid = threadIdx.x;                                                   \
    st.nthreads = blockDim.x;                                               \
    st.spill = (AG_GLOBAL uint2*)(spill + (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * STACK_SPILL);

// ------------------------------------------------------------------ ray-batch kernels
struct BatchJob {
    const float4* nodes;
    const float4* tris;
    RaysSoA rays;
    HitsSoA hits;
    uint8_t* occ;
    AGD bool begin(uint32_t i, TravLane& L) {
        f3 o{rays.ox[i], rays.oy[i], rays.oz[i]}, d{rays.dx[i], rays.dy[i], rays.dz[i]};
        lane_begin(L, nodes, tris, o, d, rays.tmax ? rays.tmax[i] : INF_F);
        return true;
    }
    AGD void finish(uint32_t i, const TravLane& L, bool mine) {
        if (!mine) return;
        if (occ) {
            occ[i] = L.prim != NO_RAY ? 1 : 0;
            return;
        }
        bool hit = L.prim != NO_RAY;
        hits.t[i] = hit ? L.t : INF_F;
        hits.u[i] = hit ? L.u : 0.0f;
        hits.v[i] = hit ? L.v : 0.0f;
        hits.prim[i] = L.prim;
    }
    AGD void drain() {}
};

__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_rays_intersect(const float4* __restrict__ nodes,
                                                               const float4* __restrict__ tris, RaysSoA rays,
                                                               HitsSoA hits, uint32_t n, uint32_t* cursor,
                                                               uint32_t refill, uint32_t spread, uint2* spill) {
    TRAV_STACK_DECL
    BatchJob J{nodes, tris, rays, hits, nullptr};
    traverse_persistent<false>(J, n, cursor, refill, st, spread);
}

__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_rays_occluded(const float4* __restrict__ nodes,
                                                              const float4* __restrict__ tris, RaysSoA rays,
                                                              uint8_t* out, uint32_t n, uint32_t* cursor,
                                                              uint32_t refill, uint32_t spread, uint2* spill) {
    TRAV_STACK_DECL
    BatchJob J{nodes, tris, rays, HitsSoA{}, out};
    traverse_persistent<true>(J, n, cursor, refill, st, spread);
}

// ------------------------------------------------------------------ cavity ambient occlusion
// cavity.rs:95-128 (ambient_occlusion) with the piece's own any-hit query
// (scene.rs:40-44 -> bvh.rs:216-245): ray i is sample k = i % samples of
// vertex v = i / samples, generated here rather than read from memory: a
// cosine-weighted direction around the vertex normal, stratified in u1 with a
// per-vertex jitter and a golden-ratio sequence in u2 rotated per vertex,
// leaving the vertex at p + n eps, reaching `dist`.  A blocked ray adds one to
// its vertex's count.
struct AoJob {
    const float4* nodes;
    const float4* tris;
    const float* pos;
    const float* nrm;
    uint32_t samples;
    float dist, eps;
    uint32_t* counts;
    AGD bool begin(uint32_t i, TravLane& L) {
        const uint32_t v = i / samples, k = i - v * samples;
        L.aux = v;
        f3 n{nrm[3 * (size_t)v], nrm[3 * (size_t)v + 1], nrm[3 * (size_t)v + 2]};
        // Vec3::normalize_or_zero: times the reciprocal length, zero unless finite and positive
        const float rcp = 1.0f / sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
        if (!(isfinite(rcp) && rcp > 0.0f)) {
            L.prim = NO_RAY;  // no normal: never occluded
            return false;
        }
        n = n * rcp;
        const Frame fr = make_frame(n);  // Vec3::any_orthonormal_pair (Duff et al. 2017)
        const f3 o = f3{pos[3 * (size_t)v], pos[3 * (size_t)v + 1], pos[3 * (size_t)v + 2]} + n * eps;
        const uint32_t h = hash32(v ^ 0x5bd1e995u);
        const float rot = (float)(h >> 8) * (1.0f / 16777216.0f);
        const float jit = (float)(hash32(h) >> 8) * (1.0f / 16777216.0f);
        const float u1 = fminf(((float)k + jit) / (float)samples, 0.999999f);
        const float x2 = (float)k * 0.618034f + rot;
        const float u2 = x2 - truncf(x2);
        const float r = sqrtf(u1), phi = 6.28318530717958647692f * u2;
        const f3 d = (fr.t * (r * cosf(phi)) + fr.b * (r * sinf(phi))) + n * sqrtf(1.0f - u1);
        lane_begin(L, nodes, tris, o, d, dist);
        return true;
    }
    AGD void finish(uint32_t i, const TravLane& L, bool mine) {
        if (mine && L.prim != NO_RAY) atomicAdd(&counts[L.aux], 1u);
    }
    AGD void drain() {}
};

__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_ao(AoJob J, uint32_t n, uint32_t* cursor, uint32_t refill,
                                                            uint32_t spread, uint2* spill) {
    TRAV_STACK_DECL
    traverse_persistent<true>(J, n, cursor, refill, st, spread);
}

// ------------------------------------------------------------------ scene pieces
// fabsf(x / y) <= 1 with x / y rounded to nearest, as the reference writes it
// (light.rs:60-62), without dividing unless |x| lies within 2^-20 of y.  For
// y in [2^-100, 2^100]: |x| <= fl(y (1 - 2^-20)) gives |x| / y < 1, and
// |x| > fl(y (1 + 2^-20)) gives |x| / y > 1 + 2^-21, whose rounding exceeds 1.
// A NaN x fails both and divides to NaN, as in the reference.
AGD bool within_unit(float x, float y) {
    const float ax = fabsf(x);
    if (y >= 0x1p-100f && y <= 0x1p100f) {
        if (ax <= y * (1.0f - 0x1p-20f)) return true;
        if (ax > y * (1.0f + 0x1p-20f)) return false;
    }
    return fabsf(x / y) <= 1.0f;
}

// light.rs:49-63.  P points into the scene block; with a wavefront-uniform P
// every field comes in by a scalar load (a whole-struct copy of DPanel is not
// reliably kept scalar by the compiler and turns into vector loads).
AGD bool panel_intersect(const DPanel* P, f3 o, f3 d, float tmax, float* tout) {
    const f3 n{uload(&P->n[0]), uload(&P->n[1]), uload(&P->n[2])};
    float denom = dot(d, n);
    if (fabsf(denom) < 1e-9f) return false;
    const f3 c{uload(&P->c[0]), uload(&P->c[1]), uload(&P->c[2])};
    float num = dot(c - o, n);
    // t = num / denom > 1e-6 needs num and denom of one sign: reject the
    // planes behind the ray before paying for the division (same result,
    // NaN included: a NaN num either fails here or divides to NaN below)
    if ((num > 0.0f) != (denom > 0.0f) || num == 0.0f) return false;
    float t = num / denom;
    if (t <= 1e-6f || t >= tmax) return false;
    f3 p = (o + d * t) - c;
    const f3 hu{uload(&P->hu[0]), uload(&P->hu[1]), uload(&P->hu[2])};
    const f3 hv{uload(&P->hv[0]), uload(&P->hv[1]), uload(&P->hv[2])};
    if (within_unit(dot(p, hu), uload(&P->hu2)) && within_unit(dot(p, hv), uload(&P->hv2))) {
        *tout = t;
        return true;
    }
    return false;
}
// light.rs:74-94 (Panel::profile)
AGD float panel_profile(const DPanel& P, f3 p) {
    if (P.feather <= 0.0f) return 1.0f;
    f3 q = p - ld3(P.c);
    float a = fminf(fabsf(dot(q, ld3(P.hu)) / P.hu2), 1.0f);
    float b = fminf(fabsf(dot(q, ld3(P.hv)) / P.hv2), 1.0f);
    float start = 1.0f - P.feather;
    float fa = 1.0f, fb = 1.0f;
    if (a > start) {
        float t = (a - start) / P.feather;
        fa = 1.0f - 0.75f * t * t * (3.0f - 2.0f * t);
    }
    if (b > start) {
        float t = (b - start) / P.feather;
        fb = 1.0f - 0.75f * t * t * (3.0f - 2.0f * t);
    }
    return fa * fb;
}
// light.rs:65-73 (Panel::le)
AGD Spec panel_le(const DPanel& P, f3 p, f3 dir, const Spec& lam) {
    if (P.emit && dot(dir, ld3(P.n)) < 0.0f) return tab_sample(P.emit, lam) * panel_profile(P, p);
    return splat(0.0f);
}
// light.rs:133-162
AGD bool floor_intersect(const DScene& S, f3 o, f3 d, float tmax, float* tout) {
    if (fabsf(d.y) < 1e-9f) return false;
    float t = (uload(&S.floor_y) - o.y) / d.y;
    if (t > 1e-5f && t < tmax) {
        *tout = t;
        return true;
    }
    return false;
}
AGD float floor_presence(const DScene& S, f3 p) {
    float r = sqrtf(p.x * p.x + p.z * p.z);
    const float f0 = uload(&S.floor_f0), f1 = uload(&S.floor_f1);
    if (r <= f0) return 1.0f;
    if (r >= f1) return 0.0f;
    float t = (r - f0) / (f1 - f0);
    return 1.0f - t * t * (3.0f - 2.0f * t);
}

// envmap.rs:77-88 LinearRgb::eval
AGD float linrgb_eval(const float* basis, const float* w, float l) {
    float x = fminf(fmaxf(l - 360.0f, 0.0f), 469.999f);
    int i = (int)x;
    float f = x - (float)i;
    float v = 0.0f;
    for (int c = 0; c < 3; c++) {
        const float* b = basis + c * AG_SPEC_N;
        v += w[c] * (b[i] * (1.0f - f) + b[i + 1] * f);
    }
    return fmaxf(v, 0.0f);
}
// envmap.rs:182-204
AGD void env_uv(const DScene& S, f3 dir, float* u, float* v) {
    f3 d = rot_y(dir, uload(&S.env_rsn), uload(&S.env_rcn));
    float uu = 0.5f + atan2f(d.x, -d.z) / (2.0f * PI_F);
    *v = acosf(fminf(fmaxf(d.y, -1.0f), 1.0f)) / PI_F;
    float r = fmodf(uu, 1.0f);
    *u = r < 0.0f ? r + 1.0f : r;
}
AGD f3 env_dir(const DScene& S, float u, float v) {
    float phi = (u - 0.5f) * 2.0f * PI_F;
    float theta = v * PI_F;
    float st = sinf(theta), ct = cosf(theta);
    f3 d{st * sinf(phi), ct, -st * cosf(phi)};
    return rot_y(d, uload(&S.env_rs), uload(&S.env_rc));
}
AGD float env_map_pdf(const DScene& S, f3 dir) {
    float u, v;
    env_uv(S, dir, &u, &v);
    float st = sinf(PI_F * v);
    if (st <= 1e-6f) return 0.0f;
    const int ew = uload(&S.env_w), eh = uload(&S.env_h);
    const float* marg = uload(&S.env_marg);
    int x = min((int)(u * (float)ew), ew - 1);
    int y = min((int)(v * (float)eh), eh - 1);
    const float* row = uload(&S.env_cond) + (size_t)y * (ew + 1);
    float pc = (row[x + 1] - row[x]) * (float)ew;
    float pm = (marg[y + 1] - marg[y]) * (float)eh;
    return pc * pm / (2.0f * PI_F * PI_F * st);
}
// number of entries <= u in a sorted array (partition_point(|c| c <= u))
AGD int count_le(const float* a, int n, float u) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a[mid] <= u)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}
AGD f3 env_map_sample(const DScene& S, float u1, float u2, float* pdf) {
    const int h = uload(&S.env_h), w = uload(&S.env_w);
    const float* marg = uload(&S.env_marg);
    int y = min(max(count_le(marg, h + 1, u1), 1), h) - 1;
    const float* cond = uload(&S.env_cond) + (size_t)y * (w + 1);
    int x = min(max(count_le(cond, w + 1, u2), 1), w) - 1;
    float fy = (u1 - marg[y]) / fmaxf(marg[y + 1] - marg[y], 1e-12f);
    float fx = (u2 - cond[x]) / fmaxf(cond[x + 1] - cond[x], 1e-12f);
    float u = ((float)x + fminf(fmaxf(fx, 0.0f), 0.999f)) / (float)w;
    float v = ((float)y + fminf(fmaxf(fy, 0.0f), 0.999f)) / (float)h;
    f3 d = env_dir(S, u, v);
    *pdf = env_map_pdf(S, d);
    return d;
}
// light.rs:110-127 (Env::eval)
AGD Spec env_eval(const DScene& S, f3 dir, const Spec& lam, const TabWin* W = nullptr) {
    if (uload(&S.env_kind) == AG_ENV_MAP) {
        float u, v;
        env_uv(S, dir, &u, &v);
        const int ew = uload(&S.env_w), eh = uload(&S.env_h);
        int x = min((int)(u * (float)ew), ew - 1);
        int y = min((int)(v * (float)eh), eh - 1);
        const float* w = uload(&S.env_weights) + ((size_t)y * ew + x) * 3;
        const float* basis = uload(&S.env_basis);
        Spec r;
        for (int i = 0; i < 4; i++) r.v[i] = linrgb_eval(basis, w, lam.v[i]);
        return r;
    }
    float y = fminf(fmaxf(dir.y, -1.0f), 1.0f);
    if (y >= 0.0f) {
        float t = powf(y, uload(&S.env_power));
        return lerp(t, tab_sample(tw(W, uload(&S.env_hor)), lam), tab_sample(tw(W, uload(&S.env_top)), lam));
    }
    float t = powf(-y, 0.5f);
    return lerp(t, tab_sample(tw(W, uload(&S.env_hor)), lam), tab_sample(tw(W, uload(&S.env_bot)), lam));
}

AGD float light_pick_pdf(const DScene& S, int panel) {
    const int nl = uload(&S.n_lights);
    for (int k = 0; k < nl; k++)
        if (uload(&S.lights[k]) == panel) return uload(&S.light_cdf[k]) - (k == 0 ? 0.0f : uload(&S.light_cdf[k - 1]));
    return 0.0f;
}

AGD DMat material_of(const DScene& S, const DMat* vmats, const DView& V, int m, const TabWin* W = nullptr) {
    const int m0 = uload(&V.mat0);
    DMat M = load_mat(m == 0 && m0 >= 0 ? &vmats[m0] : &S.mat[m]);
    M.tn = tw(W, M.tn);
    M.tk = tw(W, M.tk);
    M.tpatina = tw(W, M.tpatina);
    M.ta = tw(W, M.ta);
    M.tb = tw(W, M.tb);
    return M;
}

// material.rs:96-138
// KM: the BSDF kinds the caller can meet (a material whose lobes lie
// outside it comes out black).
template <uint32_t KM = KIND_ALL>
AGD Bsdf material_bsdf(const DMat& M, const Spec& lam, float mask, float u, f3 p_obj, f3 n_world, const Frame& fr,
                       const DView& V) {
    Bsdf b;
    b.alpha = M.alpha;
    b.eta1 = M.eta;
    b.w = 0.0f;
    b.kind = B_BLACK;
    switch (M.kind) {
        case AG_MAT_METAL: {
            if constexpr (!has_kind(KM, B_CONDUCTOR)) break;
            if (has_kind(KM, B_COATED) && M.has_patina) {
                float t = fminf(fmaxf((mask - M.p_lo) / fmaxf(M.p_hi - M.p_lo, 1e-4f), 0.0f), 1.0f);
                float w = M.p_strength * t * t * (3.0f - 2.0f * t);
                if (u < w) {
                    b.kind = B_COATED;
                    b.alpha = M.p_alpha;
                    b.eta1 = M.p_eta;
                    b.a = tab_sample(M.tpatina, lam);
                    return b;
                }
            }
            b.kind = B_CONDUCTOR;
            for (int i = 0; i < 4; i++) {
                b.a.v[i] = grid_eval(M.tn, lam.v[i]);
                b.b.v[i] = grid_eval(M.tk, lam.v[i]);
            }
            b.f0 = fresnel_conductor(1.0f, b.a, b.b);
            return b;
        }
        case AG_MAT_GEM:
            if constexpr (!has_kind(KM, B_DIELECTRIC)) break;
            b.kind = B_DIELECTRIC;
            b.a = tab_sample(M.ta, lam);
            return b;
        case AG_MAT_GEM_OPAQUE:
        case AG_MAT_PLASTIC:
            if constexpr (!has_kind(KM, B_COATED)) break;
            b.kind = B_COATED;
            b.a = tab_sample(M.ta, lam);
            return b;
        case AG_MAT_DIFFUSE:
            if constexpr (!has_kind(KM, B_DIFFUSE)) break;
            b.kind = B_DIFFUSE;
            b.a = tab_sample(M.ta, lam);
            return b;
        case AG_MAT_OPAL: {
            if constexpr (!has_kind(KM, B_OPAL)) break;
            f3 dir;
            float spacing, brightness;
            opal_domain(p_obj / M.opal_cell, &dir, &spacing, &brightness);
            f3 nw = normalize(n_world * 1.2f + rot_y(dir, uload(&V.ts), uload(&V.tc)));
            f3 n = to_local(fr, nw);
            if (n.z < 0.0f) n = -n;
            b.kind = B_OPAL;
            b.on = n;
            b.alpha = 0.22f * 0.22f;
            b.spacing = spacing;
            b.strength = M.opal_strength * brightness;
            b.b = lam;
            b.a = tab_sample(M.ta, lam);
            b.w = 0.6f;
            return b;
        }
        default: break;
    }
    b.kind = B_BLACK;
    return b;
}

// material.rs:157-177
AGD void guide_rgb(const DMat& M, float cavity, float out[3]) {
    const float L[3] = {610.0f, 550.0f, 465.0f};
    switch (M.kind) {
        case AG_MAT_METAL:
            for (int i = 0; i < 3; i++) {
                float n = grid_eval(M.tn, L[i]), k = grid_eval(M.tk, L[i]);
                out[i] = ((n - 1.0f) * (n - 1.0f) + k * k) / ((n + 1.0f) * (n + 1.0f) + k * k);
            }
            if (M.has_patina) {
                float t = fminf(fmaxf((cavity - M.p_lo) / fmaxf(M.p_hi - M.p_lo, 1e-4f), 0.0f), 1.0f);
                float w = M.p_strength * t * t * (3.0f - 2.0f * t);
                for (int i = 0; i < 3; i++) out[i] = out[i] * (1.0f - w) + tab_eval(M.tpatina, L[i]) * w;
            }
            return;
        case AG_MAT_GEM:
        case AG_MAT_GEM_OPAQUE: out[0] = out[1] = out[2] = 1.0f; return;
        case AG_MAT_OPAL:
        case AG_MAT_PLASTIC:
        case AG_MAT_DIFFUSE:
            for (int i = 0; i < 3; i++) out[i] = tab_eval(M.ta, L[i]);
            return;
        default: out[0] = out[1] = out[2] = 0.0f; return;
    }
}

// ------------------------------------------------------------------ wavefront kernels
__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

AGD int find_segment(const DSeg* segs, int nseg, uint32_t path) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[mid].begin <= path)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// render.rs:67-95 (per sample: pixel seed, jitter, lens, wavelengths, camera ray)
__global__ __launch_bounds__(256) void k_raygen(const DSeg* __restrict__ segs, int nseg, uint32_t npaths,
                                                const DView* __restrict__ views, DPaths P, DQueues Q) {
    __shared__ uint32_t s_sobol[SOBOL_TAB_WORDS];
    const AG_LDS uint32_t* sob = (AG_LDS uint32_t*)s_sobol;
    sobol_tables_fill((AG_LDS uint32_t*)s_sobol);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) Q.counters[CNT_RAYS] = npaths;
    for (uint32_t path = blockIdx.x * blockDim.x + threadIdx.x; path < npaths; path += gridDim.x * blockDim.x) {
        const DSeg& sg = segs[find_segment(segs, nseg, path)];
        uint32_t local = path - sg.begin;
        const uint32_t q = local % sg.npix, k = local / sg.npix;
        uint32_t pix = sg.p0 + q;
        uint32_t si = sg.s0 + k;
        // The camera ray's queue slot: the segment's pixel range cut into
        // Q.bands bands of rows, band-major, samples then pixels inside a band.
        // One band per XCD slice of the queue: each XCD's traversal then walks
        // the part of the tree one strip of the image sees, through its own L2.
        uint32_t slot = path;
        if (Q.bands > 1u) {
            const uint64_t nb = Q.bands, np = sg.npix;
            uint32_t bd = (uint32_t)min<uint64_t>(nb - 1, (uint64_t)q * nb / np);
            while (bd + 1 < nb && (uint64_t)(bd + 1) * np / nb <= q) bd++;
            while (bd > 0 && (uint64_t)bd * np / nb > q) bd--;
            const uint32_t lo = (uint32_t)((uint64_t)bd * np / nb), hi = (uint32_t)((uint64_t)(bd + 1) * np / nb);
            slot = sg.begin + sg.ns * lo + k * (hi - lo) + (q - lo);
        }
        const DView& V = views[sg.view];
        uint32_t x = pix % V.width, y = pix / V.width;
        uint32_t ps = hash32((y * V.width + x) ^ V.seed_h);
        float jx = tent(rng_dim(sob, ps, si, 0), 1.0f);
        float jy = tent(rng_dim(sob, ps, si, 1), 1.0f);
        float lu = rng_dim(sob, ps, si, 2), lv = rng_dim(sob, ps, si, 3);
        Spec lam, pdf;
        sample_wavelengths(rng_dim(sob, ps, si, 4), lam, pdf);
        float sx = (((float)x + 0.5f + jx) * V.aspect_sx) - 1.0f;
        float sy = 1.0f - (((float)y + 0.5f + jy) * V.aspect_sy);
        // camera.rs:57-70
        float cx = (sx + V.shift0) * V.half_w;
        float cy = (sy + V.shift1) * V.half_h;
        f3 dir = normalize(ld3(V.fwd) + ld3(V.right) * cx + ld3(V.up) * cy);
        f3 org = ld3(V.pos);
        if (V.lens_radius > 0.0f) {
            float ft = V.focus_dist / dot(dir, ld3(V.fwd));
            f3 pf = org + dir * ft;
            float dx, dy;
            disk(lu, lv, &dx, &dy);
            org = ld3(V.pos) + ld3(V.right) * (dx * V.lens_radius) + ld3(V.up) * (dy * V.lens_radius);
            dir = normalize(pf - org);
        }
        // throughput, flags, previous pdf, generator state and guides are
        // implicit at depth 0: the first shading sets them
        P.L[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.lam[path] = to4(lam);
        P.pdf[path] = to4(pdf);
        P.pix_seed[path] = ps;
        P.si[path] = si;
        Q.ray_o[0][slot] = make_float4(org.x, org.y, org.z, __uint_as_float(path));
        Q.ray_d[0][slot] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(ray_tag(sg.view, 0u)));
    }
}

// scene.rs:169-218 (closest hit over the piece, the panels and the floor):
// the piece in object space (ray turned by -turntable), then panels and floor
// in world space.  Finished rays are staged per wavefront in LDS and appended
// to the material buckets 64 or more at a time (one atomic per key per flush).
// Staging entries per wavefront (ray index, and its key beside it): flushed
// once FLUSH_AT are staged, with one atomic per key present, so a bucket
// counter sees one atomic per FLUSH_AT rays or so instead of one per 64.
constexpr uint32_t FLUSH_AT = 512;
constexpr int STAGE = FLUSH_AT + 64;

struct ExtendJob {
    const DView* views;
    DPaths P;
    DQueues Q;
    int depth;
    uint32_t* cnt;
    const float4* ro;
    const float4* rd;
    uint32_t* stage;  // this wavefront's STAGE ray indices
    uint8_t* skey;    // and their bucket keys
    uint32_t* kcnt;   // per bucket key: staged rays, then ranks (MAX_KEYS)
    uint32_t* kbase;  // per bucket key: the flush's first slot (MAX_KEYS)
    uint32_t staged;  // uniform
    // Panels and floor first (world space, strict <: panels in index order,
    // then the floor), then the piece with t_max just above their nearest
    // hit, so the piece wins ties as it does when tested first.  Until the
    // piece is hit, L.u / L.v hold the panel-or-floor code and distance.
    AGD bool begin(uint32_t i, TravLane& L) {
        float4 a = ro[i], b = rd[i];
        if (__float_as_uint(a.w) == DEAD_PATH) {  // a path that ended: nothing to trace or shade
            L.aux = NO_RAY;
            return false;
        }
        const uint32_t vi = tag_view(__float_as_uint(b.w));
        per_uniform(vi, [&](uint32_t v) {
            const DScene* Sp = uload(&views[v].scene);
            const float turn = uload(&views[v].turn_deg), tsn = uload(&views[v].tsn), tcn = uload(&views[v].tcn);
            f3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
            float tmax = INF_F;
            uint32_t code = HIT_MISS;
            const int np = uload(&Sp->n_panels);
            for (int k = 0; k < np; k++) {
                const DPanel* pk = &Sp->panel[k];
                if (depth == 0 && !uload(&pk->camera_visible)) continue;
                float t;
                if (panel_intersect(pk, o, d, tmax, &t)) {
                    tmax = t;
                    code = HIT_PANEL | (uint32_t)k;
                }
            }
            if (uload(&Sp->floor_present)) {
                const float fy = uload(&Sp->floor_y);
                if (fabsf(d.y) >= 1e-9f) {  // light.rs:133-162, as floor_intersect
                    float t = (fy - o.y) / d.y;
                    if (t > 1e-5f && t < tmax) {
                        tmax = t;
                        code = HIT_FLOOR;
                    }
                }
            }
            if (turn != 0.0f) {
                o = rot_y(o, tsn, tcn);
                d = rot_y(d, tsn, tcn);
            }
            // next float above tmax (> 0 here: panel t > 1e-6, floor t > 1e-5)
            float tpiece = tmax < INF_F ? __uint_as_float(__float_as_uint(tmax) + 1u) : INF_F;
            lane_begin(L, uload(&Sp->nodes), uload(&Sp->tris), o, d, tpiece);
            L.u = __uint_as_float(code);
            L.v = tmax;
        });
        L.aux = vi;
        return true;
    }
    AGD void finish(uint32_t i, const TravLane& L, bool mine) {
        int key = -1;
        if (mine && L.aux != NO_RAY) {
            per_uniform(L.aux, [&](uint32_t v) {
                const DScene* Sp = uload(&views[v].scene);
                float4 h;
                int m = -1;  // the surface's material
                bool floor = false;
                if (L.prim != NO_RAY) {
                    h = make_float4(L.t, L.u, L.v, __uint_as_float(L.prim));
                    m = (int)uload(&Sp->tmat)[L.prim];
                } else {
                    uint32_t code = __float_as_uint(L.u);
                    h = make_float4(L.v, 0.0f, 0.0f, L.u);
                    floor = code == HIT_FLOOR;
                    if (floor) m = uload(&Sp->floor_mat);
                    key = code == HIT_MISS ? KEY_MISS : KEY_PANEL;
                }
                if (m >= 0) {
                    key = 2 + m;
                    if (Q.sort == 4) {
                        // the shading family of the view's material m; floor
                        // hits go to the diffuse kernel or the general one
                        const int f = (int)((uload(&views[v].fam) >> (2 * m)) & 3u);
                        key = f == 0 && !floor ? KEY_METAL : f == 1 ? KEY_DIFFUSE : KEY_OTHER;
                    }
                }
                Q.hit[i] = h;
            });
        }
        if (Q.sort == 0 && key > 0) key = 0;  // unsorted shading: one bucket in arrival order
        if (Q.sort == 3 && key > 2) key = 2;  // misses, panel hits, then every surface hit in arrival order
        uint64_t m = __ballot(key >= 0);
        if (key >= 0) {
            const uint32_t e = staged + lane_prefix(m);
            stage[e] = i;
            skey[e] = (uint8_t)key;
        }
        staged += (uint32_t)__popcll(m);
        if (staged >= FLUSH_AT) flush();
    }
    AGD void flush() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (Q.sort == 2) {
            // Local sort: the staged rays (neighbouring pixels, one traversal
            // wavefront's) go to bucket 0 as one contiguous run ordered by key,
            // so a shading wavefront mostly sees one material while the queue
            // keeps its spatial order run by run.
            uint32_t base = 0;
            if (lane_id() == 0) base = atomicAdd(cnt + CNT_KEY, staged);
            base = __shfl(base, 0);
            uint32_t off = 0;
            for (int k = 0; k < Q.n_keys && off < staged; k++) {
                for (uint32_t c = 0; c < staged; c += 64) {
                    const uint32_t e = c + lane_id();
                    const bool mine = e < staged && (int)skey[e] == k;
                    const uint64_t m = __ballot(mine);
                    if (mine) Q.bucket[base + off + lane_prefix(m)] = stage[e];
                    off += (uint32_t)__popcll(m);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            staged = 0;
            return;
        }
        // per key: count the staged rays (LDS atomics), one global atomic for
        // all of them, then each ray's slot from a fresh per-key rank
        for (int k = lane_id(); k < Q.n_keys; k += 64) kcnt[k] = 0u;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t c = 0; c < staged; c += 64) {
            const uint32_t e = c + lane_id();
            if (e < staged) atomicAdd(&kcnt[skey[e]], 1u);
        }
        __builtin_amdgcn_wave_barrier();
        for (int k = lane_id(); k < Q.n_keys; k += 64) {
            const uint32_t n = kcnt[k];
            kbase[k] = n ? atomicAdd(cnt + CNT_KEY + k * CNT_PAD, n) : 0u;
            kcnt[k] = 0u;
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t c = 0; c < staged; c += 64) {
            const uint32_t e = c + lane_id();
            if (e < staged) {
                const uint32_t k = skey[e];
                Q.bucket[(size_t)k * Q.cap + kbase[k] + atomicAdd(&kcnt[k], 1u)] = stage[e];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        staged = 0;
    }
    AGD void drain() {
        if (staged) flush();
    }
};

__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_extend(const DView* __restrict__ views, DPaths P, DQueues Q, int depth,
                                                       uint32_t refill, uint32_t spread, uint2* spill) {
    TRAV_STACK_DECL
    __shared__ uint32_t s_stage[TRAV_BLOCK / 64][STAGE];
    __shared__ uint8_t s_skey[TRAV_BLOCK / 64][STAGE];
    __shared__ uint32_t s_kcnt[TRAV_BLOCK / 64][MAX_KEYS], s_kbase[TRAV_BLOCK / 64][MAX_KEYS];
    uint32_t* cnt = Q.counters + depth * CNT_STRIDE;
    ExtendJob J{views, P, Q, depth, cnt, Q.ray_o[depth & 1], Q.ray_d[depth & 1], s_stage[threadIdx.x / 64],
                s_skey[threadIdx.x / 64], s_kcnt[threadIdx.x / 64], s_kbase[threadIdx.x / 64], 0u};
    traverse_persistent<false>(J, min(cnt[CNT_RAYS], Q.qcap), Q.cursors + (size_t)depth * CURSORS_PER_DEPTH, refill, st, spread, Q.slices);
}

struct ShadeArgs {
    const DView* views;
    const DMat* vmats;
    uint32_t max_depth;   // the launch's deepest view (each path stops at its own view's)
    const float* energy;  // the GGX energy table (32 x 32), staged into LDS per block
    uint32_t tab_cap;     // floats of dynamic LDS for the first view's scene tables
    int32_t k0, k1;       // the bucket keys this launch shades
};

// Shading families (sort mode 4, keys KEY_MISS .. KEY_OTHER): each family's
// buckets are shaded by a kernel compiled for that family alone, with fewer
// registers than the kernel that handles every material.
enum ShadeFam { FAM_ALL = 0, FAM_SKY = 1, FAM_METAL = 2, FAM_DIFFUSE = 3 };
constexpr uint32_t fam_kinds(int fam) {
    return fam == FAM_ALL       ? KIND_ALL
           : fam == FAM_METAL   ? kind_bit(B_CONDUCTOR) | kind_bit(B_COATED)
           : fam == FAM_DIFFUSE ? kind_bit(B_DIFFUSE)
                                : 0u;
}

// integrator.rs:50-221, one bounce, over the hits in buckets [A.k0, A.k1).
// FAM_ALL shades every material and is compiled at two occupancies
// (k_shade<FAM_ALL, W>, W waves per SIMD): at 2 it keeps every value in
// registers; at 3 the compiler fits 168 registers and spills a few to
// scratch, trading those reloads for more waves to hide memory latency.  The
// family instances compile only their own hit codes and lobes.
template <int FAM>
AGD void shade_bounce(const ShadeArgs& A, const DPaths& P, const DQueues& Q, int depth) {
    constexpr uint32_t KM = fam_kinds(FAM);
    constexpr bool SKY = FAM == FAM_ALL || FAM == FAM_SKY;         // misses and panel hits
    constexpr bool FLOOR = FAM == FAM_ALL || FAM == FAM_DIFFUSE;   // floor hits
    constexpr bool SURF = FAM != FAM_SKY;                          // triangle hits
    constexpr bool EMITS = SKY || FLOOR;  // a miss, a panel or an absent floor adds emission
    constexpr bool ENERGY = has_kind(KM, B_CONDUCTOR);  // E(mu, alpha) of the conductor lobe
    uint32_t* cnt = Q.counters + depth * CNT_STRIDE;
    uint32_t* cnt_next = Q.counters + (depth + 1) * CNT_STRIDE;
    const int k0 = A.k0, k1 = A.k1;
    // in-order emission: the next queue is as long as this depth's hit list
    // (written by the launch over the first bucket)
    if (Q.ordered && k0 == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t tot = 0;
        for (int k = 0; k < Q.n_keys; k++) tot += cnt[CNT_KEY + k * CNT_PAD];
        cnt_next[CNT_RAYS] = (uint32_t)depth + 1 < A.max_depth ? tot : 0u;
    }
    // Wavefront w shades runs of RUN batches of 64 consecutive hits: batch t
    // is hits entry(t) .. entry(t) + 63, runs dealt to the wavefronts in turn
    // (RUN 1: a plain grid stride).  With chunked queues RUN is a chunk's
    // worth, so one chunk holds the continuations of consecutive hits and the
    // next queue stays spatially ordered chunk by chunk.
    uint32_t n_hits = 0;
    for (int k = k0; k < k1; k++) n_hits += cnt[CNT_KEY + k * CNT_PAD];
    const uint32_t chunk = n_hits >= SHADE_CHUNK_MIN ? SHADE_CHUNK : 0u;
    const uint32_t RUN = chunk ? chunk / 64u : 1u;
    const uint32_t n_wave = gridDim.x * (blockDim.x / 64u), wave_g = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u;
    auto entry = [&](uint32_t t) -> uint64_t {
        return ((uint64_t)(t / RUN) * n_wave + wave_g) * (uint64_t)(RUN * 64u) + (uint64_t)(t % RUN) * 64u;
    };
    // blocks past this launch's hit count have nothing to shade (uniform per
    // block, before any barrier)
    if ((uint64_t)blockIdx.x * (blockDim.x / 64u) * RUN * 64u >= n_hits) return;
    __shared__ uint32_t pre[MAX_KEYS + 1];
    __shared__ uint32_t s_first;  // hits in the buckets before k0 (in-order emission's slots)
    // microfacet.rs:93-140 E(mu, sqrt(alpha)): read by every rough lobe's
    // multiple-scattering term, several times per hit
    __shared__ float s_energy[ENERGY ? 32 * 32 : 1];
    if (ENERGY)
        for (int i = threadIdx.x; i < 32 * 32; i += blockDim.x) s_energy[i] = A.energy[i];
    __shared__ uint32_t s_sobol[SURF ? SOBOL_TAB_WORDS : 1];
    const AG_LDS uint32_t* sob = (AG_LDS uint32_t*)s_sobol;
    if (SURF) sobol_tables_fill((AG_LDS uint32_t*)s_sobol);
    // the first view's scene tables (materials, panel emission, sky spectra)
    // in LDS, as many as the launch gave room for: the n, k and spectra
    // lookups of every hit read them there
    extern __shared__ float s_tab[];
    TabWin W;
    {
        const DScene* S0 = uload(&A.views[0].scene);
        W.g = uload(&S0->win);
        W.n = min(uload(&S0->win_n), A.tab_cap);
        W.lds = s_tab;
        for (uint32_t q = threadIdx.x; q < W.n; q += blockDim.x) s_tab[q] = W.g[q];
    }
    if (threadIdx.x == 0) {
        uint32_t first = 0;
        for (int k = 0; k < k0; k++) first += cnt[CNT_KEY + k * CNT_PAD];
        s_first = first;
        pre[0] = 0;
        for (int k = k0; k < k1; k++) pre[k - k0 + 1] = pre[k - k0] + cnt[CNT_KEY + k * CNT_PAD];
    }
    __syncthreads();
    const int nk = k1 - k0;
    const uint32_t total = pre[nk];
    const uint32_t jfirst = s_first;
    const float4* ro = Q.ray_o[depth & 1];
    const float4* rd = Q.ray_d[depth & 1];
    float4* no = Q.ray_o[(depth + 1) & 1];
    float4* nd = Q.ray_d[(depth + 1) & 1];
    // Software pipeline over a wavefront's batches: each iteration issues the
    // next batch's ray and hit loads and the bucket entry of the one after,
    // so two of a hit's dependent memory round trips (bucket -> ray -> path
    // state) overlap the previous hit's shading.
    auto slot_of = [&](uint32_t jj) -> uint32_t {
        int key = 0;
        while (key + 1 < nk && pre[key + 1] <= jj) key++;
        return Q.bucket[(size_t)(k0 + key) * Q.cap + (jj - pre[key])];
    };
    const uint32_t lane = lane_id();
    const uint64_t j0 = entry(0) + lane, j1 = entry(1) + lane;
    float4 pa = make_float4(0.0f, 0.0f, 0.0f, 0.0f), pb = pa, ph = pa;
    if (j0 < total) {
        const uint32_t r0 = slot_of((uint32_t)j0);
        pa = ro[r0];
        pb = rd[r0];
        ph = Q.hit[r0];
    }
    uint32_t r_next = j1 < total ? slot_of((uint32_t)j1) : 0u;
    // the next depth's rays and this depth's shadow rays: appended at the
    // loop's converged point, in reserved chunks for large launches
    ChunkAppend rays_q{&cnt_next[CNT_RAYS], chunk, 0u, 0u, false};
    ChunkAppend shadow_q{&cnt[CNT_SHADOW], chunk, 0u, 0u, false};
    uint32_t n_live = 0, n_shadow = 0;  // wavefront-uniform counts of what was emitted
    for (uint32_t t = 0; entry(t) < total; t++) {
        const uint64_t jw = entry(t) + lane, jn = entry(t + 1) + lane, jnn = entry(t + 2) + lane;
        const uint32_t j = (uint32_t)min(jw, (uint64_t)total);  // >= total: no hit for this lane
        bool emit_ray = false;
        f3 nro, nrd;
        uint32_t path = 0, vidx = 0;
        bool sh_emit = false;  // this hit's next-event record, written after the loop body
        float4 sho, shd, shc, shx;
        const float4 a = pa, b = pb, hh = ph;
        if (jn < total) {
            pa = ro[r_next];
            pb = rd[r_next];
            ph = Q.hit[r_next];
        }
        if (jnn < total) r_next = slot_of((uint32_t)jnn);
        if (j < total) {
            path = __float_as_uint(a.w);
            f3 o{a.x, a.y, a.z}, d{b.x, b.y, b.z};
            float t = hh.x;
            uint32_t code = __float_as_uint(hh.w);
            // Radiance and the previous bounce's pdf are read only by the lanes
            // that may add emission (a miss, a light, the floor: every special
            // code has the top bit), up front with the rest of the state; the
            // hero pdfs only at a secondary-wavelength termination: a surface
            // hit moves ~70 bytes less than with the whole state.
            // Camera rays (depth 0) start from the state ray generation leaves
            // implicit: throughput 1, a specular previous bounce with pdf 1, no
            // medium, the generator seeded from the pixel and sample.
            const bool d0 = depth == 0;
            const bool lit = EMITS && (code & HIT_PANEL) != 0u;
            Spec beta = d0 ? splat(1.0f) : from4(P.beta[path]);
            const Spec lam = from4(P.lam[path]);
            Spec L = lit ? from4(P.L[path]) : splat(0.0f);
            const float prev_in = lit ? (d0 ? 1.0f : P.prev_pdf[path]) : 0.0f;
            float prev_pdf = 0.0f;
            uint32_t flags = d0 ? (1u | (255u << 8)) : P.flags[path];
            // the sampler's pixel seed, sample index and generator: only a
            // surface or floor hit draws from them (a miss or a light ends here)
            const bool draws = SURF && (!lit || code == HIT_FLOOR);
            const uint32_t ps = draws ? P.pix_seed[path] : 0u, si = draws ? P.si[path] : 0u;
            uint64_t rng = draws ? (d0 ? rng_init(ps, si) : P.rng[path]) : 0u;
            float4 g_alb = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g_nrm = g_alb;  // depth 0: the guides
            vidx = tag_view(__float_as_uint(b.w));
            const DView& V = A.views[vidx];  // fields read below with uload
            const DScene& S = *uload(&V.scene);
            const float turn = uload(&V.turn_deg), vts = uload(&V.ts), vtc = uload(&V.tc);
            const float env_pick = uload(&S.env_pick);
            const bool env_map = uload(&S.env_kind) == AG_ENV_MAP;
            bool specular = flags & 1u, rough = (flags & 2u) != 0;
            int medium = (int)((flags >> 8) & 255u);
            const bool clamp = rough;
            const float cm = uload(&V.clamp_indirect);
            bool alive = true;
            if (code != HIT_MISS && medium != 255) {
                Spec sig = tab_sample(tw(&W, uload(&S.mat[medium].tb)), lam);
                for (int i = 0; i < 4; i++) beta.v[i] = beta.v[i] * expf(-sig.v[i] * t);
            }
            auto env_weight = [&](f3 dir) -> float {
                if (env_map && !specular) {
                    const float pp = prev_in;
                    float pl = env_pick * env_map_pdf(S, dir);
                    return pp * pp / (pp * pp + pl * pl);
                }
                return 1.0f;
            };
            // emission seen by this bounce, added to the path's radiance
            auto emit = [&](const Spec& c) {
                add_clamped(L, c, clamp, cm);
                P.L[path] = to4(L);
            };
            f3 p, p_obj, ng, ns;
            int mat_id = 0;
            float cavity = 0.0f;
            if (SKY && code == HIT_MISS) {
                emit(beta * env_eval(S, d, lam, &W) * env_weight(d));
                alive = false;
            } else if (FLOOR && code == HIT_FLOOR) {
                p = o + d * t;
                if (next_random(rng) >= floor_presence(S, p)) {
                    emit(beta * env_eval(S, d, lam, &W) * env_weight(d));
                    alive = false;
                } else {
                    f3 n = d.y < 0.0f ? f3{0.0f, 1.0f, 0.0f} : f3{-0.0f, -1.0f, -0.0f};
                    p_obj = p;
                    ng = n;
                    ns = n;
                    mat_id = uload(&S.floor_mat);
                }
            } else if (SKY && (code & HIT_PANEL)) {
                const DPanel Pn = load_panel(&S.panel[code & 0xFFFFu], &W);
                Spec le = panel_le(Pn, o + d * t, d, lam);
                if (!is_black(le)) {
                    if (specular) {
                        emit(beta * le);
                    } else {
                        const float pp = prev_in;
                        float cos_l = fmaxf(fabsf(dot(-d, ld3(Pn.n))), 1e-8f);
                        float p_light =
                            (1.0f - env_pick) * light_pick_pdf(S, (int)(code & 0xFFFFu)) * t * t / (cos_l * Pn.area);
                        float w = pp * pp / (pp * pp + p_light * p_light);
                        emit(beta * le * w);
                    }
                }
                alive = false;
            } else if (SURF) {
                // scene.rs:173-201: attributes in object space, then to world
                uint32_t prim = code;
                const uint32_t* idx = uload(&S.idx);
                const float* pos = uload(&S.pos);
                const float* nrm = uload(&S.nrm);
                const float* cav = uload(&S.cavity);
                uint32_t i0 = idx[3 * prim], i1 = idx[3 * prim + 1], i2 = idx[3 * prim + 2];
                f3 pa = ld3(pos + 3 * i0), pb = ld3(pos + 3 * i1), pc = ld3(pos + 3 * i2);
                f3 ngo = normalize(cross(pb - pa, pc - pa));
                float hu = hh.y, hv = hh.z;
                float w = 1.0f - hu - hv;
                mat_id = uload(&S.tmat)[prim];
                const DMat Mh = material_of(S, A.vmats, V, mat_id, &W);
                f3 nso = ngo;
                if (Mh.kind != AG_MAT_GEM && nrm) {
                    f3 na = ld3(nrm + 3 * i0), nb = ld3(nrm + 3 * i1), nc = ld3(nrm + 3 * i2);
                    f3 nn = na * w + nb * hu + nc * hv;
                    nso = length_sq(nn) > 1e-12f ? normalize(nn) : ngo;
                }
                if (cav) cavity = cav[i0] * w + cav[i1] * hu + cav[i2] * hv;
                p_obj = pa * w + pb * hu + pc * hv;
                if (turn != 0.0f) {
                    p = rot_y(p_obj, vts, vtc);
                    ng = rot_y(ngo, vts, vtc);
                    ns = rot_y(nso, vts, vtc);
                } else {
                    p = p_obj;
                    ng = ngo;
                    ns = nso;
                }
            }
            if (SURF && alive) {
                const DMat M = material_of(S, A.vmats, V, mat_id, &W);
                f3 wo_w = -d;
                bool transmissive = M.kind == AG_MAT_GEM;
                f3 nsu = ns, ngu = ng;
                if (!transmissive && dot(ng, wo_w) < 0.0f) {
                    nsu = -ns;
                    ngu = -ng;
                }
                if (!transmissive && dot(nsu, wo_w) <= 0.0f) nsu = ngu;
                Frame fr = make_frame(nsu);
                f3 wo = to_local(fr, wo_w);
                const uint32_t dim = 8u + 8u * (uint32_t)depth;
                Bsdf bs = material_bsdf<KM>(M, lam, cavity, rng_dim(sob, ps, si, dim), p_obj, nsu, fr, V);
                if (d0) {
                    float g[3];
                    guide_rgb(M, cavity, g);
                    g_alb = make_float4(g[0], g[1], g[2], 0.0f);
                    g_nrm = make_float4(nsu.x, nsu.y, nsu.z, 0.0f);
                }
                // next-event estimation (integrator.rs:130-178)
                if (!bsdf_is_delta<KM>(bs)) {
                    float ul = rng_dim(sob, ps, si, dim + 1), u1 = rng_dim(sob, ps, si, dim + 2), u2 = rng_dim(sob, ps, si, dim + 3);
                    if (env_map && ul < env_pick) {
                        float pdf_dir;
                        f3 dir = env_map_sample(S, u1, u2, &pdf_dir);
                        if (pdf_dir > 0.0f && dot(dir, ngu) > 0.0f) {
                            f3 wi = to_local(fr, dir);
                            Spec f = bsdf_eval<KM>(bs, s_energy, wo, wi);
                            if (!is_black(f)) {
                                float p_light = env_pick * pdf_dir;
                                float p_bsdf = bsdf_pdf<KM>(bs, wo, wi);
                                float w = p_light * p_light / (p_light * p_light + p_bsdf * p_bsdf);
                                f3 org = offset_origin(p, ngu, dir);
                                Spec le = env_eval(S, dir, lam, &W);
                                Spec base = beta * f * le;
                                sh_emit = true;
                                sho = make_float4(org.x, org.y, org.z, __uint_as_float(path));
                                shd = make_float4(dir.x, dir.y, dir.z, __uint_as_float(0xFFFFFFFFu));
                                shc = to4(base);
                                shx = make_float4(fabsf(wi.z) * w, p_light, rough ? cm : -1.0f, __uint_as_float(vidx << 1 | 1u));
                            }
                        }
                    } else if (uload(&S.n_lights) > 0) {
                        const int nl = uload(&S.n_lights);
                        float uu = env_pick > 0.0f ? (ul - env_pick) / (1.0f - env_pick) : ul;
                        int k = 0;
                        while (k < nl && uload(&S.light_cdf[k]) < uu) k++;
                        k = min(k, nl - 1);
                        float pick = uload(&S.light_cdf[k]) - (k == 0 ? 0.0f : uload(&S.light_cdf[k - 1]));
                        int li = uload(&S.lights[k]);
                        pick = pick * (1.0f - env_pick);
                        const DPanel Pn = load_panel(&S.panel[li], &W);
                        f3 lp = ld3(Pn.c) + ld3(Pn.hu) * (2.0f * u1 - 1.0f) + ld3(Pn.hv) * (2.0f * u2 - 1.0f);
                        f3 to = lp - p;
                        float d2 = length_sq(to);
                        f3 dir = to / sqrtf(d2);
                        float cos_l = dot(-dir, ld3(Pn.n));
                        if (cos_l > 0.0f && dot(dir, ngu) > 0.0f) {
                            f3 wi = to_local(fr, dir);
                            Spec f = bsdf_eval<KM>(bs, s_energy, wo, wi);
                            if (!is_black(f)) {
                                float p_light = pick * d2 / (cos_l * Pn.area);
                                float p_bsdf = bsdf_pdf<KM>(bs, wo, wi);
                                float w = p_light * p_light / (p_light * p_light + p_bsdf * p_bsdf);
                                f3 org = offset_origin(p, ngu, dir);
                                Spec le = panel_le(Pn, lp, dir, lam);
                                Spec c = beta * f * le * (fabsf(wi.z) * w / p_light);
                                sh_emit = true;
                                sho = make_float4(org.x, org.y, org.z, __uint_as_float(path));
                                shd = make_float4(lp.x, lp.y, lp.z, __uint_as_float((uint32_t)li));
                                shc = to4(c);
                                shx = make_float4(0.0f, 0.0f, rough ? cm : -1.0f, __uint_as_float(vidx << 1));
                            }
                        }
                    }
                }
                // continue the path (integrator.rs:180-218)
                BSample sm;
                float u0 = rng_dim(sob, ps, si, dim + 4), u1 = rng_dim(sob, ps, si, dim + 5), u2 = rng_dim(sob, ps, si, dim + 6);
                bool ok = bsdf_sample<KM>(bs, s_energy, wo, u0, u1, u2, &sm);
                if (ok && (sm.pdf <= 0.0f || is_black(sm.f))) ok = false;
                f3 wi_w;
                if (ok) {
                    wi_w = to_world(fr, sm.wi);
                    float side = dot(wi_w, ngu) * dot(wo_w, ngu);
                    if ((!sm.transmit && side <= 0.0f) || (sm.transmit && side >= 0.0f)) ok = false;
                }
                if (ok) {
                    if (sm.term2) {  // kept whether or not the path goes on: the film weighs L by it
                        Spec pdf = from4(P.pdf[path]);
                        terminate_secondary(pdf);
                        P.pdf[path] = to4(pdf);
                    }
                    float sc = fabsf(sm.wi.z) / sm.pdf;
                    for (int i = 0; i < 4; i++) beta.v[i] = beta.v[i] * (sm.f.v[i] * sc);
                    if (!is_finite(beta)) ok = false;
                }
                if (ok) {
                    specular = sm.delta;
                    if (!sm.delta) rough = true;
                    prev_pdf = sm.pdf;
                    if (sm.transmit) medium = (wo.z > 0.0f && M.kind == AG_MAT_GEM) ? mat_id : 255;
                    nro = offset_origin(p, ngu, wi_w);
                    nrd = wi_w;
                    if ((uint32_t)depth >= uload(&V.rr_depth)) {
                        float survive = fminf(max_value(beta), 0.95f);
                        if (survive <= 0.0f || next_random(rng) >= survive)
                            ok = false;
                        else
                            beta = beta / survive;
                    }
                }
                if (ok && (uint32_t)depth + 1 < uload(&V.max_depth)) emit_ray = true;
            }
            // every camera ray's guides (zero where it saw no surface)
            if (d0) {
                P.galb[path] = g_alb;
                P.gnrm[path] = g_nrm;
            }
            // the state the next bounce reads (a path that ended is read
            // again only for its radiance, hero pdfs and guides)
            if (emit_ray) {
                P.beta[path] = to4(beta);
                P.prev_pdf[path] = prev_pdf;
                P.flags[path] = (specular ? 1u : 0u) | (rough ? 2u : 0u) | ((uint32_t)medium << 8);
                P.rng[path] = rng;
            }
        }
        {
            const uint32_t ss = shadow_q.append(sh_emit);
            if (sh_emit && ss < Q.qcap) {
                Q.sh_o[ss] = sho;
                Q.sh_d[ss] = shd;
                Q.sh_c[ss] = shc;
                Q.sh_x[ss] = shx;
            }
            n_shadow += (uint32_t)__popcll(__ballot(sh_emit));
        }
        if (Q.ordered) {
            // hit j's continuation goes to slot j: the next queue keeps this
            // depth's hit order whatever order the blocks run in
            if (j < total && (uint32_t)depth + 1 < A.max_depth) {
                no[jfirst + j] = make_float4(nro.x, nro.y, nro.z, __uint_as_float(emit_ray ? path : DEAD_PATH));
                if (emit_ray) nd[jfirst + j] = make_float4(nrd.x, nrd.y, nrd.z, __uint_as_float(ray_tag(vidx, (uint32_t)depth + 1)));
            }
        } else {
            const uint32_t slot = rays_q.append(emit_ray);
            if (emit_ray && slot < Q.qcap) {
                no[slot] = make_float4(nro.x, nro.y, nro.z, __uint_as_float(path));
                nd[slot] = make_float4(nrd.x, nrd.y, nrd.z, __uint_as_float(ray_tag(vidx, (uint32_t)depth + 1)));
            }
        }
        n_live += (uint32_t)__popcll(__ballot(emit_ray));
    }
    // dead tails of the reserved chunks; the emitted counts, one atomic each per wavefront
    rays_q.close(no, Q.qcap);
    shadow_q.close(Q.sh_o, Q.qcap);
    if (lane_id() == 0) {
        if (n_live) atomicAdd(&cnt_next[CNT_LIVE], n_live);
        if (n_shadow) atomicAdd(&cnt[CNT_SHLIVE], n_shadow);
    }
}

template <int FAM, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) void k_shade(ShadeArgs A, DPaths P,
                                                                                        DQueues Q, int depth) {
    shade_bounce<FAM>(A, P, Q, depth);
}

// scene.rs:120-140 (sky_visibility) and scene.rs:220-243 (visible): the
// piece first (any hit, object space), then the other panels and the floor.
struct ShadowJob {
    const DView* views;
    DPaths P;
    DQueues Q;
    // The other panels and the floor first (world space); a blocked ray
    // skips the piece.  For a sky ray that reaches the floor the
    // transmittance is 1 - presence (scene.rs:131-139).  L.u holds that
    // transmittance, L.aux the path.
    AGD bool begin(uint32_t i, TravLane& L) {
        float4 a = Q.sh_o[i], b = Q.sh_d[i];
        uint32_t path = __float_as_uint(a.w);
        if (path == DEAD_PATH) {  // an unused slot of a reserved chunk
            L.aux = path;
            L.prim = 1;
            return false;
        }
        f3 from{a.x, a.y, a.z};
        const uint32_t xw = __float_as_uint(Q.sh_x[i].w);
        const uint32_t kind = xw & 1u;
        bool blocked = false;
        per_uniform(xw >> 1, [&](uint32_t v) {
            const DScene* Sp = uload(&views[v].scene);
            const float turn = uload(&views[v].turn_deg), tsn = uload(&views[v].tsn), tcn = uload(&views[v].tcn);
            const int np = uload(&Sp->n_panels);
            const bool has_floor = uload(&Sp->floor_present) != 0;
            const float fy = uload(&Sp->floor_y);
            f3 dir;
            float tmax, vis = 1.0f;
            if (kind == 0u) {
                f3 dd = f3{b.x, b.y, b.z} - from;
                float dist = length(dd);
                dir = dd / dist;
                tmax = dist * (1.0f - 1e-4f);
                int ignore = (int)__float_as_uint(b.w);
                for (int k = 0; k < np && !blocked; k++) {
                    float t;
                    if (k != ignore && panel_intersect(&Sp->panel[k], from, dir, tmax, &t)) blocked = true;
                }
            } else {
                dir = f3{b.x, b.y, b.z};
                tmax = INF_F;
                for (int k = 0; k < np && !blocked; k++) {
                    float t;
                    if (panel_intersect(&Sp->panel[k], from, dir, INF_F, &t)) blocked = true;
                }
            }
            if (!blocked && has_floor && fabsf(dir.y) >= 1e-9f) {  // light.rs:133-162, as floor_intersect
                float t = (fy - from.y) / dir.y;
                if (t > 1e-5f && t < tmax) {
                    if (kind == 0u) {
                        blocked = true;
                    } else {  // floor_presence, light.rs:150-162
                        f3 p = from + dir * t;
                        float r = sqrtf(p.x * p.x + p.z * p.z), f0 = uload(&Sp->floor_f0), f1 = uload(&Sp->floor_f1);
                        float pres = 1.0f;
                        if (r >= f1) {
                            pres = 0.0f;
                        } else if (r > f0) {
                            float q = (r - f0) / (f1 - f0);
                            pres = 1.0f - q * q * (3.0f - 2.0f * q);
                        }
                        vis = 1.0f - pres;
                    }
                }
            }
            if (kind != 0u && !(vis > 0.0f)) blocked = true;
            f3 oo = from, od = dir;
            if (turn != 0.0f) {
                oo = rot_y(from, tsn, tcn);
                od = rot_y(dir, tsn, tcn);
            }
            lane_begin(L, uload(&Sp->nodes), uload(&Sp->tris), oo, od, tmax);
            L.u = vis;
        });
        L.aux = path;
        if (blocked) {
            L.prim = 1;
            return false;
        }
        return true;
    }
    AGD void finish(uint32_t i, const TravLane& L, bool mine) {
        if (!mine || L.prim != NO_RAY) return;  // blocked
        float4 x = Q.sh_x[i];
        Spec c = from4(Q.sh_c[i]);
        if (__float_as_uint(x.w) & 1u) c = c * (x.x * L.u / x.y);
        uint32_t path = L.aux;
        Spec Lr = from4(P.L[path]);
        add_clamped(Lr, c, x.z >= 0.0f, x.z);
        P.L[path] = to4(Lr);
    }
    AGD void drain() {}
};

__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_shadow(const DView* __restrict__ views, DPaths P, DQueues Q, int depth,
                                                       uint32_t refill, uint32_t spread, uint2* spill) {
    TRAV_STACK_DECL
    uint32_t* cnt = Q.counters + depth * CNT_STRIDE;
    ShadowJob J{views, P, Q};
    traverse_persistent<true>(J, min(cnt[CNT_SHADOW], Q.qcap), Q.cursors + (size_t)depth * CURSORS_PER_DEPTH + CURSOR_SET, refill, st,
                              spread, Q.slices);
}

// The shadow rays of one depth, then the extension rays of the next, in one
// persistent launch.  The two queues are independent: a shadow ray adds to its
// path's radiance (P.L) and reads only the shadow queue, an extension ray
// writes its hit record and material bucket.  Waves that run out of shadow
// rays go straight on to the extension queue, so the two launches' latency
// tails overlap; that is where a small pass spends most of its bounces.  The
// next shade, which adds emission to P.L, runs after this launch, so every
// path's radiance sums in the same order as with separate launches.
__global__ __launch_bounds__(TRAV_BLOCK) TRAV_WAVES void k_shadow_extend(const DView* __restrict__ views, DPaths P,
                                                              DQueues Q, int depth, uint32_t refill_shadow,
                                                              uint32_t refill_extend, uint32_t spread,
                                                              uint2* spill) {
    TRAV_STACK_DECL
    __shared__ uint32_t s_stage[TRAV_BLOCK / 64][STAGE];
    __shared__ uint8_t s_skey[TRAV_BLOCK / 64][STAGE];
    __shared__ uint32_t s_kcnt[TRAV_BLOCK / 64][MAX_KEYS], s_kbase[TRAV_BLOCK / 64][MAX_KEYS];
    {
        uint32_t* cnt = Q.counters + depth * CNT_STRIDE;
        ShadowJob J{views, P, Q};
        traverse_persistent<true>(J, min(cnt[CNT_SHADOW], Q.qcap), Q.cursors + (size_t)depth * CURSORS_PER_DEPTH + CURSOR_SET,
                                  refill_shadow, st, spread, Q.slices);
    }
    {
        const int nd = depth + 1;
        uint32_t* cnt = Q.counters + nd * CNT_STRIDE;
        ExtendJob J{views, P, Q, nd, cnt, Q.ray_o[nd & 1], Q.ray_d[nd & 1], s_stage[threadIdx.x / 64],
                    s_skey[threadIdx.x / 64], s_kcnt[threadIdx.x / 64], s_kbase[threadIdx.x / 64], 0u};
        traverse_persistent<false>(J, min(cnt[CNT_RAYS], Q.qcap), Q.cursors + (size_t)nd * CURSORS_PER_DEPTH, refill_extend, st, spread, Q.slices);
    }
}

// render.rs:96-111: per pixel, samples in order, into the pass accumulators
__global__ __launch_bounds__(256) void k_accumulate(const DSeg* __restrict__ segs, int nseg,
                                                    const DView* __restrict__ views, const float* __restrict__ cmf_g,
                                                    DPaths P) {
    const float s = 1.0f / (4.0f * 106.85692f);
    // the CIE 1931 table (471 x 3) in LDS: every path reads it for 4 wavelengths
    __shared__ float s_cmf[3 * AG_SPEC_N];
    for (int i = threadIdx.x; i < 3 * AG_SPEC_N; i += blockDim.x) s_cmf[i] = cmf_g[i];
    __syncthreads();
    const float* cmf = s_cmf;
    for (int g = 0; g < nseg; g++) {
        const DSeg sg = segs[g];
        const DView& V = views[sg.view];
        const size_t plane = (size_t)uload(&V.width) * uload(&V.height) * 3;
        float* const acc0 = uload(&V.acc);  // a global pointer (see uload)
        for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < sg.npix; q += gridDim.x * blockDim.x) {
            uint32_t pix = sg.p0 + q;
            float* ac = acc0 + (size_t)pix * 3;
            float* aa = ac + plane;
            float* an = ac + 2 * plane;
            float acc[3] = {ac[0], ac[1], ac[2]}, aacc[3] = {aa[0], aa[1], aa[2]}, nacc[3] = {an[0], an[1], an[2]};
            for (uint32_t k = 0; k < sg.ns; k++) {
                uint32_t path = sg.begin + k * sg.npix + q;
                Spec L = from4(P.L[path]), lam = from4(P.lam[path]), pdf = from4(P.pdf[path]);
                float xyz[3] = {0.0f, 0.0f, 0.0f};
                for (int i = 0; i < 4; i++) {
                    float p = pdf.v[i];
                    if (p == 0.0f || L.v[i] == 0.0f) continue;
                    float c[3];
                    xyz_bar(cmf, lam.v[i], c);
                    float w = L.v[i] / p;
                    xyz[0] += c[0] * w;
                    xyz[1] += c[1] * w;
                    xyz[2] += c[2] * w;
                }
                xyz[0] *= s;
                xyz[1] *= s;
                xyz[2] *= s;
                const float* m = V.cam_m;
                float rgb[3] = {m[0] * xyz[0] + m[1] * xyz[1] + m[2] * xyz[2], m[3] * xyz[0] + m[4] * xyz[1] + m[5] * xyz[2],
                                m[6] * xyz[0] + m[7] * xyz[1] + m[8] * xyz[2]};
                if (isfinite(rgb[0]) && isfinite(rgb[1]) && isfinite(rgb[2])) {
                    acc[0] += rgb[0];
                    acc[1] += rgb[1];
                    acc[2] += rgb[2];
                }
                float4 ga = P.galb[path], gn = P.gnrm[path];
                aacc[0] += ga.x; aacc[1] += ga.y; aacc[2] += ga.z;
                nacc[0] += gn.x; nacc[1] += gn.y; nacc[2] += gn.z;
            }
            ac[0] = acc[0]; ac[1] = acc[1]; ac[2] = acc[2];
            aa[0] = aacc[0]; aa[1] = aacc[1]; aa[2] = aacc[2];
            an[0] = nacc[0]; an[1] = nacc[1]; an[2] = nacc[2];
        }
    }
}

// ray tallies of one wave into 64-bit totals: [extension rays, shadow rays]
// totals: [0] extension rays, [1] shadow rays, [2 + d] extension rays at
// depth d, [2 + TALLY_DEPTHS + d] shadow rays at depth d (d < TALLY_DEPTHS)
__global__ void k_tally(const uint32_t* __restrict__ counters, uint32_t depth_slots, unsigned long long* totals) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        unsigned long long e = 0, s = 0;
        for (uint32_t d = 0; d < depth_slots; d++) {
            // live rays: the queue length at depth 0 (camera rays), the
            // continuation count after that (an in-order queue has dead slots)
            uint32_t ed = d == 0 ? counters[CNT_RAYS] : counters[d * CNT_STRIDE + CNT_LIVE];
            uint32_t sd = counters[d * CNT_STRIDE + CNT_SHLIVE];
            e += ed;
            s += sd;
            if (d < (uint32_t)TALLY_DEPTHS) {
                totals[2 + d] += ed;
                totals[2 + TALLY_DEPTHS + d] += sd;
            }
        }
        totals[0] += e;
        totals[1] += s;
    }
}

// film += pass accumulators; accumulators back to zero
__global__ void k_film_commit(float* __restrict__ film, float* __restrict__ acc, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        film[i] += acc[i];
        acc[i] = 0.0f;
    }
}

// Per-wavelength BSDF evaluation on given inputs (material.rs:96-138 +
// bsdf.rs:53-239) in the local frame with +Z the shading normal; the hit
// context is p_obj = (0.1, 0.2, 0.3), n_world = +Z.  out: 18 floats per probe.
__global__ void k_bsdf_probe(const DScene* __restrict__ Sp, ProbeIn in, float* out, uint32_t n) {
    const DScene& S = *Sp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const DMat& M = S.mat[in.mat[i]];
        Spec lam{{in.lam[4 * i], in.lam[4 * i + 1], in.lam[4 * i + 2], in.lam[4 * i + 3]}};
        DView V{};
        V.ts = 0.0f;
        V.tc = 1.0f;
        f3 nz{0.0f, 0.0f, 1.0f};
        Frame fr = make_frame(nz);
        const float* u = in.u + 4 * i;
        Bsdf b = material_bsdf(M, lam, in.mask[i], u[0], f3{0.1f, 0.2f, 0.3f}, nz, fr, V);
        f3 wo = ld3(in.wo + 3 * i), wi = ld3(in.wi + 3 * i);
        Spec e = bsdf_eval(b, S.energy, wo, wi);
        float* o = out + 18 * (size_t)i;
        for (int k = 0; k < 4; k++) o[k] = e.v[k];
        o[4] = bsdf_pdf(b, wo, wi);
        BSample sm{};
        bool ok = bsdf_sample(b, S.energy, wo, u[1], u[2], u[3], &sm);
        o[5] = ok ? 1.0f : 0.0f;
        o[6] = ok ? sm.wi.x : 0.0f;
        o[7] = ok ? sm.wi.y : 0.0f;
        o[8] = ok ? sm.wi.z : 0.0f;
        for (int k = 0; k < 4; k++) o[9 + k] = ok ? sm.f.v[k] : 0.0f;
        o[13] = ok ? sm.pdf : 0.0f;
        o[14] = ok && sm.delta ? 1.0f : 0.0f;
        o[15] = ok && sm.transmit ? 1.0f : 0.0f;
        o[16] = ok && sm.term2 ? 1.0f : 0.0f;
        o[17] = bsdf_is_delta(b) ? 1.0f : 0.0f;
    }
}

// ------------------------------------------------------------------ launchers
void launch_zero_u32(hipStream_t s, uint32_t* p, uint32_t n) {
    hipLaunchKernelGGL(k_zero_u32, dim3(std::max(1u, std::min((n + 255) / 256, 1024u))), dim3(256), 0, s, p, n);
}

void launch_rays_intersect(hipStream_t s, int grid, const float4* nodes, const float4* tris, RaysSoA r, HitsSoA h,
                           uint32_t n, uint32_t* cursor, uint32_t refill, uint32_t spread, uint2* spill) {
    hipLaunchKernelGGL(k_rays_intersect, dim3(grid), dim3(TRAV_BLOCK), 0, s, nodes, tris, r, h, n, cursor, refill, spread,
                       spill);
}
void launch_rays_occluded(hipStream_t s, int grid, const float4* nodes, const float4* tris, RaysSoA r, uint8_t* out,
                          uint32_t n, uint32_t* cursor, uint32_t refill, uint32_t spread, uint2* spill) {
    hipLaunchKernelGGL(k_rays_occluded, dim3(grid), dim3(TRAV_BLOCK), 0, s, nodes, tris, r, out, n, cursor, refill,
                       spread, spill);
}
int traversal_blocks_per_cu(bool batch) {
    int lo = 1 << 30;
    const void* render[3] = {(const void*)k_extend, (const void*)k_shadow, (const void*)k_shadow_extend};
    const void* rays[3] = {(const void*)k_rays_intersect, (const void*)k_rays_occluded, (const void*)k_ao};
    const void* const* ks = batch ? rays : render;
    for (int i = 0; i < 3; i++) {
        const void* k = ks[i];
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k, TRAV_BLOCK, 0) != hipSuccess || b < 1) b = 1;
        lo = std::min(lo, b);
    }
    return lo;
}
void launch_ao(hipStream_t s, int grid, const float4* nodes, const float4* tris, const float* pos, const float* nrm,
               uint32_t samples, float dist, float eps, uint32_t* counts, uint32_t n, uint32_t* cursor,
               uint32_t refill, uint32_t spread, uint2* spill) {
    AoJob J{nodes, tris, pos, nrm, samples, dist, eps, counts};
    hipLaunchKernelGGL(k_ao, dim3(grid), dim3(TRAV_BLOCK), 0, s, J, n, cursor, refill, spread, spill);
}
void launch_raygen(hipStream_t s, int grid, const DSeg* segs, int nseg, uint32_t npaths, const DView* views, DPaths P,
                   DQueues Q) {
    hipLaunchKernelGGL(k_raygen, dim3(grid), dim3(256), 0, s, segs, nseg, npaths, views, P, Q);
}
void launch_extend(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth, uint32_t refill,
                   uint32_t spread, uint2* spill) {
    hipLaunchKernelGGL(k_extend, dim3(grid), dim3(TRAV_BLOCK), 0, s, views, P, Q, depth, refill, spread, spill);
}
void launch_shade(hipStream_t s, int grid, const DView* views, const DMat* vmats, uint32_t max_depth,
                  const float* energy, DPaths P, DQueues Q, int depth, int waves, uint32_t tab_floats) {
    const uint32_t cap = std::min(tab_floats, SHADE_TAB_FLOATS);
    const size_t lds = cap * sizeof(float);
    ShadeArgs A{views, vmats, max_depth, energy, cap, 0, Q.n_keys};
    if (Q.sort == 4) {
        // one launch per family, in key order (misses and panel hits, metal,
        // diffuse, the rest); the metal kernel at the shading occupancy knob,
        // the others at the occupancy their registers allow without spills
        auto keys = [&](int k0, int k1) {
            A.k0 = k0;
            A.k1 = k1;
        };
        keys(KEY_MISS, KEY_PANEL + 1);
        hipLaunchKernelGGL((k_shade<FAM_SKY, 6>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        keys(KEY_METAL, KEY_METAL + 1);
        if (waves >= 3) hipLaunchKernelGGL((k_shade<FAM_METAL, 3>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        else hipLaunchKernelGGL((k_shade<FAM_METAL, 2>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        keys(KEY_DIFFUSE, KEY_DIFFUSE + 1);
        hipLaunchKernelGGL((k_shade<FAM_DIFFUSE, 3>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        keys(KEY_OTHER, Q.n_keys);
        if (waves >= 3) hipLaunchKernelGGL((k_shade<FAM_ALL, 3>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        else hipLaunchKernelGGL((k_shade<FAM_ALL, 2>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
        return;
    }
    if (waves >= 3) hipLaunchKernelGGL((k_shade<FAM_ALL, 3>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
    else hipLaunchKernelGGL((k_shade<FAM_ALL, 2>), dim3(grid), dim3(256), lds, s, A, P, Q, depth);
}
void launch_shadow(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth, uint32_t refill,
                   uint32_t spread, uint2* spill) {
    hipLaunchKernelGGL(k_shadow, dim3(grid), dim3(TRAV_BLOCK), 0, s, views, P, Q, depth, refill, spread, spill);
}
void launch_shadow_extend(hipStream_t s, int grid, const DView* views, DPaths P, DQueues Q, int depth,
                          uint32_t refill_shadow, uint32_t refill_extend, uint32_t spread, uint2* spill) {
    hipLaunchKernelGGL(k_shadow_extend, dim3(grid), dim3(TRAV_BLOCK), 0, s, views, P, Q, depth, refill_shadow,
                       refill_extend, spread, spill);
}
void launch_accumulate(hipStream_t s, int grid, const DSeg* segs, int nseg, const DView* views, const float* cmf,
                       DPaths P) {
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(256), 0, s, segs, nseg, views, cmf, P);
}
void launch_bsdf_probe(hipStream_t s, const DScene* S, ProbeIn in, float* out, uint32_t n) {
    hipLaunchKernelGGL(k_bsdf_probe, dim3(std::max(1u, std::min((n + 255) / 256, 1024u))), dim3(256), 0, s, S, in, out, n);
}
void launch_tally(hipStream_t s, const uint32_t* counters, uint32_t depth_slots, unsigned long long* totals) {
    hipLaunchKernelGGL(k_tally, dim3(1), dim3(64), 0, s, counters, depth_slots, totals);
}
void launch_film_commit(hipStream_t s, float* film, float* acc, size_t n) {
    unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_film_commit, dim3(std::max(1u, g)), dim3(256), 0, s, film, acc, n);
}

}  // namespace ag
