// Shared by the mesh kernels (mesh.hip, normals.hip): device buffers, the
// context workspace arena, the device mesh, the exclusive scan and the
// lock-free union-find.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "host_internal.hpp"

namespace ag {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int BLK = 256;
inline unsigned grid_of(size_t n) { return (unsigned)std::max<size_t>(1, (n + BLK - 1) / BLK); }

#define MCHK(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return fail(AG_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)


// Union-find over the classes.  parent[x] <= x always holds (a node only
// ever moves under a smaller one), and only roots are swapped, by atomicCAS.
// Loads are plain (workgroup scope, may hit a stale L1 line): a stale parent
// is still an ancestor, and a stale "is a root" answer only makes the CAS
// fail and the loop retry with the fresh value the CAS returns.
__device__ inline uint32_t uf_load(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Root of x with path halving.
__device__ inline uint32_t uf_find(uint32_t* parent, uint32_t x) {
    for (;;) {
        const uint32_t p = uf_load(&parent[x]);
        if (p == x) return x;
        const uint32_t gp = uf_load(&parent[p]);
        if (gp == p) return p;
        __hip_atomic_store(&parent[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

// Unites the sets of a and b; returns the root of the union as seen here.
__device__ inline uint32_t uf_unite(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return a;
        if (a > b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        const uint32_t old = atomicCAS(&parent[b], b, a);  // hook the larger root under the smaller
        if (old == b) return a;
        b = old;
    }
}

// Device buffer owned by a mesh.
struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    ~Buf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        release();
        if (n == 0) n = 4;
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

// 256-byte-aligned pieces of the context's mesh workspace (grow-only, kept
// by the context between calls, so a load allocates only its outputs).
struct Arena {
    char* base = nullptr;
    size_t used = 0;
    size_t add(size_t bytes) {
        const size_t o = used;
        used += (bytes + 255) & ~(size_t)255;
        return o;
    }
    template <class T>
    T* at(size_t off) const {
        return reinterpret_cast<T*>(base + off);
    }
    int bind(ag_ctx* ctx) {
        void* p = nullptr;
        const int rc = ctx_workspace(ctx, used, &p);
        base = static_cast<char*>(p);
        return rc;
    }
};


// Exclusive scan of n words (in != out); *total (device) gets the sum.
void device_scan_u32(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* work, uint32_t* total);
size_t device_scan_work_words(size_t n);

}  // namespace ag

struct ag_mesh {
    ag_ctx* ctx = nullptr;
    ag::Buf pos, tris;  // capacity >= nv * 3 floats, nt * 3 words
    ag::Buf nrm;        // nv * 3 floats once crease normals are computed
    uint32_t nv = 0, nt = 0;
    bool has_nrm = false;
};


namespace ag {
// Drops triangles with a repeated index, in order (weld.rs:104-112); returns
// the number dropped through *dropped.  Synchronizes.
int mesh_drop_collapsed(ag_mesh* m, uint32_t* dropped);
// Rejects out-of-range triangle indices; lo / hi = bounds of all vertices.
int mesh_check_bounds(ag_mesh* m, float lo[3], float hi[3]);
}  // namespace ag
