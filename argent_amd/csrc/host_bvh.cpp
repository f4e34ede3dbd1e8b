// Host BVH build for the device traversal kernels: binned SAH over all three
// axes (32 bins, parallel above 16K primitives, a node-traversal cost equal to
// one triangle test so leaves keep up to four triangles), collapsed to 8-wide
// nodes by repeatedly opening the child with the largest surface area (the
// policy of the reference's 4-wide collapse, bvh.rs:303-378), with child
// boxes quantized to 8 bits per plane against the node's box (conservative),
// laid out for 16-byte vector loads.  Topology differs from the reference's
// BVH; closest-hit results do not depend on it, except for a rare grazing ray
// whose Moller-Trumbore hit lies outside its triangle by more than the
// traversal's box widening.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <limits>

#include "host_internal.hpp"

namespace ag {
namespace {

struct Box {
    float lo[3], hi[3];
    Box() {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::numeric_limits<float>::infinity();
            hi[a] = -std::numeric_limits<float>::infinity();
        }
    }
    void grow(const float* p) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], p[a]);
            hi[a] = std::fmax(hi[a], p[a]);
        }
    }
    void grow(const Box& b) {
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], b.lo[a]);
            hi[a] = std::fmax(hi[a], b.hi[a]);
        }
    }
    float area() const {
        float e[3];
        for (int a = 0; a < 3; a++) e[a] = std::fmax(hi[a] - lo[a], 0.0f);
        return 2.0f * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
    }
};

struct Prim {
    Box b;
    float c[3];
    uint32_t id;
};

struct BNode {  // binary node
    Box b;
    int32_t left = -1, right = -1;  // children (inner)
    uint32_t first = 0, count = 0;  // leaf range in prim order
};

constexpr int BINS = 32;

struct Builder {
    std::vector<Prim>& prims;
    std::vector<BNode> nodes;
    std::mutex mu;

    explicit Builder(std::vector<Prim>& p) : prims(p) {}

    int alloc() {
        std::lock_guard<std::mutex> g(mu);
        nodes.emplace_back();
        return (int)nodes.size() - 1;
    }

    void build(int ni, uint32_t begin, uint32_t end, int depth) {
        Box bounds, cb;
        for (uint32_t i = begin; i < end; i++) {
            bounds.grow(prims[i].b);
            cb.grow(prims[i].c);
        }
        uint32_t n = end - begin;
        auto make_leaf = [&]() {
            std::lock_guard<std::mutex> g(mu);
            nodes[ni].b = bounds;
            nodes[ni].first = begin;
            nodes[ni].count = n;
        };
        if (n <= 1) {
            make_leaf();
            return;
        }
        // binned SAH over all axes
        float best_cost = std::numeric_limits<float>::infinity();
        int best_axis = -1, best_split = 0;
        for (int a = 0; a < 3; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (!(ext > 1e-12f)) continue;
            float k = (float)BINS / ext;
            Box bb[BINS];
            uint32_t bn[BINS] = {0};
            for (uint32_t i = begin; i < end; i++) {
                int b = std::min((int)((prims[i].c[a] - cb.lo[a]) * k), BINS - 1);
                bn[b]++;
                bb[b].grow(prims[i].b);
            }
            float la[BINS];
            uint32_t ln[BINS];
            Box acc;
            uint32_t cnt = 0;
            for (int i = 0; i < BINS - 1; i++) {
                acc.grow(bb[i]);
                cnt += bn[i];
                la[i] = cnt ? acc.area() : 0.0f;
                ln[i] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (int i = BINS - 1; i >= 1; i--) {
                acc.grow(bb[i]);
                cnt += bn[i];
                if (ln[i - 1] == 0 || cnt == 0) continue;
                float cost = la[i - 1] * (float)ln[i - 1] + acc.area() * (float)cnt;
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = a;
                    best_split = i;
                }
            }
        }
        // SAH with a traversal cost of one triangle test: C = A + A_l N_l + A_r N_r
        best_cost += bounds.area();
        float leaf_cost = bounds.area() * (float)n;
        if (n <= (uint32_t)BVH_MAX_LEAF && (best_axis < 0 || best_cost >= leaf_cost)) {
            make_leaf();
            return;
        }
        uint32_t mid;
        if (best_axis < 0) {
            mid = begin + n / 2;  // coincident centroids
        } else {
            int a = best_axis;
            float k = (float)BINS / (cb.hi[a] - cb.lo[a]);
            float lo = cb.lo[a];
            auto it = std::partition(prims.begin() + begin, prims.begin() + end, [&](const Prim& p) {
                return std::min((int)((p.c[a] - lo) * k), BINS - 1) < best_split;
            });
            mid = (uint32_t)(it - prims.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        int l = alloc(), r = alloc();
        {
            std::lock_guard<std::mutex> g(mu);
            nodes[ni].b = bounds;
            nodes[ni].left = l;
            nodes[ni].right = r;
        }
        if (n > 16384 && depth < 12) {
            auto fut = std::async(std::launch::async, [&, l, begin, mid, depth] { build(l, begin, mid, depth + 1); });
            build(r, mid, end, depth + 1);
            fut.get();
        } else {
            build(l, begin, mid, depth + 1);
            build(r, mid, end, depth + 1);
        }
    }
};

}  // namespace

void quantize_axis(float p, float s, float lo, float hi, uint8_t* qlo, uint8_t* qhi) {
    float fl = std::floor((lo - p) / s), fh = std::ceil((hi - p) / s);
    int l = (int)std::clamp(fl, 0.0f, 255.0f), h = (int)std::clamp(fh, 0.0f, 255.0f);
    while (l > 0 && std::fma((float)l, s, p) > lo) l--;
    while (h < 255 && std::fma((float)h, s, p) < hi) h++;
    *qlo = (uint8_t)l;
    *qhi = (uint8_t)h;
}

// Exponent byte of the smallest power-of-two scale s with p + 255 s >= hi in
// float fma arithmetic.
static uint8_t axis_exponent(float p, float hi) {
    float ext = hi - p;
    int e = 1;
    if (ext > 0.0f) {
        int ex;
        std::frexp(ext / 255.0f, &ex);  // ext / 255 < 2^ex
        e = std::clamp(ex + 127, 1, 254);
    }
    for (;; e++) {
        float s;
        uint32_t bits = (uint32_t)e << 23;
        std::memcpy(&s, &bits, 4);
        if (std::fma(255.0f, s, p) >= hi || e >= 254) return (uint8_t)e;
    }
}

void build_wide_bvh(const float* pos, const uint32_t* tri, uint32_t n_tris, WideBvh* out) {
    out->nodes.clear();
    out->tris.clear();
    out->n_nodes = out->n_leaves = out->depth = out->max_stack = 0;
    if (n_tris == 0) {
        // one node without children: every query misses
        out->nodes.assign(BVH_NODE_FLOATS, 0.0f);
        uint32_t w = 127u | 127u << 8 | 127u << 16;
        std::memcpy(&out->nodes[3], &w, 4);
        out->n_nodes = 1;
        out->depth = 1;
        return;
    }
    std::vector<Prim> prims(n_tris);
    for (uint32_t i = 0; i < n_tris; i++) {
        Prim& p = prims[i];
        for (int k = 0; k < 3; k++) p.b.grow(pos + 3 * (size_t)tri[3 * (size_t)i + k]);
        for (int a = 0; a < 3; a++) p.c[a] = (p.b.lo[a] + p.b.hi[a]) * 0.5f;
        p.id = i;
    }
    Builder B(prims);
    B.nodes.reserve(2 * (size_t)n_tris);
    B.alloc();
    B.build(0, 0, n_tris, 0);
    const std::vector<BNode>& bn = B.nodes;
    for (int a = 0; a < 3; a++) {
        out->lo[a] = bn[0].b.lo[a];
        out->hi[a] = bn[0].b.hi[a];
    }
    auto area = [](const BNode& n) {
        float e[3];
        for (int a = 0; a < 3; a++) e[a] = n.b.hi[a] - n.b.lo[a];
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    };
    auto put_tri = [&](uint32_t id) {
        const float* p0 = pos + 3 * (size_t)tri[3 * (size_t)id + 0];
        const float* p1 = pos + 3 * (size_t)tri[3 * (size_t)id + 1];
        const float* p2 = pos + 3 * (size_t)tri[3 * (size_t)id + 2];
        float e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        float e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
        float t[12] = {p0[0], p0[1], p0[2], e1[0], e1[1], e1[2], e2[0], e2[1], e2[2], 0.0f, 0.0f, 0.0f};
        std::memcpy(&t[9], &id, 4);
        out->tris.insert(out->tris.end(), t, t + 12);
    };
    std::vector<float>& W = out->nodes;
    out->tris.reserve((size_t)n_tris * 12);
    uint32_t max_depth = 0, leaves = 0, max_stack = 0;
    // fill node `me` from binary subtree `root`; inner children get consecutive
    // slots allocated here, then are filled depth first
    std::function<void(uint32_t, int, uint32_t, uint32_t)> fill = [&](uint32_t me, int root, uint32_t depth,
                                                                      uint32_t stack) {
        max_depth = std::max(max_depth, depth + 1);
        std::vector<int> kids;
        if (bn[root].left < 0)
            kids.push_back(root);
        else {
            kids.push_back(bn[root].left);
            kids.push_back(bn[root].right);
        }
        while (kids.size() < (size_t)BVH_WIDTH) {
            int best = -1;
            float ba = -1.0f;
            for (size_t i = 0; i < kids.size(); i++) {
                const BNode& n = bn[kids[i]];
                if (n.left >= 0) {
                    float a = area(n);
                    if (best < 0 || a > ba) {
                        best = (int)i;
                        ba = a;
                    }
                }
            }
            if (best < 0) break;
            int k = kids[best];
            kids[best] = kids.back();
            kids.pop_back();
            kids.push_back(bn[k].left);
            kids.push_back(bn[k].right);
        }
        Box nb;
        for (int k : kids) nb.grow(bn[k].b);
        float* nd = &W[(size_t)me * BVH_NODE_FLOATS];
        uint8_t e[3];
        for (int a = 0; a < 3; a++) {
            nd[a] = nb.lo[a];
            e[a] = axis_exponent(nb.lo[a], nb.hi[a]);
        }
        uint32_t child_base = (uint32_t)(W.size() / BVH_NODE_FLOATS);
        uint32_t tri_base = (uint32_t)(out->tris.size() / 12);
        uint32_t imask = 0, n_inner = 0, toff = 0;
        uint8_t meta[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint8_t q[6][8];
        std::memset(q, 0, sizeof(q));
        for (size_t c = 0; c < kids.size(); c++) {
            const BNode& n = bn[kids[c]];
            for (int a = 0; a < 3; a++) {
                float s;
                uint32_t bits = (uint32_t)e[a] << 23;
                std::memcpy(&s, &bits, 4);
                quantize_axis(nb.lo[a], s, n.b.lo[a], n.b.hi[a], &q[a][c], &q[3 + a][c]);
            }
            if (n.left >= 0) {
                imask |= 1u << c;
                meta[c] = (uint8_t)n_inner++;
            } else {
                meta[c] = (uint8_t)(n.count << 5 | toff);
                for (uint32_t i = 0; i < n.count; i++) put_tri(prims[n.first + i].id);
                toff += n.count;
                leaves++;
            }
        }
        uint32_t w = (uint32_t)e[0] | (uint32_t)e[1] << 8 | (uint32_t)e[2] << 16 | imask << 24;
        std::memcpy(&nd[3], &w, 4);
        uint32_t hdr[4] = {child_base, tri_base, 0, 0};
        std::memcpy(&hdr[2], meta, 8);
        std::memcpy(&nd[4], hdr, 16);
        std::memcpy(&nd[8], q, 48);
        // a traversal holds at most `stack` entries here plus all other children
        max_stack = std::max(max_stack, stack + (uint32_t)kids.size() - 1);
        W.resize(W.size() + (size_t)n_inner * BVH_NODE_FLOATS, 0.0f);
        uint32_t r = 0;
        for (size_t c = 0; c < kids.size(); c++)
            if (bn[kids[c]].left >= 0) fill(child_base + r++, kids[c], depth + 1, stack + (uint32_t)kids.size() - 1);
    };
    W.assign(BVH_NODE_FLOATS, 0.0f);
    fill(0, 0, 0, 0);
    out->n_nodes = (uint32_t)(W.size() / BVH_NODE_FLOATS);
    out->n_leaves = leaves;
    out->depth = max_depth;
    out->max_stack = max_stack;
}

}  // namespace ag
