// TEST INFRASTRUCTURE ONLY — the oracle's mesh ingest: a CPU restatement of
// argent-mesh's binary STL reader (io/stl.rs:34-62), the invalid-triangle
// filter (io.rs:206-224), vertex welding (weld.rs:38-330) and compaction
// (weld.rs:66-91), used by the tests as the checker of the HIP path.  Every
// function names the reference lines it follows.  Single-threaded; the
// reference's parallel loops all have order-independent results.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace orm {

using V3 = float[3];

static uint32_t f_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}
static float le_f32(const uint8_t* p) {
    float f;
    std::memcpy(&f, p, 4);  // little-endian host
    return f;
}

// Rust's `f64 as i64`: saturating, NaN -> 0.
static int64_t rust_i64(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 9223372036854775807.0) return INT64_MAX;
    if (x <= -9223372036854775808.0) return INT64_MIN;
    return (int64_t)x;
}

// Rust's f32::min / f32::max (LLVM minnum/maxnum as lowered for x86-64): a
// NaN operand, signalling or quiet, yields the other operand.  glibc's
// fmin/fmax differ: they return NaN for a signalling NaN.
template <class T>
static T rmin(T a, T b) {
    if (b != b) return a;
    if (a != a) return b;
    return b < a ? b : a;
}
template <class T>
static T rmax(T a, T b) {
    if (b != b) return a;
    if (a != a) return b;
    return b > a ? b : a;
}

// weld.rs:150-154: -0.0 and +0.0 must collide.
static uint32_t canon_bits(float x) {
    volatile float zero = 0.0f;
    return f_bits(x + zero);
}

// util.rs:58-87: union by smaller root, so every root is its set's minimum.
struct UnionFind {
    std::vector<uint32_t> parent;
    explicit UnionFind(size_t n) : parent(n) {
        for (size_t i = 0; i < n; i++) parent[i] = (uint32_t)i;
    }
    uint32_t find(uint32_t x) {
        while (parent[x] != x) {
            uint32_t gp = parent[parent[x]];
            parent[x] = gp;
            x = gp;
        }
        return x;
    }
    void unite(uint32_t a, uint32_t b) {
        uint32_t ra = find(a), rb = find(b);
        if (ra == rb) return;
        if (ra < rb) parent[rb] = ra;
        else parent[ra] = rb;
    }
};

struct Mesh {
    std::vector<float> pos;        // 3 per vertex
    std::vector<uint32_t> tris;    // 3 per triangle
    size_t nv() const { return pos.size() / 3; }
    size_t nt() const { return tris.size() / 3; }
};

// util.rs:132-164 (parts are not carried: one part per mesh)
static size_t retain_triangles(Mesh& m, const std::vector<uint8_t>& keep) {
    size_t w = 0, removed = 0;
    for (size_t t = 0; t < m.nt(); t++) {
        if (!keep[t]) {
            removed++;
            continue;
        }
        for (int k = 0; k < 3; k++) m.tris[3 * w + k] = m.tris[3 * t + k];
        w++;
    }
    m.tris.resize(3 * w);
    return removed;
}

// weld.rs:104-112
static size_t drop_collapsed(Mesh& m) {
    std::vector<uint8_t> keep(m.nt());
    for (size_t t = 0; t < m.nt(); t++) {
        const uint32_t* q = &m.tris[3 * t];
        keep[t] = q[0] != q[1] && q[1] != q[2] && q[0] != q[2];
    }
    return retain_triangles(m, keep);
}

// weld.rs:114-124
static std::vector<uint8_t> referenced(const Mesh& m) {
    std::vector<uint8_t> used(m.nv(), 0);
    for (uint32_t v : m.tris)
        if (v < used.size()) used[v] = 1;
    return used;
}

// weld.rs:126-148; f32::min/max ignore a NaN operand
static void bounds_of(const Mesh& m, const std::vector<uint8_t>& used, float lo[3], float hi[3]) {
    for (int k = 0; k < 3; k++) {
        lo[k] = 3.40282347e38f;
        hi[k] = -3.40282347e38f;
    }
    for (size_t i = 0; i < m.nv(); i++) {
        if (!used[i]) continue;
        for (int k = 0; k < 3; k++) {
            lo[k] = rmin(lo[k], m.pos[3 * i + k]);
            hi[k] = rmax(hi[k], m.pos[3 * i + k]);
        }
    }
    if (lo[0] > hi[0])
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = 0.0f;
}

struct WeldOut {
    std::vector<float> pos;
    std::vector<uint32_t> remap;
    size_t exact = 0, tol = 0;
};

// weld.rs:156-330 (weld_positions_impl)
static WeldOut weld_positions(const std::vector<float>& P, const std::vector<uint8_t>* used, float eps) {
    const size_t n = P.size() / 3;
    auto is_used = [&](size_t i) { return !used || (*used)[i]; };

    // ---- stage 1: exact duplicates, by sorting (coordinate bits, index)
    struct Key {
        uint32_t x, y, z, i;
        bool operator<(const Key& o) const {
            if (x != o.x) return x < o.x;
            if (y != o.y) return y < o.y;
            if (z != o.z) return z < o.z;
            return i < o.i;
        }
    };
    std::vector<Key> keys;
    keys.reserve(n);
    for (size_t i = 0; i < n; i++)
        if (is_used(i))
            keys.push_back({canon_bits(P[3 * i]), canon_bits(P[3 * i + 1]), canon_bits(P[3 * i + 2]), (uint32_t)i});
    std::sort(keys.begin(), keys.end());
    const size_t used_count = keys.size();
    std::vector<uint32_t> rep(n, UINT32_MAX);
    for (size_t a = 0; a < used_count;) {
        size_t b = a;
        while (b < used_count && keys[b].x == keys[a].x && keys[b].y == keys[a].y && keys[b].z == keys[a].z) {
            rep[keys[b].i] = keys[a].i;  // run head = smallest index
            b++;
        }
        a = b;
    }
    std::vector<uint32_t> uid(n, UINT32_MAX);
    std::vector<float> uniq;
    for (size_t i = 0; i < n; i++) {
        if (rep[i] == UINT32_MAX) continue;
        if (rep[i] == i) {
            uid[i] = (uint32_t)(uniq.size() / 3);
            uniq.insert(uniq.end(), &P[3 * i], &P[3 * i] + 3);
        } else {
            uid[i] = uid[rep[i]];
        }
    }
    WeldOut out;
    const size_t m = uniq.size() / 3;
    out.exact = used_count - m;

    // ---- stage 2: tolerance, through a grid of cells >= 8 eps
    std::vector<uint32_t> root;
    if (eps > 0.0f && m > 1) {
        float lo[3] = {3.40282347e38f, 3.40282347e38f, 3.40282347e38f};
        float hi[3] = {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f};
        for (size_t u = 0; u < m; u++)
            for (int k = 0; k < 3; k++) {
                lo[k] = rmin(lo[k], uniq[3 * u + k]);
                hi[k] = rmax(hi[k], uniq[3 * u + k]);
            }
        double ext = 0.0;
        for (int k = 0; k < 3; k++) ext = rmax(ext, (double)(hi[k] - lo[k]));
        const double h = rmax(rmax(8.0 * (double)eps, ext / (double)(1ull << 20)), 1e-30);
        auto cell_of = [&](const float* p, int64_t c[3]) {
            for (int k = 0; k < 3; k++) c[k] = rust_i64(std::floor(((double)p[k] - (double)lo[k]) / h));
        };
        auto pack = [](const int64_t c[3]) -> uint64_t {
            return ((uint64_t)c[0] << 42) | ((uint64_t)c[1] << 21) | (uint64_t)c[2];
        };
        std::vector<std::pair<uint64_t, uint32_t>> cells(m);
        for (size_t u = 0; u < m; u++) {
            int64_t c[3];
            cell_of(&uniq[3 * u], c);
            cells[u] = {pack(c), (uint32_t)u};
        }
        std::sort(cells.begin(), cells.end());
        const double e = (double)eps, eps2 = e * e;
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (size_t i = 0; i < m; i++) {
            const float* p = &uniq[3 * i];
            int64_t c[3];
            cell_of(p, c);
            int64_t offs[3][2] = {{0, 0}, {0, 0}, {0, 0}};
            int nofs[3] = {1, 1, 1};
            for (int k = 0; k < 3; k++) {
                const double base = (double)lo[k] + (double)c[k] * h;
                const double f = (double)p[k] - base;
                if (f < e && c[k] > 0) offs[k][nofs[k]++] = -1;
                else if (h - f < e && c[k] < (1 << 21) - 1) offs[k][nofs[k]++] = 1;
            }
            for (int a = 0; a < nofs[0]; a++)
                for (int b = 0; b < nofs[1]; b++)
                    for (int d = 0; d < nofs[2]; d++) {
                        const int64_t cc[3] = {c[0] + offs[0][a], c[1] + offs[1][b], c[2] + offs[2][d]};
                        const uint64_t key = pack(cc);
                        auto it = std::lower_bound(cells.begin(), cells.end(), std::make_pair(key, (uint32_t)0));
                        for (; it != cells.end() && it->first == key; ++it) {
                            const uint32_t j = it->second;
                            if (j >= i) continue;
                            const float* q = &uniq[3 * j];
                            const double dx = (double)p[0] - (double)q[0];
                            const double dy = (double)p[1] - (double)q[1];
                            const double dz = (double)p[2] - (double)q[2];
                            if (dx * dx + dy * dy + dz * dz <= eps2) pairs.push_back({(uint32_t)i, j});
                        }
                    }
        }
        if (!pairs.empty()) {
            UnionFind uf(m);
            for (auto& pr : pairs) uf.unite(pr.first, pr.second);
            root.resize(m);
            for (size_t u = 0; u < m; u++) root[u] = uf.find((uint32_t)u);
        }
    }
    if (root.empty()) {
        out.pos = std::move(uniq);
        out.remap = std::move(uid);
        return out;
    }
    std::vector<uint32_t> cid(m);
    for (size_t u = 0; u < m; u++) {
        if (root[u] == u) {
            cid[u] = (uint32_t)(out.pos.size() / 3);
            out.pos.insert(out.pos.end(), &uniq[3 * u], &uniq[3 * u] + 3);
        } else {
            cid[u] = cid[root[u]];
        }
    }
    out.tol = m - out.pos.size() / 3;
    out.remap.resize(n);
    for (size_t i = 0; i < n; i++) out.remap[i] = uid[i] == UINT32_MAX ? UINT32_MAX : cid[uid[i]];
    return out;
}

// weld.rs:38-62.  report: input, output vertices, exact, tolerance merges, dropped triangles
static void weld(Mesh& m, float rel_tolerance, uint64_t rep[5], float* tol_mm) {
    auto used = referenced(m);
    float lo[3], hi[3];
    bounds_of(m, used, lo, hi);
    const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    const float diag = std::sqrt(dx * dx + dy * dy + dz * dz);
    const float eps = (float)((double)rmax(rel_tolerance, 0.0f) * (double)diag);
    WeldOut w = weld_positions(m.pos, &used, eps);
    rep[0] = m.nv();
    rep[1] = w.pos.size() / 3;
    rep[2] = w.exact;
    rep[3] = w.tol;
    *tol_mm = eps;
    m.pos = std::move(w.pos);
    for (auto& v : m.tris) v = w.remap[v];
    rep[4] = drop_collapsed(m);
}

// weld.rs:66-91
static size_t compact(Mesh& m) {
    const size_t dropped = drop_collapsed(m);
    auto used = referenced(m);
    bool all = true;
    for (auto u : used) all = all && u;
    if (all) return dropped;
    std::vector<uint32_t> remap(m.nv(), UINT32_MAX);
    std::vector<float> pos;
    for (size_t i = 0; i < m.nv(); i++)
        if (used[i]) {
            remap[i] = (uint32_t)(pos.size() / 3);
            pos.insert(pos.end(), &m.pos[3 * i], &m.pos[3 * i] + 3);
        }
    m.pos = std::move(pos);
    for (auto& v : m.tris) v = remap[v];
    return dropped;
}

// io/stl.rs:34-62: triangle count after the truncation rule
static int64_t stl_count(const uint8_t* b, size_t len, uint64_t* trailing, uint64_t* declared_out) {
    if (len < 84) return -1;
    uint32_t declared;
    std::memcpy(&declared, b + 80, 4);
    const uint64_t available = (len - 84) / 50;
    *declared_out = declared;
    if (declared <= available) {
        *trailing = len - 84 - 50 * (uint64_t)declared;
        return declared;
    }
    *trailing = 0;
    return (int64_t)available;
}

static void stl_parse(const uint8_t* b, size_t n, Mesh& m) {
    m.pos.resize(9 * n);
    m.tris.resize(3 * n);
    for (size_t t = 0; t < n; t++) {
        const uint8_t* rec = b + 84 + 50 * t;
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) m.pos[9 * t + 3 * k + a] = le_f32(rec + 12 + 12 * k + 4 * a);
        for (int k = 0; k < 3; k++) m.tris[3 * t + k] = (uint32_t)(3 * t + k);
    }
}

// io.rs:206-224
static size_t drop_invalid(Mesh& m) {
    const size_t nv = m.nv();
    std::vector<uint8_t> bad(nv);
    for (size_t i = 0; i < nv; i++)
        bad[i] = !(std::isfinite(m.pos[3 * i]) && std::isfinite(m.pos[3 * i + 1]) && std::isfinite(m.pos[3 * i + 2]));
    std::vector<uint8_t> keep(m.nt());
    for (size_t t = 0; t < m.nt(); t++) {
        bool ok = true;
        for (int k = 0; k < 3; k++) {
            const uint32_t v = m.tris[3 * t + k];
            ok = ok && v < nv && !bad[v];
        }
        keep[t] = ok;
    }
    return retain_triangles(m, keep);
}

static bool copy_out(const Mesh& m, float* pos, uint32_t* nv, uint32_t* tris, uint32_t* nt) {
    *nv = (uint32_t)m.nv();
    *nt = (uint32_t)m.nt();
    if (pos) std::memcpy(pos, m.pos.data(), m.pos.size() * 4);
    if (tris) std::memcpy(tris, m.tris.data(), m.tris.size() * 4);
    return true;
}

}  // namespace orm

using namespace orm;

extern "C" int64_t or_stl_count(const uint8_t* b, size_t len) {
    uint64_t trailing, declared;
    return stl_count(b, len, &trailing, &declared);
}

// Weld a mesh (weld.rs:38-62).  out_pos holds nv*3 floats, out_tris nt*3.
extern "C" int or_weld(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float rel_tolerance,
                       float* out_pos, uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t rep[5],
                       float* tol_mm) {
    Mesh m;
    m.pos.assign(pos, pos + 3 * (size_t)nv);
    m.tris.assign(tris, tris + 3 * (size_t)nt);
    for (uint32_t v : m.tris)
        if (v >= nv) return 1;  // the reference indexes out of bounds here (weld.rs:55)
    weld(m, rel_tolerance, rep, tol_mm);
    copy_out(m, out_pos, out_nv, out_tris, out_nt);
    return 0;
}

// Remove collapsed triangles and unreferenced vertices (weld.rs:66-91).
extern "C" int or_compact(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out_pos,
                          uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t* dropped) {
    Mesh m;
    m.pos.assign(pos, pos + 3 * (size_t)nv);
    m.tris.assign(tris, tris + 3 * (size_t)nt);
    for (uint32_t v : m.tris)
        if (v >= nv) return 1;
    *dropped = compact(m);
    copy_out(m, out_pos, out_nv, out_tris, out_nt);
    return 0;
}

// Binary STL through io.rs:96-122 (parse, drop_invalid, unit scale, weld or
// compact).  rep: raw triangles, raw vertices, invalid triangles dropped,
// trailing bytes, declared count, then the five weld counts (weld != 0).
// Returns 1 for a file shorter than its header, 2 for no triangles, 3 for
// only degenerate triangles (io.rs:101-103, :121-123).
extern "C" int or_load_stl(const uint8_t* b, size_t len, int32_t do_weld, float rel_tolerance, float unit_scale,
                           float* out_pos, uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt, uint64_t rep[10],
                           float* tol_mm) {
    uint64_t trailing, declared;
    const int64_t n = stl_count(b, len, &trailing, &declared);
    if (n < 0) return 1;
    Mesh m;
    stl_parse(b, (size_t)n, m);
    std::memset(rep, 0, 10 * sizeof(uint64_t));
    rep[0] = m.nt();
    rep[1] = m.nv();
    rep[3] = trailing;
    rep[4] = declared;
    *tol_mm = 0.0f;
    if (m.nt() == 0) return 2;
    rep[2] = drop_invalid(m);
    if (unit_scale != 1.0f && std::isfinite(unit_scale) && unit_scale > 0.0f)
        for (auto& c : m.pos) c *= unit_scale;
    if (do_weld) weld(m, rel_tolerance, rep + 5, tol_mm);
    else compact(m);
    if (m.nt() == 0) return 3;
    copy_out(m, out_pos, out_nv, out_tris, out_nt);
    return 0;
}
