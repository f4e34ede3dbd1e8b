// TEST INFRASTRUCTURE ONLY: CPU restatement of the patina mask's mesh steps,
// argent-mesh cavity.rs -- concavity (cavity.rs:130-243), the Laplacian
// smoothing with crease copies welded (cavity.rs:245-296, with
// weld::weld_positions at eps 0, weld.rs:92-95 / :151-154 / :158-228) and the
// blend of patina_mask (cavity.rs:49-80).  The ambient-occlusion step is
// or_ambient_occlusion (oracle.cpp).  Only the tests load this, as the checker
// of the device pipeline.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

#include "oracle.h"

namespace {

struct D3 {
    double x = 0, y = 0, z = 0;
    D3() = default;
    D3(double a, double b, double c) : x(a), y(b), z(c) {}
    D3 operator+(const D3& o) const { return {x + o.x, y + o.y, z + o.z}; }
    D3 operator-(const D3& o) const { return {x - o.x, y - o.y, z - o.z}; }
    D3 operator*(double s) const { return {x * s, y * s, z * s}; }
    D3 operator/(double s) const { return {x / s, y / s, z / s}; }
    double dot(const D3& o) const { return x * o.x + y * o.y + z * o.z; }
    D3 cross(const D3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
    double length_squared() const { return dot(*this); }
    double length() const { return std::sqrt(length_squared()); }
    // glam normalize_or_zero: times the reciprocal length when finite and positive
    D3 normalize_or_zero() const {
        const double r = 1.0 / length();
        return (std::isfinite(r) && r > 0.0) ? (*this) * r : D3{};
    }
};

D3 ld3(const float* p) { return D3{(double)p[0], (double)p[1], (double)p[2]}; }

// Rust `f64 as u64`: saturating, NaN to 0
uint64_t sat_u64(double q) {
    if (!(q > 0.0)) return 0;
    if (q >= 18446744073709551615.0) return UINT64_MAX;
    return (uint64_t)q;
}

float smoothstep(float e0, float e1, float x) {  // cavity.rs:44-47
    float t = (x - e0) / (e1 - e0);
    t = std::min(std::max(t, 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// weld::weld_positions(positions, 0.0): canonical id per vertex, numbered in
// first-occurrence order of the exact (sign-of-zero-blind) coordinates.
std::vector<uint32_t> exact_canon(const float* pos, uint32_t nv, uint32_t* n_canon) {
    auto bits = [](float x) {
        float y = x + 0.0f;  // -0.0 and +0.0 collide (weld.rs:151-154)
        uint32_t u;
        std::memcpy(&u, &y, 4);
        return u;
    };
    std::vector<uint32_t> order(nv);
    std::iota(order.begin(), order.end(), 0u);
    auto key = [&](uint32_t i, int a) { return bits(pos[3 * (size_t)i + a]); };
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        for (int k = 0; k < 3; k++)
            if (key(a, k) != key(b, k)) return key(a, k) < key(b, k);
        return a < b;
    });
    std::vector<uint32_t> rep(nv);
    for (uint32_t s = 0; s < nv;) {
        uint32_t e = s;
        while (e < nv && key(order[e], 0) == key(order[s], 0) && key(order[e], 1) == key(order[s], 1) &&
               key(order[e], 2) == key(order[s], 2))
            e++;
        for (uint32_t k = s; k < e; k++) rep[order[k]] = order[s];  // run head: smallest index
        s = e;
    }
    std::vector<uint32_t> uid(nv);
    uint32_t next = 0;
    for (uint32_t i = 0; i < nv; i++) uid[i] = rep[i] == i ? next++ : uid[rep[i]];
    *n_canon = next;
    return uid;
}

}  // namespace

// cavity.rs:130-243
extern "C" int or_concavity(const float* pos, const float* nrm, uint32_t nv, const uint32_t* tris, uint32_t nt,
                            float radius, float* out) {
    if (nv == 0) return 0;
    if (std::isnan(radius) || radius <= 0.0f) {
        std::fill(out, out + nv, 0.0f);
        return 0;
    }
    std::vector<double> varea(nv, 0.0);
    for (uint32_t t = 0; t < nt; t++) {
        const uint32_t* tr = tris + 3 * (size_t)t;
        const D3 a = ld3(pos + 3 * (size_t)tr[0]), b = ld3(pos + 3 * (size_t)tr[1]), c = ld3(pos + 3 * (size_t)tr[2]);
        const double ar = 0.5 * (b - a).cross(c - a).length() / 3.0;
        for (int k = 0; k < 3; k++) varea[tr[k]] += ar;
    }
    // Mesh::bounds (lib.rs): f32 min / max over every position
    float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
    for (uint32_t v = 0; v < nv; v++)
        for (int k = 0; k < 3; k++) lo[k] = std::min(lo[k], pos[3 * (size_t)v + k]);
    const D3 lod{(double)lo[0], (double)lo[1], (double)lo[2]};
    const double h = (double)radius / 2.0;
    const uint64_t lim21 = (1ull << 21) - 1;
    auto key_of = [&](const float* p) {
        const D3 q = (ld3(p) - lod) / h;
        const uint64_t c0 = std::min(sat_u64(std::max(q.x, 0.0)), lim21), c1 = std::min(sat_u64(std::max(q.y, 0.0)), lim21),
                       c2 = std::min(sat_u64(std::max(q.z, 0.0)), lim21);
        return (c0 << 42) | (c1 << 21) | c2;
    };
    std::vector<std::pair<uint64_t, uint32_t>> order(nv);
    for (uint32_t i = 0; i < nv; i++) order[i] = {key_of(pos + 3 * (size_t)i), i};
    std::sort(order.begin(), order.end());
    struct Cell {
        uint64_t key;
        double m;
        D3 p, n;
    };
    std::vector<Cell> cells;
    std::vector<uint32_t> cell_of(nv);
    for (uint32_t i = 0; i < nv;) {
        const uint64_t k = order[i].first;
        double m = 0.0;
        D3 p, n;
        uint32_t e = i;
        for (; e < nv && order[e].first == k; e++) {
            const uint32_t v = order[e].second;
            const double w = std::max(varea[v], 1e-12);
            m += w;
            p = p + ld3(pos + 3 * (size_t)v) * w;
            n = n + ld3(nrm + 3 * (size_t)v) * w;
            cell_of[v] = (uint32_t)cells.size();
        }
        cells.push_back(Cell{k, m, p / m, n.normalize_or_zero()});
        i = e;
    }
    const int64_t reach = (int64_t)std::ceil((double)radius / h) + 1;
    const int64_t lim = (1ll << 21) - 1;
    const double max_d = (double)radius + 2.0 * h;
    std::vector<std::vector<uint32_t>> nbrs(cells.size());
    for (size_t ci = 0; ci < cells.size(); ci++) {
        const Cell& c = cells[ci];
        const int64_t base[3] = {(int64_t)(c.key >> 42), (int64_t)((c.key >> 21) & lim21), (int64_t)(c.key & lim21)};
        for (int64_t dx = -reach; dx <= reach; dx++)
            for (int64_t dy = -reach; dy <= reach; dy++)
                for (int64_t dz = -reach; dz <= reach; dz++) {
                    const int64_t q[3] = {base[0] + dx, base[1] + dy, base[2] + dz};
                    bool out_of = false;
                    for (int k = 0; k < 3; k++) out_of |= q[k] < 0 || q[k] > lim;
                    if (out_of) continue;
                    const uint64_t key = ((uint64_t)q[0] << 42) | ((uint64_t)q[1] << 21) | (uint64_t)q[2];
                    auto it = std::lower_bound(cells.begin(), cells.end(), key,
                                               [](const Cell& a, uint64_t k) { return a.key < k; });
                    if (it != cells.end() && it->key == key && (it->p - c.p).length() <= max_d)
                        nbrs[ci].push_back((uint32_t)(it - cells.begin()));
                }
    }
    const double r2 = (double)radius * (double)radius;
    const double sg = 0.5 * (double)radius;
    const double inv2s2 = 1.0 / (2.0 * (sg * sg));
    for (uint32_t v = 0; v < nv; v++) {
        const D3 p = ld3(pos + 3 * (size_t)v);
        const D3 n = ld3(nrm + 3 * (size_t)v).normalize_or_zero();
        if (n.x == 0.0 && n.y == 0.0 && n.z == 0.0) {
            out[v] = 0.0f;
            continue;
        }
        double sw = 0.0, st = 0.0;
        D3 sd;
        for (uint32_t j : nbrs[cell_of[v]]) {
            const Cell& c = cells[j];
            const D3 d = c.p - p;
            const double dist2 = d.length_squared();
            if (dist2 > r2 || c.n.dot(n) < -0.5) continue;
            const double w = c.m * std::exp(-dist2 * inv2s2);
            sw += w;
            sd = sd + d * w;
            const double dn = d.dot(n);
            st += w * (dist2 - dn * dn);
        }
        if (sw <= 0.0 || st <= 1e-18) {
            out[v] = 0.0f;
            continue;
        }
        const double hgt = (sd / sw).dot(n);
        const double dt2 = st / sw;
        const double mean_curv = 2.0 * hgt / dt2;
        out[v] = (float)std::min(std::max(mean_curv * (double)radius, -4.0), 4.0);
    }
    return 0;
}

// cavity.rs:245-296: x <- x + lambda (mean of neighbours - x), crease copies as one
extern "C" int or_smooth(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* values,
                         uint32_t iterations, float lambda) {
    if (nv == 0) return 0;
    uint32_t nc = 0;
    const std::vector<uint32_t> canon = exact_canon(pos, nv, &nc);
    std::vector<double> sum(nc, 0.0);
    std::vector<uint32_t> cnt(nc, 0);
    for (uint32_t v = 0; v < nv; v++) {
        sum[canon[v]] += (double)values[v];
        cnt[canon[v]]++;
    }
    std::vector<float> x(nc);
    for (uint32_t c = 0; c < nc; c++) x[c] = (float)(sum[c] / (double)std::max(cnt[c], 1u));
    if (iterations > 0 && nt > 0) {
        std::vector<uint64_t> edges;
        edges.reserve(6 * (size_t)nt);
        for (uint32_t t = 0; t < nt; t++) {
            const uint64_t c0 = canon[tris[3 * (size_t)t]], c1 = canon[tris[3 * (size_t)t + 1]],
                           c2 = canon[tris[3 * (size_t)t + 2]];
            edges.push_back((c0 << 32) | c1);
            edges.push_back((c1 << 32) | c0);
            edges.push_back((c1 << 32) | c2);
            edges.push_back((c2 << 32) | c1);
            edges.push_back((c2 << 32) | c0);
            edges.push_back((c0 << 32) | c2);
        }
        std::sort(edges.begin(), edges.end());
        edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
        std::vector<uint32_t> off(nc + 1, 0);
        for (uint64_t e : edges) off[(e >> 32) + 1]++;
        for (uint32_t i = 0; i < nc; i++) off[i + 1] += off[i];
        std::vector<float> y = x;
        for (uint32_t it = 0; it < iterations; it++) {
            for (uint32_t i = 0; i < nc; i++) {
                const uint32_t a = off[i], b = off[i + 1];
                const float xi = x[i];
                if (b > a) {
                    float s = 0.0f;
                    for (uint32_t k = a; k < b; k++) {
                        const uint32_t j = (uint32_t)edges[k];
                        if (j != i) s += x[j];
                    }
                    const float m = s / (float)(b - a);
                    y[i] = xi + lambda * (m - xi);
                } else {
                    y[i] = xi;
                }
            }
            std::swap(x, y);
        }
    }
    for (uint32_t v = 0; v < nv; v++) values[v] = x[canon[v]];
    return 0;
}

// cavity.rs:64-79: the contrast-shaped blend of AO and concavity, smoothing,
// clamp.  ao / cav: per-vertex inputs (zeros when a term is off).
extern "C" int or_patina_finish(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, const float* ao,
                                const float* cav, float concavity_weight, uint32_t smooth_iterations, float* out) {
    const float w = std::min(std::max(concavity_weight, 0.0f), 1.0f);
    for (uint32_t v = 0; v < nv; v++) {
        const float at = smoothstep(0.04f, 0.55f, ao[v]);
        const float ct = smoothstep(0.02f, 0.5f, cav[v]);
        out[v] = smoothstep(0.05f, 0.9f, (1.0f - w) * at + w * ct);
    }
    or_smooth(pos, nv, tris, nt, out, smooth_iterations, 0.5f);
    for (uint32_t v = 0; v < nv; v++) out[v] = std::min(std::max(out[v], 0.0f), 1.0f);
    return 0;
}
