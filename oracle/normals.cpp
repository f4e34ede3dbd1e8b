// TEST INFRASTRUCTURE ONLY — the oracle's crease-aware normals: a CPU
// restatement of argent-mesh normals.rs (compute with one crease angle,
// :47-220; fan clustering :250-542; winding repair :544-610; shells and
// inside-out flips :612-785) with the helpers of util.rs it calls, used by
// the tests as the checker of the HIP path.  Single-threaded; the
// reference's parallel loops have order-independent results.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace orn {

using D3 = std::array<double, 3>;

static D3 d3(const float* p) { return {(double)p[0], (double)p[1], (double)p[2]}; }
static D3 sub(const D3& a, const D3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
static D3 cross(const D3& a, const D3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
static double dot(const D3& a, const D3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static double len(const D3& a) { return std::sqrt(dot(a, a)); }

// normals.rs:243-246
static void normalize_or_z(const D3& a, float* out) {
    const double l = len(a);
    if (l > 1e-300) {
        for (int k = 0; k < 3; k++) out[k] = (float)(a[k] / l);
    } else {
        out[0] = 0.0f;
        out[1] = 0.0f;
        out[2] = 1.0f;
    }
}

// util.rs:39-50
static void tri_normal_area(const std::vector<float>& P, const uint32_t* t, float* n, float* area) {
    const D3 a = d3(&P[3 * t[0]]), b = d3(&P[3 * t[1]]), c = d3(&P[3 * t[2]]);
    const D3 x = cross(sub(b, a), sub(c, a));
    const double l = len(x);
    if (l > 0.0 && std::isfinite(l)) {
        for (int k = 0; k < 3; k++) n[k] = (float)(x[k] / l);
        *area = (float)(0.5 * l);
    } else {
        n[0] = n[1] = n[2] = 0.0f;
        *area = 0.0f;
    }
}

struct Mesh {
    std::vector<float> pos;       // 3 per vertex
    std::vector<uint32_t> tris;   // 3 per triangle
    std::vector<float> nrm;       // 3 per vertex (output)
    size_t nv() const { return pos.size() / 3; }
    size_t nt() const { return tris.size() / 3; }
};

// util.rs:96-128
struct VertexFaces {
    std::vector<uint32_t> off, faces;
    VertexFaces(size_t nv, const std::vector<uint32_t>& T) : off(nv + 1, 0), faces(T.size()) {
        for (uint32_t v : T) off[v + 1]++;
        for (size_t i = 0; i < nv; i++) off[i + 1] += off[i];
        std::vector<uint32_t> cur(off.begin(), off.end() - 1);
        for (size_t f = 0; f < T.size() / 3; f++)
            for (int k = 0; k < 3; k++) faces[cur[T[3 * f + k]]++] = (uint32_t)f;
    }
};

struct UF16 {
    std::vector<uint16_t> p;
    uint16_t find(uint16_t x) {
        while (p[x] != x) {
            uint16_t g = p[p[x]];
            p[x] = g;
            x = g;
        }
        return x;
    }
    void unite(uint16_t a, uint16_t b) {
        uint16_t ra = find(a), rb = find(b);
        if (ra == rb) return;
        if (ra < rb) p[rb] = ra;
        else p[ra] = rb;
    }
};

struct UF32 {
    std::vector<uint32_t> p;
    explicit UF32(size_t n) : p(n) {
        for (size_t i = 0; i < n; i++) p[i] = (uint32_t)i;
    }
    uint32_t find(uint32_t x) {
        while (p[x] != x) {
            uint32_t g = p[p[x]];
            p[x] = g;
            x = g;
        }
        return x;
    }
    void unite(uint32_t a, uint32_t b) {
        uint32_t ra = find(a), rb = find(b);
        if (ra == rb) return;
        if (ra < rb) p[rb] = ra;
        else p[ra] = rb;
    }
};

constexpr size_t MAX_FAN = 60000;  // normals.rs:303

struct Pass {
    std::vector<uint16_t> labels;  // per CSR entry
    std::vector<uint32_t> first;   // nv + 1
    std::vector<float> normals;    // 3 per fan
    std::vector<uint32_t> vboundary;
    uint64_t boundary = 0, nonmanifold = 0, inconsistent = 0;
};

// normals.rs:375-542 (cluster_vertex), uniform thresholds (cos_join, cos_merge)
static uint16_t cluster_vertex(uint32_t v, const uint32_t* faces, size_t k, const Mesh& m,
                               const std::vector<float>& fn, double cos_join, double cos_merge, uint16_t* labels,
                               Pass& out) {
    if (k == 0) return 0;
    const D3 pv = d3(&m.pos[3 * v]);
    struct Ent {
        uint32_t w;
        uint16_t j;
        bool fwd;
        bool operator<(const Ent& o) const {
            if (w != o.w) return w < o.w;
            if (j != o.j) return j < o.j;
            return fwd < o.fwd;
        }
    };
    std::vector<Ent> ent;
    std::vector<double> angle(k);
    for (size_t j = 0; j < k; j++) {
        const uint32_t* t = &m.tris[3 * faces[j]];
        const int slot = t[0] == v ? 0 : (t[1] == v ? 1 : 2);
        const uint32_t a = t[(slot + 1) % 3], b = t[(slot + 2) % 3];
        const D3 ea = sub(d3(&m.pos[3 * a]), pv), eb = sub(d3(&m.pos[3 * b]), pv);
        angle[j] = std::atan2(len(cross(ea, eb)), dot(ea, eb));
        if (k <= MAX_FAN) {
            ent.push_back({a, (uint16_t)j, true});
            ent.push_back({b, (uint16_t)j, false});
        }
    }
    auto fn64 = [&](uint32_t f) -> D3 { return d3(&fn[3 * f]); };
    if (k > MAX_FAN) {
        D3 sum = {0, 0, 0};
        for (size_t j = 0; j < k; j++) {
            const D3 n = fn64(faces[j]);
            for (int c = 0; c < 3; c++) sum[c] += angle[j] * n[c];
        }
        std::fill(labels, labels + k, 0);
        float o[3];
        normalize_or_z(sum, o);
        out.normals.insert(out.normals.end(), o, o + 3);
        return 1;
    }
    std::sort(ent.begin(), ent.end());
    UF16 uf;
    uf.p.resize(k);
    for (size_t i = 0; i < k; i++) uf.p[i] = (uint16_t)i;
    for (size_t i = 0; i < ent.size();) {
        const uint32_t w = ent[i].w;
        size_t e = i + 1;
        while (e < ent.size() && ent[e].w == w) e++;
        const size_t mm = e - i;
        if (v < w) {
            if (mm == 1) {
                out.boundary++;
                out.vboundary[v]++;
            } else if (mm == 2) {
                if (ent[i].fwd == ent[i + 1].fwd) out.inconsistent++;
            } else {
                out.nonmanifold++;
            }
        }
        if (mm >= 2)
            for (size_t x = i; x < e; x++)
                for (size_t y = x + 1; y < e; y++) {
                    const Ent &ex = ent[x], &ey = ent[y];
                    if (ex.fwd == ey.fwd || ex.j == ey.j) continue;
                    const uint32_t fx = faces[ex.j], fy = faces[ey.j];
                    if (dot(fn64(fx), fn64(fy)) >= cos_join) uf.unite(ex.j, ey.j);
                }
        i = e;
    }
    // compact fan labels in face order
    std::vector<uint16_t> label(k), remap(k, UINT16_MAX);
    uint16_t nclusters = 0;
    for (size_t j = 0; j < k; j++) {
        const uint16_t r = uf.find((uint16_t)j);
        if (remap[r] == UINT16_MAX) remap[r] = nclusters++;
        label[j] = remap[r];
    }
    std::vector<D3> sums;
    std::vector<double> weight;
    auto accumulate = [&](size_t n) {
        sums.assign(n, D3{0, 0, 0});
        weight.assign(n, 0.0);
        for (size_t j = 0; j < k; j++) {
            const size_t l = label[j];
            const D3 nf = fn64(faces[j]);
            const double a = angle[j];
            for (int c = 0; c < 3; c++) sums[l][c] += a * nf[c];
            weight[l] += a * len(nf);
        }
    };
    accumulate(nclusters);
    if (nclusters > 1) {
        const size_t n = nclusters;
        UF16 cu;
        cu.p.resize(n);
        for (size_t i = 0; i < n; i++) cu.p[i] = (uint16_t)i;
        std::vector<D3> unit(n);
        std::vector<bool> has(n);
        for (size_t c = 0; c < n; c++) {
            const double l = len(sums[c]);
            has[c] = l > 1e-12;
            if (has[c]) unit[c] = {sums[c][0] / l, sums[c][1] / l, sums[c][2] / l};
        }
        // max_by with total_cmp: the last of equal maxima wins
        int heaviest = -1;
        for (size_t c = 0; c < n; c++)
            if (has[c] && (heaviest < 0 || !(weight[c] < weight[heaviest]))) heaviest = (int)c;
        for (size_t a = 0; a < n; a++) {
            if (!has[a]) {
                if (heaviest >= 0) cu.unite((uint16_t)a, (uint16_t)heaviest);
            } else {
                for (size_t b = a + 1; b < n; b++)
                    if (has[b] && dot(unit[a], unit[b]) >= cos_merge) cu.unite((uint16_t)a, (uint16_t)b);
            }
        }
        std::vector<uint16_t> rm(n, UINT16_MAX);
        uint16_t mcount = 0;
        for (size_t c = 0; c < n; c++) {
            const uint16_t r = cu.find((uint16_t)c);
            if (rm[r] == UINT16_MAX) rm[r] = mcount++;
        }
        if (mcount < n) {
            for (size_t j = 0; j < k; j++) label[j] = rm[cu.find(label[j])];
            nclusters = mcount;
            accumulate(nclusters);
        }
    }
    std::copy(label.begin(), label.end(), labels);
    for (size_t c = 0; c < nclusters; c++) {
        float o[3];
        normalize_or_z(sums[c], o);
        out.normals.insert(out.normals.end(), o, o + 3);
    }
    return nclusters;
}

// normals.rs:306-373
static Pass cluster_pass(const Mesh& m, const VertexFaces& vf, const std::vector<float>& fn, double cj, double cm) {
    Pass p;
    const size_t nv = m.nv();
    p.labels.assign(vf.faces.size(), 0);
    p.first.assign(1, 0);
    p.vboundary.assign(nv, 0);
    for (size_t v = 0; v < nv; v++) {
        const size_t a = vf.off[v], b = vf.off[v + 1];
        const uint16_t cnt = cluster_vertex((uint32_t)v, &vf.faces[a], b - a, m, fn, cj, cm, &p.labels[a], p);
        p.first.push_back(p.first.back() + cnt);
    }
    return p;
}

// normals.rs:548-610
static size_t repair_winding(std::vector<uint32_t>& T, const VertexFaces& vf) {
    const size_t nt = T.size() / 3;
    std::vector<uint8_t> flip(nt, 0), seen(nt, 0);
    std::vector<uint32_t> stack, comp;
    size_t total = 0;
    for (size_t seed = 0; seed < nt; seed++) {
        if (seen[seed]) continue;
        seen[seed] = 1;
        stack.push_back((uint32_t)seed);
        comp.clear();
        size_t nflip = 0;
        while (!stack.empty()) {
            const uint32_t f = stack.back();
            stack.pop_back();
            comp.push_back(f);
            const uint32_t* t = &T[3 * f];
            for (int e = 0; e < 3; e++) {
                uint32_t a = t[e], b = t[(e + 1) % 3];
                if (flip[f]) std::swap(a, b);
                size_t nb = SIZE_MAX;
                int count = 0;
                for (uint32_t i = vf.off[a]; i < vf.off[a + 1]; i++) {
                    const uint32_t g = vf.faces[i];
                    const uint32_t* tg = &T[3 * g];
                    if (g != f && (tg[0] == b || tg[1] == b || tg[2] == b)) {
                        count++;
                        nb = g;
                    }
                }
                if (count != 1 || seen[nb]) continue;
                const uint32_t* tg = &T[3 * nb];
                bool has_ab = false;
                for (int k = 0; k < 3; k++) has_ab = has_ab || (tg[k] == a && tg[(k + 1) % 3] == b);
                seen[nb] = 1;
                flip[nb] = has_ab;
                nflip += has_ab;
                stack.push_back((uint32_t)nb);
            }
        }
        if (nflip * 2 > comp.size()) {
            for (uint32_t f : comp) flip[f] = !flip[f];
            nflip = comp.size() - nflip;
        }
        total += nflip;
    }
    for (size_t f = 0; f < nt; f++)
        if (flip[f]) std::swap(T[3 * f + 1], T[3 * f + 2]);
    return total;
}

// normals.rs:761-785
static bool ray_hits(const Mesh& m, const uint32_t* t, const D3& p, const D3& d) {
    const D3 a = d3(&m.pos[3 * t[0]]), b = d3(&m.pos[3 * t[1]]), c = d3(&m.pos[3 * t[2]]);
    const D3 e1 = sub(b, a), e2 = sub(c, a);
    const D3 pv = cross(d, e2);
    const double det = dot(e1, pv);
    if (std::fabs(det) < 1e-300) return false;
    const double inv = 1.0 / det;
    const D3 tv = sub(p, a);
    const double u = dot(tv, pv) * inv;
    if (!(u >= 0.0 && u <= 1.0)) return false;
    const D3 q = cross(tv, e1);
    const double w = dot(d, q) * inv;
    if (w < 0.0 || u + w > 1.0) return false;
    return dot(e2, q) * inv > 0.0;
}

struct Report {
    uint64_t shells = 0, flipped = 0, cavity = 0, repaired = 0, boundary = 0, nonmanifold = 0, inconsistent = 0,
             degenerate = 0, split = 0;
};

// normals.rs:612-757 (Shells::build, decide_flips)
static std::vector<uint8_t> decide_flips(const Mesh& m, const std::vector<uint32_t>& comp, size_t ns,
                                         const std::vector<uint32_t>& vboundary, Report& rep) {
    std::vector<uint8_t> flip(ns, 0);
    if (ns == 0) return flip;
    std::vector<uint32_t> refv(ns, UINT32_MAX);
    std::vector<std::array<float, 3>> lo(ns, {3.40282347e38f, 3.40282347e38f, 3.40282347e38f}),
        hi(ns, {-3.40282347e38f, -3.40282347e38f, -3.40282347e38f});
    for (size_t v = 0; v < comp.size(); v++) {
        const uint32_t c = comp[v];
        if (c == UINT32_MAX) continue;
        if (refv[c] == UINT32_MAX) refv[c] = (uint32_t)v;
        for (int k = 0; k < 3; k++) {
            const float x = m.pos[3 * v + k];
            lo[c][k] = x < lo[c][k] ? x : lo[c][k];  // finite positions: f32::min
            hi[c][k] = x > hi[c][k] ? x : hi[c][k];
        }
    }
    std::vector<size_t> faces(ns, 0);
    std::vector<double> vol(ns, 0.0);
    for (size_t f = 0; f < m.nt(); f++) {
        const uint32_t* t = &m.tris[3 * f];
        const uint32_t c = comp[t[0]];
        faces[c]++;
        const D3 o = d3(&m.pos[3 * refv[c]]);
        const D3 a = sub(d3(&m.pos[3 * t[0]]), o), b = sub(d3(&m.pos[3 * t[1]]), o), d = sub(d3(&m.pos[3 * t[2]]), o);
        vol[c] += dot(a, cross(b, d)) / 6.0;
    }
    std::vector<uint64_t> boundary(ns, 0);
    for (size_t v = 0; v < comp.size(); v++)
        if (comp[v] != UINT32_MAX) boundary[comp[v]] += vboundary[v];
    auto closed = [&](size_t c) { return (double)boundary[c] <= 0.01 * 1.5 * (double)faces[c]; };
    std::vector<size_t> cand;
    for (size_t c = 0; c < ns; c++)
        if (closed(c) && vol[c] < 0.0) cand.push_back(c);
    if (cand.empty()) return flip;
    if (cand.size() > 256) {
        double total = 0.0;
        for (size_t c = 0; c < ns; c++)
            if (closed(c)) total += vol[c];
        if (total < 0.0)
            for (size_t c : cand) flip[c] = 1;
        else
            rep.cavity += cand.size();
        for (auto f : flip) rep.flipped += f;
        return flip;
    }
    auto contains = [&](size_t o, size_t i) {
        if (o == i) return false;
        for (int k = 0; k < 3; k++)
            if (!(lo[o][k] <= lo[i][k] && hi[o][k] >= hi[i][k])) return false;
        return true;
    };
    uint64_t budget = 300000000;
    for (size_t c : cand) {
        std::vector<size_t> enc;
        for (size_t o = 0; o < ns; o++)
            if (closed(o) && contains(o, c)) enc.push_back(o);
        if (enc.empty()) {
            flip[c] = 1;
            continue;
        }
        uint64_t cost = 0;
        for (size_t o : enc) cost += faces[o];
        if (cost > budget) {
            rep.cavity++;
            continue;
        }
        budget -= cost;
        const D3 p = d3(&m.pos[3 * refv[c]]);
        const D3 dir = {0.5773502691896258, 0.5773519, 0.5773486};  // f64 literals (normals.rs:736)
        size_t crossings = 0;
        for (size_t f = 0; f < m.nt(); f++) {
            const uint32_t* t = &m.tris[3 * f];
            const uint32_t sc = comp[t[0]];
            if (std::find(enc.begin(), enc.end(), (size_t)sc) == enc.end()) continue;
            crossings += ray_hits(m, t, p, dir);
        }
        if (crossings % 2 == 1) rep.cavity++;
        else flip[c] = 1;
    }
    for (auto f : flip) rep.flipped += f;
    return flip;
}

// normals.rs:94-99 (Thresholds::of_deg)
static void thresholds(float deg, double* cj, double* cm) {
    const float d = std::min(std::max(deg, 0.05f), 180.0f);
    const double pi = 3.14159265358979323846;
    *cj = std::cos((double)d * (pi / 180.0));
    const float h = std::min(std::max(0.5f * d, 0.05f), 15.0f);
    *cm = std::cos((double)h * (pi / 180.0));
}

// normals.rs:120-220 (compute_impl, uniform crease)
static void compute(Mesh& m, float crease_deg, Report& rep) {
    {
        std::vector<uint32_t> keep;
        for (size_t f = 0; f < m.nt(); f++) {
            const uint32_t* t = &m.tris[3 * f];
            if (t[0] != t[1] && t[1] != t[2] && t[0] != t[2]) keep.insert(keep.end(), t, t + 3);
        }
        m.tris.swap(keep);
    }
    const size_t nv = m.nv();
    if (m.nt() == 0) {
        m.nrm.assign(3 * nv, 0.0f);
        for (size_t v = 0; v < nv; v++) m.nrm[3 * v + 2] = 1.0f;
        return;
    }
    double cj, cm;
    thresholds(crease_deg, &cj, &cm);
    const VertexFaces vf(nv, m.tris);
    auto face_normals = [&](std::vector<float>& fn, std::vector<float>* area) {
        fn.assign(3 * m.nt(), 0.0f);
        if (area) area->assign(m.nt(), 0.0f);
        for (size_t f = 0; f < m.nt(); f++) {
            float a;
            tri_normal_area(m.pos, &m.tris[3 * f], &fn[3 * f], &a);
            if (area) (*area)[f] = a;
        }
    };
    std::vector<float> fn, area;
    face_normals(fn, &area);
    for (float a : area) rep.degenerate += a <= 0.0f;
    Pass pass = cluster_pass(m, vf, fn, cj, cm);
    if (pass.inconsistent > 0) {
        const size_t flipped = repair_winding(m.tris, vf);
        if (flipped > 0) {
            rep.repaired = flipped;
            face_normals(fn, nullptr);
            pass = cluster_pass(m, vf, fn, cj, cm);
        }
    }
    rep.boundary = pass.boundary;
    rep.nonmanifold = pass.nonmanifold;
    rep.inconsistent = pass.inconsistent;
    // shells
    UF32 uf(nv);
    std::vector<uint8_t> used(nv, 0);
    for (size_t f = 0; f < m.nt(); f++) {
        const uint32_t* t = &m.tris[3 * f];
        uf.unite(t[0], t[1]);
        uf.unite(t[0], t[2]);
        used[t[0]] = used[t[1]] = used[t[2]] = 1;
    }
    std::vector<uint32_t> comp(nv, UINT32_MAX), id_of_root(nv, UINT32_MAX);
    uint32_t count = 0;
    for (size_t v = 0; v < nv; v++) {
        if (!used[v]) continue;
        const uint32_t r = uf.find((uint32_t)v);
        if (id_of_root[r] == UINT32_MAX) id_of_root[r] = count++;
        comp[v] = id_of_root[r];
    }
    rep.shells = count;
    const std::vector<uint8_t> flip = decide_flips(m, comp, count, pass.vboundary, rep);
    bool any = false;
    for (auto f : flip) any = any || f;
    if (any) {
        for (size_t f = 0; f < m.nt(); f++) {
            uint32_t* t = &m.tris[3 * f];
            if (flip[comp[t[0]]]) std::swap(t[1], t[2]);
        }
        for (size_t v = 0; v < nv; v++)
            if (comp[v] != UINT32_MAX && flip[comp[v]])
                for (uint32_t c = pass.first[v]; c < pass.first[v + 1]; c++)
                    for (int k = 0; k < 3; k++) pass.normals[3 * c + k] = -pass.normals[3 * c + k];
    }
    // assemble split vertices
    std::vector<uint32_t> extra(nv + 1, 0);
    for (size_t v = 0; v < nv; v++) {
        const uint32_t cnt = pass.first[v + 1] - pass.first[v];
        extra[v + 1] = extra[v] + (cnt > 0 ? cnt - 1 : 0);
    }
    rep.split = extra[nv];
    std::vector<float> normals(3 * nv);
    for (size_t v = 0; v < nv; v++) {
        if (pass.first[v + 1] > pass.first[v])
            for (int k = 0; k < 3; k++) normals[3 * v + k] = pass.normals[3 * pass.first[v] + k];
        else {
            normals[3 * v] = 0.0f;
            normals[3 * v + 1] = 0.0f;
            normals[3 * v + 2] = 1.0f;
        }
    }
    for (size_t v = 0; v < nv; v++)
        for (uint32_t c = pass.first[v] + 1; c < pass.first[v + 1]; c++) {
            normals.insert(normals.end(), &pass.normals[3 * c], &pass.normals[3 * c] + 3);
            m.pos.insert(m.pos.end(), m.pos.begin() + 3 * v, m.pos.begin() + 3 * v + 3);
        }
    const std::vector<uint32_t> src = m.tris;
    for (size_t f = 0; f < m.nt(); f++)
        for (int k = 0; k < 3; k++) {
            const uint32_t v = src[3 * f + k];
            const uint32_t* beg = &vf.faces[vf.off[v]];
            const uint32_t* end = &vf.faces[vf.off[v + 1]];
            const size_t j = std::lower_bound(beg, end, (uint32_t)f) - beg;
            const uint32_t label = pass.labels[vf.off[v] + j];
            m.tris[3 * f + k] = label == 0 ? v : (uint32_t)nv + extra[v] + label - 1;
        }
    m.nrm = std::move(normals);
}

}  // namespace orn

using namespace orn;

// normals::compute (normals.rs:47-49).  out_pos / out_nrm hold up to
// (nv + 3 nt) * 3 floats, out_tris nt * 3; rep: shells, flipped shells,
// cavity shells, repaired faces, boundary, non-manifold and inconsistent
// edges, degenerate faces, split vertices.
extern "C" int or_normals(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float crease_deg,
                          float* out_pos, float* out_nrm, uint32_t* out_nv, uint32_t* out_tris, uint32_t* out_nt,
                          uint64_t* rep9) {
    Mesh m;
    m.pos.assign(pos, pos + 3 * (size_t)nv);
    m.tris.assign(tris, tris + 3 * (size_t)nt);
    for (uint32_t v : m.tris)
        if (v >= nv) return 1;
    Report r;
    compute(m, crease_deg, r);
    *out_nv = (uint32_t)m.nv();
    *out_nt = (uint32_t)m.nt();
    std::memcpy(out_pos, m.pos.data(), m.pos.size() * 4);
    std::memcpy(out_nrm, m.nrm.data(), m.nrm.size() * 4);
    std::memcpy(out_tris, m.tris.data(), m.tris.size() * 4);
    const uint64_t v[9] = {r.shells, r.flipped, r.cavity, r.repaired, r.boundary, r.nonmanifold, r.inconsistent,
                           r.degenerate, r.split};
    std::memcpy(rep9, v, sizeof v);
    return 0;
}

// normals::smooth_vertex_normals (normals.rs:226-242): per vertex, the face
// normals of its triangles weighted by the corner angle, summed in f64 in
// triangle then corner order, normalized (or +z).  out: 3 * nv floats.
extern "C" int or_smooth_vertex_normals(const float* pos, uint32_t nv, const uint32_t* tris, uint32_t nt, float* out) {
    const std::vector<float> P(pos, pos + 3 * (size_t)nv);
    for (size_t i = 0; i < 3 * (size_t)nt; i++)
        if (tris[i] >= nv) return 1;
    std::vector<double> acc(3 * (size_t)nv, 0.0);
    for (uint32_t f = 0; f < nt; f++) {
        const uint32_t* t = tris + 3 * (size_t)f;
        float fn[3], area;
        tri_normal_area(P, t, fn, &area);
        for (int k = 0; k < 3; k++) {
            const uint32_t v = t[k];
            const D3 p = d3(&P[3 * (size_t)v]);
            const D3 ea = sub(d3(&P[3 * (size_t)t[(k + 1) % 3]]), p);
            const D3 eb = sub(d3(&P[3 * (size_t)t[(k + 2) % 3]]), p);
            const double ang = std::atan2(len(cross(ea, eb)), dot(ea, eb));
            for (int i = 0; i < 3; i++) acc[3 * (size_t)v + i] += ang * (double)fn[i];
        }
    }
    for (uint32_t v = 0; v < nv; v++)
        normalize_or_z(D3{acc[3 * (size_t)v], acc[3 * (size_t)v + 1], acc[3 * (size_t)v + 2]}, out + 3 * (size_t)v);
    return 0;
}
