// TEST INFRASTRUCTURE ONLY — the oracle.  A CPU restatement of the
// reference's render core, used by the tests as the checker of the HIP path
// and by the benchmark as the timed CPU baseline.  Every function names the
// reference file:line it follows.  Built with -ffp-contract=off so every
// expression rounds as the reference's (x86-64, no FMA) code does.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <future>
#include <limits>
#include <thread>
#include <vector>

#include "../argent_amd/csrc/tables/spectral_data.h"

namespace orc {

const float PI = 3.14159265358979323846f;
const float FRAC_1_PI = 0.318309886183790671538f;
const float FRAC_PI_2 = 1.57079632679489661923f;
const float FRAC_PI_4 = 0.785398163397448309616f;
const float INF = std::numeric_limits<float>::infinity();
const int N = 4;

// ------------------------------------------------------------------ glam Vec3A
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    V3 operator+(V3 o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(V3 o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(float s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(float s) const { return V3(x / s, y / s, z / s); }
    float dot(V3 o) const { return (x * o.x + y * o.y) + z * o.z; }
    V3 cross(V3 o) const { return V3(y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x); }
    float length_squared() const { return dot(*this); }
    float length() const { return std::sqrt(dot(*this)); }
    V3 normalize() const { return *this * (1.0f / length()); }
    V3 abs() const { return V3(std::fabs(x), std::fabs(y), std::fabs(z)); }
    float max_element() const { return std::fmax(std::fmax(x, y), z); }
};
static V3 ld(const float* p) { return V3(p[0], p[1], p[2]); }

struct Ray {
    V3 o, d;
    V3 at(float t) const { return o + d * t; }
};

// geom.rs:26-31
V3 offset_origin(V3 p, V3 ng, V3 dir) {
    float scale = 1e-4f * (1.0f + p.abs().max_element());
    V3 n = ng.dot(dir) >= 0.0f ? ng : -ng;
    return p + n * scale;
}
// geom.rs:33-66
struct Frame {
    V3 t, b, n;
    explicit Frame(V3 nn) : n(nn) {
        float sign = std::copysign(1.0f, nn.z);
        float a = -1.0f / (sign + nn.z);
        float bb = nn.x * nn.y * a;
        t = V3(1.0f + sign * nn.x * nn.x * a, sign * bb, -sign * nn.x);
        b = V3(bb, sign + nn.y * nn.y * a, -nn.y);
    }
    V3 to_local(V3 v) const { return V3(v.dot(t), v.dot(b), v.dot(n)); }
    V3 to_world(V3 v) const { return t * v.x + b * v.y + n * v.z; }
};
V3 reflect(V3 wo, V3 n) { return -wo + n * (2.0f * wo.dot(n)); }
// geom.rs:75-85
bool refract(V3 wi, V3 n, float eta, V3* out) {
    float cos_i = n.dot(wi);
    float sin2_i = std::fmax(1.0f - cos_i * cos_i, 0.0f);
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return false;
    float cos_t = std::sqrt(1.0f - sin2_t);
    *out = (-wi) / eta + n * (cos_i / eta - cos_t);
    return true;
}
// geom.rs:94-97
V3 rot_y(V3 v, float deg) {
    float r = deg * (PI / 180.0f);
    float s = std::sin(r), c = std::cos(r);
    return V3(c * v.x + s * v.z, v.y, -s * v.x + c * v.z);
}

// ------------------------------------------------------------------ spec.rs
struct Spec {
    float v[N];
    static Spec splat(float s) { return Spec{{s, s, s, s}}; }
    Spec operator+(const Spec& o) const { return Spec{{v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2], v[3] + o.v[3]}}; }
    Spec operator*(const Spec& o) const { return Spec{{v[0] * o.v[0], v[1] * o.v[1], v[2] * o.v[2], v[3] * o.v[3]}}; }
    Spec operator*(float s) const { return Spec{{v[0] * s, v[1] * s, v[2] * s, v[3] * s}}; }
    Spec operator/(float s) const { return Spec{{v[0] / s, v[1] / s, v[2] / s, v[3] / s}}; }
    float max_value() const {
        float m = -std::numeric_limits<float>::max();
        for (float a : v) m = std::fmax(m, a);
        return m;
    }
    float min_value() const {
        float m = std::numeric_limits<float>::max();
        for (float a : v) m = std::fmin(m, a);
        return m;
    }
    float average() const {
        float s = 0.0f;
        for (float a : v) s += a;
        return s / (float)N;
    }
    bool is_black() const { return v[0] == 0.0f && v[1] == 0.0f && v[2] == 0.0f && v[3] == 0.0f; }
    bool is_finite() const { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]) && std::isfinite(v[3]); }
    static Spec lerp(float t, const Spec& a, const Spec& b) { return a * (1.0f - t) + b * t; }
};

struct Wavelengths {
    float lambda[N], pdf[N];
    // spec.rs:209-220 (Rust's f32::atanh is 0.5 * ln_1p(2x / (1 - x)))
    static float sample_visible_wavelength(float u) {
        float x = 0.8569106f - 1.827502f * u;
        return 538.0f - 138.88889f * (0.5f * std::log1p((2.0f * x) / (1.0f - x)));
    }
    static float visible_pdf(float l) {
        if (!(l >= 360.0f && l <= 830.0f)) return 0.0f;
        float c = std::cosh(0.0072f * (l - 538.0f));
        return 0.0039398042f / (c * c);
    }
    // spec.rs:157-169
    static Wavelengths sample_visible(float u) {
        Wavelengths w;
        for (int i = 0; i < N; i++) {
            float up = u + (float)i / (float)N;
            if (up > 1.0f) up -= 1.0f;
            w.lambda[i] = sample_visible_wavelength(up);
            w.pdf[i] = visible_pdf(w.lambda[i]);
        }
        return w;
    }
    bool secondary_terminated() const { return pdf[1] == 0.0f && pdf[2] == 0.0f && pdf[3] == 0.0f; }
    // spec.rs:187-195
    void terminate_secondary() {
        if (secondary_terminated()) return;
        for (int i = 1; i < N; i++) pdf[i] = 0.0f;
        pdf[0] /= (float)N;
    }
};

// ------------------------------------------------------------------ sampler.rs
uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
const uint32_t SOBOL_DIRS[4][32] = {
    {0x80000000u, 0x40000000u, 0x20000000u, 0x10000000u, 0x08000000u, 0x04000000u, 0x02000000u, 0x01000000u,
     0x00800000u, 0x00400000u, 0x00200000u, 0x00100000u, 0x00080000u, 0x00040000u, 0x00020000u, 0x00010000u,
     0x00008000u, 0x00004000u, 0x00002000u, 0x00001000u, 0x00000800u, 0x00000400u, 0x00000200u, 0x00000100u,
     0x00000080u, 0x00000040u, 0x00000020u, 0x00000010u, 0x00000008u, 0x00000004u, 0x00000002u, 0x00000001u},
    {0x80000000u, 0xc0000000u, 0xa0000000u, 0xf0000000u, 0x88000000u, 0xcc000000u, 0xaa000000u, 0xff000000u,
     0x80800000u, 0xc0c00000u, 0xa0a00000u, 0xf0f00000u, 0x88880000u, 0xcccc0000u, 0xaaaa0000u, 0xffff0000u,
     0x80008000u, 0xc000c000u, 0xa000a000u, 0xf000f000u, 0x88008800u, 0xcc00cc00u, 0xaa00aa00u, 0xff00ff00u,
     0x80808080u, 0xc0c0c0c0u, 0xa0a0a0a0u, 0xf0f0f0f0u, 0x88888888u, 0xccccccccu, 0xaaaaaaaau, 0xffffffffu},
    {0x80000000u, 0xc0000000u, 0x60000000u, 0x90000000u, 0xe8000000u, 0x5c000000u, 0x8e000000u, 0xc5000000u,
     0x68800000u, 0x9cc00000u, 0xee600000u, 0x55900000u, 0x80680000u, 0xc09c0000u, 0x60ee0000u, 0x90550000u,
     0xe8808000u, 0x5cc0c000u, 0x8e606000u, 0xc5909000u, 0x6868e800u, 0x9c9c5c00u, 0xeeee8e00u, 0x5555c500u,
     0x8000e880u, 0xc0005cc0u, 0x60008e60u, 0x9000c590u, 0xe8006868u, 0x5c009c9cu, 0x8e00eeeeu, 0xc5005555u},
    {0x80000000u, 0xc0000000u, 0x20000000u, 0x50000000u, 0xf8000000u, 0x74000000u, 0xa2000000u, 0x93000000u,
     0xd8800000u, 0x25400000u, 0x59e00000u, 0xe6d00000u, 0x78080000u, 0xb40c0000u, 0x82020000u, 0xc3050000u,
     0x208f8000u, 0x51474000u, 0xfbea2000u, 0x75d93000u, 0xa0858800u, 0x914e5400u, 0xdbe79e00u, 0x25db6d00u,
     0x58800080u, 0xe54000c0u, 0x79e00020u, 0xb6d00050u, 0x800800f8u, 0xc00c0074u, 0x200200a2u, 0x50050093u}};
uint32_t reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0f0f0f0fu) | ((x & 0x0f0f0f0fu) << 4);
    x = ((x >> 8) & 0x00ff00ffu) | ((x & 0x00ff00ffu) << 8);
    return (x >> 16) | (x << 16);
}
uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
uint32_t nested_scramble(uint32_t x, uint32_t seed) { return reverse_bits(lk_permute(reverse_bits(x), seed)); }
// sample_4d(index, set, seed)[k] in place of sobol_burley (sampler.rs:45-51):
// Burley 2020 Owen-scrambled Sobol, shuffled index, Joe-Kuo directions.
float rng_dim(uint32_t seed, uint32_t index, uint32_t dim) {
    if (dim >= 256u || index >= 65536u) {
        uint32_t h = hash32(seed ^ hash32(index ^ hash32(dim ^ 0x2f0b6e35u)));
        return (float)(h >> 8) * (1.0f / 16777216.0f);
    }
    uint32_t set = dim / 4u, k = dim % 4u;
    uint32_t s = hash32(seed ^ hash32(set * 0x9e3779b9u + 0x632be5abu));
    uint32_t i = nested_scramble(index, s);
    uint32_t x = 0;
    for (int b = 0; b < 32; b++)
        if ((i >> b) & 1u) x ^= SOBOL_DIRS[k][b];
    uint32_t hk = hash32(k + 1u);
    x = nested_scramble(x, s ^ (hk + (s << 6) + (s >> 2)));
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}
struct Sampler {
    uint32_t index, seed, dim;
    uint64_t rng;
    Sampler(uint32_t pixel_seed, uint32_t sample_index)
        : index(sample_index), seed(pixel_seed), dim(0),
          rng((((uint64_t)pixel_seed << 32) | (uint64_t)sample_index) ^ 0x9e3779b97f4a7c15ull) {}
    float next() {
        uint32_t d = dim++;
        return rng_dim(seed, index, d);
    }
    // sampler.rs:64-70
    float next_random() {
        rng = rng * 6364136223846793005ull + 1442695040888963407ull;
        uint64_t x = ((rng >> 18) ^ rng) >> 27;
        uint32_t rot = (uint32_t)(rng >> 59);
        uint32_t xx = (uint32_t)x;
        uint32_t v = (xx >> rot) | (xx << ((32u - rot) & 31u));
        return (float)(v >> 8) * (1.0f / (float)(1u << 24));
    }
    void set_dim(uint32_t d) { dim = d; }
};
float tent(float u, float r) {
    u = 2.0f * u;
    return u < 1.0f ? r * (std::sqrt(u) - 1.0f) : r * (1.0f - std::sqrt(2.0f - u));
}
void disk(float u, float v, float* x, float* y) {
    float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
    if (a == 0.0f && b == 0.0f) {
        *x = 0.0f;
        *y = 0.0f;
        return;
    }
    float r, phi;
    if (std::fabs(a) > std::fabs(b)) {
        r = a;
        phi = FRAC_PI_4 * (b / a);
    } else {
        r = b;
        phi = FRAC_PI_2 - FRAC_PI_4 * (a / b);
    }
    *x = r * std::cos(phi);
    *y = r * std::sin(phi);
}
V3 cosine_hemisphere(float u, float v) {
    float x, y;
    disk(u, v, &x, &y);
    return V3(x, y, std::sqrt(std::fmax(1.0f - x * x - y * y, 0.0f)));
}

// ------------------------------------------------------------------ tables
// tab.rs:44-52
float tab_eval(const float* values, float l) {
    float x = std::clamp(l - 360.0f, 0.0f, 470.0f);
    size_t i = std::min((size_t)x, (size_t)469);
    float f = x - (float)i;
    return values[i] * (1.0f - f) + values[i + 1] * f;
}
Spec tab_sample(const float* v, const Wavelengths& wl) {
    Spec r;
    for (int i = 0; i < N; i++) r.v[i] = tab_eval(v, wl.lambda[i]);
    return r;
}
// spectrum.rs:80-93 with the lambda list 360 + 5 i
float piecewise(const float* value, float l) {
    static float lam[95];
    static bool init = [] {
        for (int i = 0; i < 95; i++) lam[i] = 360.0f + 5.0f * (float)i;
        return true;
    }();
    (void)init;
    if (l <= lam[0]) return value[0];
    if (l >= lam[94]) return value[94];
    size_t i = (size_t)(std::upper_bound(lam, lam + 95, l) - lam) - 1;
    float t = (l - lam[i]) / (lam[i + 1] - lam[i]);
    return value[i] * (1.0f - t) + value[i + 1] * t;
}
// cie.rs:38-49
void xyz_bar(float lambda, float out[3]) {
    float x = lambda - 360.0f;
    if (!(x >= 0.0f && x <= 470.0f)) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    size_t i = std::min((size_t)x, (size_t)469);
    float t = x - (float)i;
    const float* a = &ag_data::CMF_1931_2DEG[3 * i];
    const float* b = a + 3;
    for (int k = 0; k < 3; k++) out[k] = a[k] + (b[k] - a[k]) * t;
}
const float CIE_Y_INTEGRAL = 106.85692f;
// cie.rs:55-70
void to_xyz(const Spec& l, const Wavelengths& wl, float xyz[3]) {
    xyz[0] = xyz[1] = xyz[2] = 0.0f;
    for (int i = 0; i < N; i++) {
        float p = wl.pdf[i];
        if (p == 0.0f || l.v[i] == 0.0f) continue;
        float c[3];
        xyz_bar(wl.lambda[i], c);
        float w = l.v[i] / p;
        xyz[0] += c[0] * w;
        xyz[1] += c[1] * w;
        xyz[2] += c[2] * w;
    }
    float s = 1.0f / ((float)N * CIE_Y_INTEGRAL);
    for (int k = 0; k < 3; k++) xyz[k] *= s;
}
// spectrum.rs:413-420 over a 1 nm table
float luminance_y_tab(const float* t) {
    double y = 0.0;
    for (int i = 0; i <= 470; i++) {
        float l = 360.0f + (float)i;
        float c[3];
        xyz_bar(l, c);
        y += (double)tab_eval(t, l) * (double)c[1];
    }
    return (float)(y / (double)CIE_Y_INTEGRAL);
}

// color.rs:8-55
typedef float M3[3][3];
void m3mul(const M3 a, const M3 b, M3 r) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float s = 0.0f;
            for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
            r[i][j] = s;
        }
}
void camera_matrix(bool has_white, const float* w, M3 out) {
    const M3 BRADFORD = {{0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}};
    const M3 BRADFORD_INV = {{0.986993f, -0.147054f, 0.159963f}, {0.432305f, 0.51836f, 0.049291f}, {-0.008529f, 0.040043f, 0.968487f}};
    const M3 XYZ_TO_SRGB = {{3.2404542f, -1.5371385f, -0.4985314f}, {-0.969266f, 1.8760108f, 0.041556f}, {0.0556434f, -0.2040259f, 1.0572252f}};
    const float D65[3] = {0.95047f, 1.0f, 1.08883f};
    if (!has_white) {
        std::memcpy(out, XYZ_TO_SRGB, sizeof(M3));
        return;
    }
    auto apply = [](const M3 m, const float* v, float* r) {
        for (int i = 0; i < 3; i++) r[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
    };
    float src[3], dst[3];
    apply(BRADFORD, w, src);
    apply(BRADFORD, D65, dst);
    M3 d = {{dst[0] / src[0], 0, 0}, {0, dst[1] / src[1], 0}, {0, 0, dst[2] / src[2]}};
    M3 t1, cat;
    m3mul(d, BRADFORD, t1);
    m3mul(BRADFORD_INV, t1, cat);
    m3mul(XYZ_TO_SRGB, cat, out);
}

// ------------------------------------------------------------------ fresnel.rs
float fresnel_dielectric(float cos_theta_i, float eta) {
    float cos_i = std::clamp(cos_theta_i, -1.0f, 1.0f);
    if (cos_i < 0.0f) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    float sin2_i = 1.0f - cos_i * cos_i;
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = std::sqrt(std::fmax(1.0f - sin2_t, 0.0f));
    float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (r_parl * r_parl + r_perp * r_perp);
}
struct C {
    float re, im;
    C add(C o) const { return {re + o.re, im + o.im}; }
    C sub(C o) const { return {re - o.re, im - o.im}; }
    C mul(C o) const { return {re * o.re - im * o.im, re * o.im + im * o.re}; }
    C scale(float s) const { return {re * s, im * s}; }
    C div(C o) const {
        float d = o.re * o.re + o.im * o.im;
        return {(re * o.re + im * o.im) / d, (im * o.re - re * o.im) / d};
    }
    float norm() const { return re * re + im * im; }
    C sqrt() const {
        float n = std::sqrt(norm());
        if (n == 0.0f) return {0.0f, 0.0f};
        float t1 = std::sqrt(0.5f * (n + std::fabs(re)));
        float t2 = 0.5f * im / t1;
        if (re >= 0.0f) return {t1, t2};
        return {std::fabs(t2), std::copysign(t1, im)};
    }
};
float fresnel_conductor1(float cos_theta_i, float eta, float k) {
    float cos_i = std::clamp(cos_theta_i, 0.0f, 1.0f);
    float sin2_i = 1.0f - cos_i * cos_i;
    C e{eta, k};
    C sin2_t = C{sin2_i, 0.0f}.div(e.mul(e));
    C cos_t = C{1.0f, 0.0f}.sub(sin2_t).sqrt();
    C ec = e.scale(cos_i);
    C r_parl = ec.sub(cos_t).div(ec.add(cos_t));
    C ci{cos_i, 0.0f};
    C ect = e.mul(cos_t);
    C r_perp = ci.sub(ect).div(ci.add(ect));
    return 0.5f * (r_parl.norm() + r_perp.norm());
}
Spec fresnel_conductor(float c, const Spec& eta, const Spec& k) {
    Spec r;
    for (int i = 0; i < N; i++) r.v[i] = fresnel_conductor1(c, eta.v[i], k.v[i]);
    return r;
}

// ------------------------------------------------------------------ microfacet.rs
struct Ggx {
    float alpha;
    static Ggx from_roughness(float r) {
        r = std::clamp(r, 0.0f, 1.0f);
        return Ggx{std::fmax(r * r, 1e-4f)};
    }
    bool is_smooth() const { return alpha < 1.5e-3f; }
    float d(V3 h) const {
        if (h.z <= 0.0f) return 0.0f;
        float a2 = alpha * alpha;
        float c2 = h.z * h.z;
        float k = c2 * (a2 - 1.0f) + 1.0f;
        return a2 / (PI * k * k);
    }
    float lambda(V3 w) const {
        float c2 = w.z * w.z;
        float t2 = std::fmax(1.0f - c2, 0.0f) / std::fmax(c2, 1e-12f);
        if (!std::isfinite(t2)) return 0.0f;
        return 0.5f * (-1.0f + std::sqrt(1.0f + alpha * alpha * t2));
    }
    float g1(V3 w) const { return 1.0f / (1.0f + lambda(w)); }
    float g2(V3 wo, V3 wi) const { return 1.0f / (1.0f + lambda(wo) + lambda(wi)); }
    float d_visible(V3 wo, V3 h) const {
        float c = std::fmax(std::fabs(wo.z), 1e-8f);
        return g1(wo) * std::fmax(wo.dot(h), 0.0f) * d(h) / c;
    }
    V3 sample_h(V3 wo, float u1, float u2) const {
        float a = alpha;
        V3 vh = V3(a * wo.x, a * wo.y, wo.z).normalize();
        float lensq = vh.x * vh.x + vh.y * vh.y;
        V3 t1 = lensq > 1e-12f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : V3(1.0f, 0.0f, 0.0f);
        V3 t2 = vh.cross(t1);
        float r = std::sqrt(u1);
        float phi = 2.0f * PI * u2;
        float p1 = r * std::cos(phi);
        float p2 = r * std::sin(phi);
        float s = 0.5f * (1.0f + vh.z);
        p2 = (1.0f - s) * std::sqrt(std::fmax(1.0f - p1 * p1, 0.0f)) + s * p2;
        V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::fmax(1.0f - p1 * p1 - p2 * p2, 0.0f));
        return V3(a * nh.x, a * nh.y, std::fmax(nh.z, 1e-6f)).normalize();
    }
};
const int EN = 32;
const float* energy_table() {
    static std::vector<float> e = [] {
        std::vector<float> t(EN * EN);
        for (int j = 0; j < EN; j++) {
            float r = (float)j / (float)(EN - 1);
            Ggx g{std::fmax(r * r, 1e-4f)};
            for (int i = 0; i < EN; i++) {
                float mu = std::fmax((float)i / (float)(EN - 1), 1e-3f);
                V3 wo(std::sqrt(1.0f - mu * mu), 0.0f, mu);
                const int n = 24;
                float sum = 0.0f;
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        float u1 = ((float)a + 0.5f) / (float)n, u2 = ((float)b + 0.5f) / (float)n;
                        V3 h = g.sample_h(wo, u1, u2);
                        V3 wi = reflect(wo, h);
                        if (wi.z > 0.0f) sum += g.g2(wo, wi) / g.g1(wo);
                    }
                t[j * EN + i] = sum / (float)(n * n);
            }
        }
        return t;
    }();
    return e.data();
}
float ggx_e(float mu, float alpha) {
    const float* e = energy_table();
    float x = std::fmin(std::clamp(mu, 0.0f, 1.0f) * (float)(EN - 1), (float)(EN - 1) - 1e-4f);
    float y = std::fmin(std::clamp(std::sqrt(alpha), 0.0f, 1.0f) * (float)(EN - 1), (float)(EN - 1) - 1e-4f);
    size_t xi = (size_t)x, yi = (size_t)y;
    float fx = x - (float)xi, fy = y - (float)yi;
    auto g = [&](size_t i, size_t j) { return e[j * EN + i]; };
    float a = g(xi, yi) * (1.0f - fx) + g(xi + 1, yi) * fx;
    float b = g(xi, yi + 1) * (1.0f - fx) + g(xi + 1, yi + 1) * fx;
    return std::clamp(a * (1.0f - fy) + b * fy, 0.05f, 1.0f);
}

// ------------------------------------------------------------------ bsdf.rs
enum Kind { CONDUCTOR, COATED, DIFFUSE, DIELECTRIC, BLACK, OPAL };
struct BsdfSample {
    V3 wi;
    Spec f;
    float pdf;
    bool delta, transmit, terminate_secondary;
};
struct Bsdf {
    Kind kind;
    Ggx ggx;
    float eta1;
    Spec a, b;  // conductor eta/k; coated base; diffuse albedo; dielectric eta; opal body / lambda
    V3 on;
    float spacing, strength, w;

    bool is_delta() const { return kind == DIELECTRIC || (kind == CONDUCTOR && ggx.is_smooth()); }

    static float coat_spec_prob(float cos_o, float eta, const Spec& base) {
        float f = fresnel_dielectric(cos_o, eta);
        float d = base.average() * (1.0f - f);
        return std::clamp(f / std::fmax(f + d, 1e-6f), 0.15f, 0.95f);
    }
    static Bsdf coated(Ggx g, float eta, const Spec& base) {
        Bsdf r{};
        r.kind = COATED;
        r.ggx = g;
        r.eta1 = eta;
        r.a = base;
        return r;
    }

    Spec eval(V3 wo, V3 wi) const {
        switch (kind) {
            case CONDUCTOR: {
                if (wo.z <= 0.0f || wi.z <= 0.0f || ggx.is_smooth()) return Spec::splat(0.0f);
                V3 h = (wo + wi).normalize();
                Spec f = fresnel_conductor(wo.dot(h), a, b);
                float base = ggx.d(h) * ggx.g2(wo, wi) / (4.0f * wo.z * wi.z);
                float e = ggx_e(wo.z, ggx.alpha);
                Spec f0 = fresnel_conductor(1.0f, a, b);
                return f * base * (Spec::splat(1.0f) + f0 * ((1.0f - e) / e));
            }
            case COATED: {
                if (wo.z <= 0.0f || wi.z <= 0.0f) return Spec::splat(0.0f);
                V3 h = (wo + wi).normalize();
                float fs = fresnel_dielectric(wo.dot(h), eta1);
                float spec = ggx.d(h) * ggx.g2(wo, wi) * fs / (4.0f * wo.z * wi.z);
                float fo = fresnel_dielectric(wo.z, eta1);
                float fi = fresnel_dielectric(wi.z, eta1);
                return a * ((1.0f - fo) * (1.0f - fi) * FRAC_1_PI) + Spec::splat(spec);
            }
            case DIFFUSE:
                if (wo.z <= 0.0f || wi.z <= 0.0f) return Spec::splat(0.0f);
                return a * FRAC_1_PI;
            case OPAL: {
                if (wo.z <= 0.0f || wi.z <= 0.0f) return Spec::splat(0.0f);
                Spec base = coated(Ggx::from_roughness(0.12f), 1.45f, a).eval(wo, wi);
                return base * (1.0f - w) + flake_eval(wo, wi) * w;
            }
            default: return Spec::splat(0.0f);
        }
    }

    float pdf(V3 wo, V3 wi) const {
        switch (kind) {
            case CONDUCTOR: {
                if (wo.z <= 0.0f || wi.z <= 0.0f || ggx.is_smooth()) return 0.0f;
                V3 h = (wo + wi).normalize();
                return ggx.d_visible(wo, h) / (4.0f * std::fmax(wo.dot(h), 1e-8f));
            }
            case COATED: {
                if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
                float ps = coat_spec_prob(wo.z, eta1, a);
                V3 h = (wo + wi).normalize();
                float p_spec = ggx.d_visible(wo, h) / (4.0f * std::fmax(wo.dot(h), 1e-8f));
                return ps * p_spec + (1.0f - ps) * wi.z * FRAC_1_PI;
            }
            case DIFFUSE:
                if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
                return wi.z * FRAC_1_PI;
            case OPAL: {
                if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
                float pb = coated(Ggx::from_roughness(0.12f), 1.45f, a).pdf(wo, wi);
                return pb * (1.0f - w) + flake_pdf(wo, wi) * w;
            }
            default: return 0.0f;
        }
    }

    bool sample(V3 wo, float u0, float u1, float u2, BsdfSample* s) const {
        switch (kind) {
            case CONDUCTOR: {
                if (wo.z <= 0.0f) return false;
                if (ggx.is_smooth()) {
                    V3 wi(-wo.x, -wo.y, wo.z);
                    *s = BsdfSample{wi, fresnel_conductor(wo.z, a, b) / std::fmax(wi.z, 1e-8f), 1.0f, true, false, false};
                    return true;
                }
                V3 h = ggx.sample_h(wo, u1, u2);
                V3 wi = reflect(wo, h);
                if (wi.z <= 0.0f) return false;
                float p = ggx.d_visible(wo, h) / (4.0f * std::fmax(wo.dot(h), 1e-8f));
                *s = BsdfSample{wi, eval(wo, wi), p, false, false, false};
                return true;
            }
            case COATED: {
                if (wo.z <= 0.0f) return false;
                float ps = coat_spec_prob(wo.z, eta1, a);
                V3 wi = u0 < ps ? reflect(wo, ggx.sample_h(wo, u1, u2)) : cosine_hemisphere(u1, u2);
                if (wi.z <= 0.0f) return false;
                float p = pdf(wo, wi);
                if (p <= 0.0f) return false;
                *s = BsdfSample{wi, eval(wo, wi), p, false, false, false};
                return true;
            }
            case DIFFUSE: {
                if (wo.z <= 0.0f) return false;
                V3 wi = cosine_hemisphere(u1, u2);
                float p = wi.z * FRAC_1_PI;
                if (p <= 0.0f) return false;
                *s = BsdfSample{wi, a * FRAC_1_PI, p, false, false, false};
                return true;
            }
            case DIELECTRIC: {
                bool dispersive = (a.max_value() - a.min_value()) > 1e-5f;
                float e0 = a.v[0];
                float cos_o = wo.z;
                float f0 = fresnel_dielectric(cos_o, e0);
                if (u0 < f0) {
                    V3 wi(-wo.x, -wo.y, wo.z);
                    float c = std::fmax(std::fabs(wi.z), 1e-8f);
                    Spec f;
                    for (int i = 0; i < N; i++) f.v[i] = fresnel_dielectric(cos_o, a.v[i]);
                    *s = BsdfSample{wi, f / c, f0, true, false, false};
                    return true;
                }
                bool entering = cos_o > 0.0f;
                float etap = entering ? e0 : 1.0f / e0;
                V3 n = entering ? V3(0.0f, 0.0f, 1.0f) : V3(-0.0f, -0.0f, -1.0f);
                V3 wi;
                if (!refract(wo, n, etap, &wi)) return false;
                float c = std::fmax(std::fabs(wi.z), 1e-8f);
                float ft = (1.0f - f0) / (c * etap * etap);
                *s = BsdfSample{wi, Spec::splat(ft), 1.0f - f0, true, true, dispersive};
                return true;
            }
            case OPAL: {
                if (wo.z <= 0.0f) return false;
                V3 wi;
                if (u0 < w) {
                    Frame fr(on);
                    V3 wol = fr.to_local(wo);
                    if (wol.z <= 0.0f) return false;
                    V3 h = ggx.sample_h(wol, u1, u2);
                    wi = fr.to_world(reflect(wol, h));
                } else {
                    BsdfSample bs;
                    if (!coated(Ggx::from_roughness(0.12f), 1.45f, a).sample(wo, (u0 - w) / std::fmax(1.0f - w, 1e-6f), u1, u2, &bs))
                        return false;
                    wi = bs.wi;
                }
                if (wi.z <= 0.0f) return false;
                float p = pdf(wo, wi);
                if (p <= 0.0f) return false;
                *s = BsdfSample{wi, eval(wo, wi), p, false, false, false};
                return true;
            }
            default: return false;
        }
    }

    // bsdf.rs:257-292
    static float bragg(float spacing_nm, float cos_t, float l) {
        float peak = 2.0f * spacing_nm * 1.37f * std::clamp(cos_t, 0.0f, 1.0f);
        float sigma = 0.045f * std::fmax(peak, 1.0f);
        float x = (l - peak) / sigma;
        return std::exp(-0.5f * x * x);
    }
    Spec flake_eval(V3 wo, V3 wi) const {
        Frame fr(on);
        V3 aa = fr.to_local(wo), bb = fr.to_local(wi);
        if (aa.z <= 0.0f || bb.z <= 0.0f) return Spec::splat(0.0f);
        V3 h = (aa + bb).normalize();
        float g = ggx.d(h) * ggx.g2(aa, bb) / (4.0f * aa.z * bb.z);
        float cos_t = aa.dot(h);
        Spec r;
        for (int i = 0; i < N; i++) r.v[i] = strength * bragg(spacing, cos_t, b.v[i]);
        return r * (g * bb.z / std::fmax(wi.z, 1e-6f));
    }
    float flake_pdf(V3 wo, V3 wi) const {
        Frame fr(on);
        V3 aa = fr.to_local(wo), bb = fr.to_local(wi);
        if (aa.z <= 0.0f || bb.z <= 0.0f) return 0.0f;
        V3 h = (aa + bb).normalize();
        return ggx.d_visible(aa, h) / (4.0f * std::fmax(aa.dot(h), 1e-8f));
    }
};

// ------------------------------------------------------------------ material.rs
uint32_t hash3(int32_t i, int32_t j, int32_t k, uint32_t salt) {
    uint32_t h = ((uint32_t)i * 0x8da6b343u) ^ ((uint32_t)j * 0xd8163841u) ^ ((uint32_t)k * 0xcb1ab31fu) ^ salt;
    return hash32(h);
}
float unit(uint32_t h) { return (float)(h >> 8) / (float)(1u << 24); }
struct OpalDomain {
    V3 dir;
    float spacing, brightness;
};
OpalDomain opal_domain(V3 p) {
    int32_t fx = (int32_t)std::floor(p.x), fy = (int32_t)std::floor(p.y), fz = (int32_t)std::floor(p.z);
    float best = std::numeric_limits<float>::max();
    int32_t bi = 0, bj = 0, bk = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int32_t i = fx + dx, j = fy + dy, k = fz + dz;
                V3 c((float)i + unit(hash3(i, j, k, 1)), (float)j + unit(hash3(i, j, k, 2)), (float)k + unit(hash3(i, j, k, 3)));
                float d2 = (c - p).length_squared();
                if (d2 < best) {
                    best = d2;
                    bi = i;
                    bj = j;
                    bk = k;
                }
            }
    float z = 2.0f * unit(hash3(bi, bj, bk, 4)) - 1.0f;
    float phi = 6.2831855f * unit(hash3(bi, bj, bk, 5));
    float r = std::sqrt(std::fmax(1.0f - z * z, 0.0f));
    return OpalDomain{V3(r * std::cos(phi), r * std::sin(phi), z), 190.0f + 75.0f * unit(hash3(bi, bj, bk, 6)),
                      0.35f + 0.65f * std::pow(unit(hash3(bi, bj, bk, 7)), 0.7f)};
}

struct HitCtx {
    V3 p_obj, n_world;
    const Frame* frame;
    float turn_deg;
};

Bsdf material_bsdf(const ag_material& M, const Wavelengths& wl, float mask, float u, const HitCtx& ctx) {
    Bsdf b{};
    switch (M.kind) {
        case AG_MAT_OPAL: {
            OpalDomain d = opal_domain(ctx.p_obj / M.opal_cell_mm);
            V3 nw = (ctx.n_world * 1.2f + rot_y(d.dir, ctx.turn_deg)).normalize();
            V3 n = ctx.frame->to_local(nw);
            if (n.z < 0.0f) n = -n;
            b.kind = OPAL;
            b.on = n;
            b.ggx = Ggx::from_roughness(0.22f);
            b.spacing = d.spacing;
            b.strength = M.opal_strength * d.brightness;
            for (int i = 0; i < N; i++) b.b.v[i] = wl.lambda[i];
            b.a = tab_sample(M.spec_a, wl);
            b.w = 0.6f;
            return b;
        }
        case AG_MAT_METAL: {
            if (M.has_patina) {
                float t = std::clamp((mask - M.patina_lo) / std::fmax(M.patina_hi - M.patina_lo, 1e-4f), 0.0f, 1.0f);
                float w = M.patina_strength * t * t * (3.0f - 2.0f * t);
                if (u < w) return Bsdf::coated(Ggx{M.patina_alpha}, M.patina_eta, tab_sample(M.patina_base, wl));
            }
            b.kind = CONDUCTOR;
            b.ggx = Ggx{M.alpha};
            for (int i = 0; i < N; i++) {
                b.a.v[i] = piecewise(M.metal_n, wl.lambda[i]);
                b.b.v[i] = piecewise(M.metal_k, wl.lambda[i]);
            }
            return b;
        }
        case AG_MAT_GEM:
            b.kind = DIELECTRIC;
            b.a = tab_sample(M.spec_a, wl);
            return b;
        case AG_MAT_GEM_OPAQUE:
        case AG_MAT_PLASTIC: return Bsdf::coated(Ggx{M.alpha}, M.eta, tab_sample(M.spec_a, wl));
        case AG_MAT_DIFFUSE:
            b.kind = DIFFUSE;
            b.a = tab_sample(M.spec_a, wl);
            return b;
        default: b.kind = BLACK; return b;
    }
}

void guide_rgb(const ag_material& M, float cavity, float out[3]) {
    const float L[3] = {610.0f, 550.0f, 465.0f};
    switch (M.kind) {
        case AG_MAT_METAL:
            for (int i = 0; i < 3; i++) {
                float n = piecewise(M.metal_n, L[i]), k = piecewise(M.metal_k, L[i]);
                out[i] = ((n - 1.0f) * (n - 1.0f) + k * k) / ((n + 1.0f) * (n + 1.0f) + k * k);
            }
            if (M.has_patina) {
                float t = std::clamp((cavity - M.patina_lo) / std::fmax(M.patina_hi - M.patina_lo, 1e-4f), 0.0f, 1.0f);
                float w = M.patina_strength * t * t * (3.0f - 2.0f * t);
                for (int i = 0; i < 3; i++) out[i] = out[i] * (1.0f - w) + tab_eval(M.patina_base, L[i]) * w;
            }
            return;
        case AG_MAT_GEM:
        case AG_MAT_GEM_OPAQUE: out[0] = out[1] = out[2] = 1.0f; return;
        case AG_MAT_OPAL:
        case AG_MAT_PLASTIC:
        case AG_MAT_DIFFUSE:
            for (int i = 0; i < 3; i++) out[i] = tab_eval(M.spec_a, L[i]);
            return;
        default: out[0] = out[1] = out[2] = 0.0f; return;
    }
}

// ------------------------------------------------------------------ bvh.rs (restated)
struct Aabb {
    V3 min, max;
    static Aabb empty() { return Aabb{V3(INF, INF, INF), V3(-INF, -INF, -INF)}; }
    void grow(V3 p) {
        min = V3(std::fmin(min.x, p.x), std::fmin(min.y, p.y), std::fmin(min.z, p.z));
        max = V3(std::fmax(max.x, p.x), std::fmax(max.y, p.y), std::fmax(max.z, p.z));
    }
    Aabb uni(const Aabb& o) const {
        return Aabb{V3(std::fmin(min.x, o.min.x), std::fmin(min.y, o.min.y), std::fmin(min.z, o.min.z)),
                    V3(std::fmax(max.x, o.max.x), std::fmax(max.y, o.max.y), std::fmax(max.z, o.max.z))};
    }
    float area() const {
        V3 e = max - min;
        e = V3(std::fmax(e.x, 0.0f), std::fmax(e.y, 0.0f), std::fmax(e.z, 0.0f));
        return 2.0f * (e.x * e.y + e.y * e.z + e.z * e.x);
    }
};
struct Node {
    float min[3];
    uint32_t index;
    float max[3];
    uint32_t count;
};
struct Tri {
    V3 v0, e1, e2;
};
struct Node4 {
    float mn[3][4], mx[3][4];
    uint32_t child[4], count[4];
};
const uint32_t LEAF = 1u << 31;
const uint32_t EMPTY = 0xFFFFFFFFu;
const int BINS = 16;
const size_t MAX_LEAF = 4;
struct BuildPrim {
    Aabb bounds;
    V3 centroid;
    uint32_t id;
};
struct Sub {
    std::vector<Node> nodes;
    std::vector<uint32_t> prims;
};

Sub build_rec(BuildPrim* prims, size_t n);

// bvh.rs:412-417
Sub leaf(BuildPrim* prims, size_t n, const Aabb& b) {
    Sub s;
    s.nodes.push_back(Node{{b.min.x, b.min.y, b.min.z}, 0, {b.max.x, b.max.y, b.max.z}, (uint32_t)n});
    for (size_t i = 0; i < n; i++) s.prims.push_back(prims[i].id);
    return s;
}

// bvh.rs:497-538
Sub combine(const Aabb& bounds, BuildPrim* prims, size_t n, size_t mid) {
    Sub ls, rs;
    if (n > 8192) {
        auto f = std::async(std::launch::async, [&] { return build_rec(prims, mid); });
        rs = build_rec(prims + mid, n - mid);
        ls = f.get();
    } else {
        ls = build_rec(prims, mid);
        rs = build_rec(prims + mid, n - mid);
    }
    Sub out;
    out.nodes.reserve(1 + ls.nodes.size() + rs.nodes.size());
    out.nodes.push_back(Node{{bounds.min.x, bounds.min.y, bounds.min.z}, 1, {bounds.max.x, bounds.max.y, bounds.max.z}, 0});
    uint32_t l_rest = 3, r_rest = l_rest + (uint32_t)(ls.nodes.size() - 1);
    auto remap = [](Node nd, uint32_t base_rest, uint32_t prim_off) {
        if (nd.count > 0)
            nd.index += prim_off;
        else
            nd.index = base_rest + (nd.index - 1);
        return nd;
    };
    uint32_t lp = 0, rp = (uint32_t)ls.prims.size();
    out.nodes.push_back(remap(ls.nodes[0], l_rest, lp));
    out.nodes.push_back(remap(rs.nodes[0], r_rest, rp));
    for (size_t i = 1; i < ls.nodes.size(); i++) out.nodes.push_back(remap(ls.nodes[i], l_rest, lp));
    for (size_t i = 1; i < rs.nodes.size(); i++) out.nodes.push_back(remap(rs.nodes[i], r_rest, rp));
    out.prims = std::move(ls.prims);
    out.prims.insert(out.prims.end(), rs.prims.begin(), rs.prims.end());
    return out;
}

// bvh.rs:419-495
Sub build_rec(BuildPrim* prims, size_t n) {
    Aabb bounds = Aabb::empty(), cb = Aabb::empty();
    for (size_t i = 0; i < n; i++) {
        bounds = bounds.uni(prims[i].bounds);
        cb.grow(prims[i].centroid);
    }
    if (n <= MAX_LEAF) return leaf(prims, n, bounds);
    V3 ext = cb.max - cb.min;
    int axis = (ext.x >= ext.y && ext.x >= ext.z) ? 0 : (ext.y >= ext.z ? 1 : 2);
    auto comp = [axis](V3 v) { return axis == 0 ? v.x : (axis == 1 ? v.y : v.z); };
    float lo = comp(cb.min);
    float span = comp(ext);
    if (span <= 1e-12f) return combine(bounds, prims, n, n / 2);
    Aabb bin_b[BINS];
    size_t bin_n[BINS] = {0};
    for (auto& b : bin_b) b = Aabb::empty();
    float k = (float)BINS / span;
    auto bin_of = [&](float c) { return std::min((size_t)((c - lo) * k), (size_t)(BINS - 1)); };
    for (size_t i = 0; i < n; i++) {
        size_t b = bin_of(comp(prims[i].centroid));
        bin_n[b]++;
        bin_b[b] = bin_b[b].uni(prims[i].bounds);
    }
    float left_area[BINS - 1];
    size_t left_n[BINS - 1];
    Aabb acc = Aabb::empty();
    size_t cnt = 0;
    for (int i = 0; i < BINS - 1; i++) {
        acc = acc.uni(bin_b[i]);
        cnt += bin_n[i];
        left_area[i] = cnt > 0 ? acc.area() : 0.0f;
        left_n[i] = cnt;
    }
    float best_cost = INF;
    size_t best_bin = 0;
    acc = Aabb::empty();
    cnt = 0;
    for (int i = BINS - 1; i >= 1; i--) {
        acc = acc.uni(bin_b[i]);
        cnt += bin_n[i];
        size_t ln = left_n[i - 1];
        if (ln == 0 || cnt == 0) continue;
        float cost = left_area[i - 1] * (float)ln + acc.area() * (float)cnt;
        if (cost < best_cost) {
            best_cost = cost;
            best_bin = (size_t)i;
        }
    }
    float leaf_cost = bounds.area() * (float)n;
    if (!std::isfinite(best_cost) || (n <= 16 && best_cost >= leaf_cost)) {
        if (n <= 16) return leaf(prims, n, bounds);
        return combine(bounds, prims, n, n / 2);
    }
    size_t i = 0, j = n;
    while (i < j) {
        if (bin_of(comp(prims[i].centroid)) < best_bin)
            i++;
        else {
            j--;
            std::swap(prims[i], prims[j]);
        }
    }
    size_t mid = (i == 0 || i == n) ? n / 2 : i;
    return combine(bounds, prims, n, mid);
}

// bvh.rs:303-378
uint32_t collapse_build(const std::vector<Node>& nodes, uint32_t root, std::vector<Node4>& wide) {
    uint32_t me = (uint32_t)wide.size();
    Node4 blank;
    for (int a = 0; a < 3; a++)
        for (int c = 0; c < 4; c++) {
            blank.mn[a][c] = INF;
            blank.mx[a][c] = -INF;
        }
    for (int c = 0; c < 4; c++) {
        blank.child[c] = EMPTY;
        blank.count[c] = 0;
    }
    wide.push_back(blank);
    auto area = [](const Node& n) {
        float e[3] = {n.max[0] - n.min[0], n.max[1] - n.min[1], n.max[2] - n.min[2]};
        return e[0] * e[1] + e[1] * e[2] + e[2] * e[0];
    };
    const Node& r = nodes[root];
    std::vector<uint32_t> kids;
    if (r.count > 0)
        kids.push_back(root);
    else {
        kids.push_back(r.index);
        kids.push_back(r.index + 1);
    }
    while (kids.size() < 4) {
        int best = -1;
        float ba = 0.0f;
        for (size_t i = 0; i < kids.size(); i++) {
            const Node& n = nodes[kids[i]];
            if (n.count == 0) {
                float a = area(n);
                if (best < 0 || a > ba) {
                    best = (int)i;
                    ba = a;
                }
            }
        }
        if (best < 0) break;
        Node n = nodes[kids[best]];
        kids[best] = kids.back();
        kids.pop_back();
        kids.push_back(n.index);
        kids.push_back(n.index + 1);
    }
    Node4 node = wide[me];
    for (size_t c = 0; c < kids.size(); c++) {
        const Node& n = nodes[kids[c]];
        for (int a = 0; a < 3; a++) {
            node.mn[a][c] = n.min[a];
            node.mx[a][c] = n.max[a];
        }
        if (n.count > 0) {
            node.child[c] = LEAF | n.index;
            node.count[c] = n.count;
        } else {
            node.child[c] = EMPTY - 1;
        }
    }
    wide[me] = node;
    for (size_t c = 0; c < kids.size(); c++)
        if (nodes[kids[c]].count == 0) {
            uint32_t w = collapse_build(nodes, kids[c], wide);
            wide[me].child[c] = w;
        }
    return me;
}

struct TriHit {
    float t, u, v;
    uint32_t prim;
};

// bvh.rs:391-409
bool tri_intersect(const Tri& tri, const Ray& ray, float* t, float* u, float* v) {
    V3 p = ray.d.cross(tri.e2);
    float det = tri.e1.dot(p);
    if (std::fabs(det) < 1e-14f) return false;
    float inv_det = 1.0f / det;
    V3 s = ray.o - tri.v0;
    float uu = s.dot(p) * inv_det;
    if (!(uu >= 0.0f && uu <= 1.0f)) return false;
    V3 q = s.cross(tri.e1);
    float vv = ray.d.dot(q) * inv_det;
    if (vv < 0.0f || uu + vv > 1.0f) return false;
    *t = tri.e2.dot(q) * inv_det;
    *u = uu;
    *v = vv;
    return true;
}

struct Bvh {
    std::vector<Node> nodes;
    std::vector<Node4> wide;
    std::vector<Tri> tris;
    std::vector<uint32_t> prim_ids;

    // bvh.rs:117-146
    void build(const float* pos, const uint32_t* tri, uint32_t nt) {
        std::vector<BuildPrim> prims(nt);
        for (uint32_t i = 0; i < nt; i++) {
            Aabb b = Aabb::empty();
            for (int k = 0; k < 3; k++) b.grow(ld(pos + 3 * (size_t)tri[3 * (size_t)i + k]));
            prims[i] = BuildPrim{b, (b.min + b.max) * 0.5f, i};
        }
        Sub sub;
        if (prims.empty())
            sub.nodes.push_back(Node{{0, 0, 0}, 0, {0, 0, 0}, 0});
        else
            sub = build_rec(prims.data(), prims.size());
        tris.resize(sub.prims.size());
        for (size_t i = 0; i < sub.prims.size(); i++) {
            const uint32_t* t = tri + 3 * (size_t)sub.prims[i];
            V3 v0 = ld(pos + 3 * (size_t)t[0]);
            tris[i] = Tri{v0, ld(pos + 3 * (size_t)t[1]) - v0, ld(pos + 3 * (size_t)t[2]) - v0};
        }
        nodes = std::move(sub.nodes);
        prim_ids = std::move(sub.prims);
        wide.clear();
        if (!nodes.empty()) collapse_build(nodes, 0, wide);
    }

    // bvh.rs:280-299, with SSE minps/maxps semantics; robust mode widens the
    // interval by 2^-16 relative on both ends.  f32 Moller-Trumbore can report
    // a hit a little outside its triangle for grazing rays, so the margin is
    // wider than the box test's own rounding needs: it keeps such hits
    // visible to the traversal whatever the tree.
    static uint32_t slab4(const Node4& n, const Ray& r, V3 inv, float t_max, bool robust, float tmin_out[4]) {
        auto mn = [](float a, float b) { return a < b ? a : b; };
        auto mx = [](float a, float b) { return a > b ? a : b; };
        uint32_t mask = 0;
        for (int c = 0; c < 4; c++) {
            float tx0 = (n.mn[0][c] - r.o.x) * inv.x, tx1 = (n.mx[0][c] - r.o.x) * inv.x;
            float ty0 = (n.mn[1][c] - r.o.y) * inv.y, ty1 = (n.mx[1][c] - r.o.y) * inv.y;
            float tz0 = (n.mn[2][c] - r.o.z) * inv.z, tz1 = (n.mx[2][c] - r.o.z) * inv.z;
            float tmin = mx(mx(mx(mn(tx0, tx1), mn(ty0, ty1)), mn(tz0, tz1)), 0.0f);
            float tmax;
            if (robust) {
                tmax = mn(mn(mx(tx0, tx1), mx(ty0, ty1)), mx(tz0, tz1));
                tmin *= 1.0f - 1.52587890625e-5f;
                tmax = mn(tmax * (1.0f + 1.52587890625e-5f), t_max);
            } else {
                tmax = mn(mn(mn(mx(tx0, tx1), mx(ty0, ty1)), mx(tz0, tz1)), t_max);
            }
            tmin_out[c] = tmin;
            if (tmin <= tmax) mask |= 1u << c;
        }
        return mask;
    }

    // bvh.rs:159-213
    // counts (optional): [0] += Node4 visits, [1] += triangle tests
    bool intersect(const Ray& ray, float t_max, bool robust, TriHit* out, uint64_t* counts = nullptr) const {
        if (tris.empty()) return false;
        V3 inv(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
        float best_t = t_max;
        bool found = false;
        uint32_t bi = 0, bprim = EMPTY;
        float bu = 0, bv = 0;
        struct E {
            uint32_t entry, count;
            float tn;
        };
        E stack[192];
        int sp = 0;
        stack[sp++] = E{0, 0, 0.0f};
        while (sp > 0) {
            E e = stack[--sp];
            if (e.tn > best_t) continue;
            if (e.entry & LEAF) {
                if (e.entry == EMPTY) continue;
                uint32_t start = e.entry & ~LEAF;
                if (counts) counts[1] += e.count;
                for (uint32_t i = start; i < start + e.count; i++) {
                    float t, u, v;
                    if (!tri_intersect(tris[i], ray, &t, &u, &v)) continue;
                    if (!(t > 1e-7f)) continue;
                    uint32_t pid = prim_ids[i];
                    bool take = robust ? (t < best_t || (t == best_t && found && pid < bprim)) : (t < best_t);
                    if (take) {
                        best_t = t;
                        bi = i;
                        bu = u;
                        bv = v;
                        bprim = pid;
                        found = true;
                    }
                }
                continue;
            }
            const Node4& n = wide[e.entry];
            if (counts) counts[0]++;
            float tmin[4];
            uint32_t mask = slab4(n, ray, inv, best_t, robust, tmin);
            std::pair<int, float> order[4];
            int k = 0;
            for (int c = 0; c < 4; c++) {
                if (!(mask & (1u << c))) continue;
                float t = tmin[c];
                int j = k;
                while (j > 0 && order[j - 1].second < t) {
                    order[j] = order[j - 1];
                    j--;
                }
                order[j] = {c, t};
                k++;
            }
            for (int q = 0; q < k; q++) {
                if (sp >= 192) return found;  // cannot happen for the trees built here
                stack[sp++] = E{n.child[order[q].first], n.count[order[q].first], order[q].second};
            }
        }
        (void)bi;
        if (!found) return false;
        *out = TriHit{best_t, bu, bv, bprim};
        return true;
    }

    // bvh.rs:216-245
    bool occluded(const Ray& ray, float t_max, bool robust) const {
        if (tris.empty()) return false;
        V3 inv(1.0f / ray.d.x, 1.0f / ray.d.y, 1.0f / ray.d.z);
        struct E {
            uint32_t entry, count;
        };
        E stack[192];
        int sp = 1;
        stack[0] = E{0, 0};
        while (sp > 0) {
            E e = stack[--sp];
            if (e.entry & LEAF) {
                if (e.entry == EMPTY) continue;
                uint32_t start = e.entry & ~LEAF;
                for (uint32_t i = start; i < start + e.count; i++) {
                    float t, u, v;
                    if (tri_intersect(tris[i], ray, &t, &u, &v) && t > 1e-7f && t < t_max) return true;
                }
                continue;
            }
            const Node4& n = wide[e.entry];
            float tmin[4];
            uint32_t mask = slab4(n, ray, inv, t_max, robust, tmin);
            for (int c = 0; c < 4; c++)
                if (mask & (1u << c)) {
                    if (sp >= 192) return false;
                    stack[sp++] = E{n.child[c], n.count[c]};
                }
        }
        return false;
    }
};

// ------------------------------------------------------------------ light.rs
struct Panel {
    V3 center, hu, hv, n;
    float area;
    const float* emission;  // null = flag
    bool camera_visible;
    float feather;
    explicit Panel(const ag_panel& p) {
        center = ld(p.center);
        hu = ld(p.hu);
        hv = ld(p.hv);
        V3 c = hu.cross(hv);
        n = c.normalize();
        area = 4.0f * c.length();
        emission = p.has_emission ? p.emission : nullptr;
        camera_visible = p.camera_visible != 0;
        feather = p.feather;
    }
    bool intersect(const Ray& ray, float t_max, float* tout) const {
        float denom = ray.d.dot(n);
        if (std::fabs(denom) < 1e-9f) return false;
        float t = (center - ray.o).dot(n) / denom;
        if (t <= 1e-6f || t >= t_max) return false;
        V3 p = ray.at(t) - center;
        float a = p.dot(hu) / hu.length_squared();
        float b = p.dot(hv) / hv.length_squared();
        if (std::fabs(a) <= 1.0f && std::fabs(b) <= 1.0f) {
            *tout = t;
            return true;
        }
        return false;
    }
    Spec le(V3 p, V3 dir, const Wavelengths& wl) const {
        if (emission && dir.dot(n) < 0.0f) return tab_sample(emission, wl) * profile(p);
        return Spec::splat(0.0f);
    }
    float profile(V3 p) const {
        if (feather <= 0.0f) return 1.0f;
        V3 q = p - center;
        float a = std::fmin(std::fabs(q.dot(hu) / hu.length_squared()), 1.0f);
        float b = std::fmin(std::fabs(q.dot(hv) / hv.length_squared()), 1.0f);
        auto f = [&](float x) {
            float start = 1.0f - feather;
            if (x <= start) return 1.0f;
            float t = (x - start) / feather;
            return 1.0f - 0.75f * t * t * (3.0f - 2.0f * t);
        };
        return f(a) * f(b);
    }
    V3 sample_point(float u, float v) const { return center + hu * (2.0f * u - 1.0f) + hv * (2.0f * v - 1.0f); }
};

struct Floor {
    float y;
    int material;
    float fade_start, fade_end;
    bool intersect(const Ray& ray, float t_max, float* tout) const {
        if (std::fabs(ray.d.y) < 1e-9f) return false;
        float t = (y - ray.o.y) / ray.d.y;
        if (t > 1e-5f && t < t_max) {
            *tout = t;
            return true;
        }
        return false;
    }
    float presence(V3 p) const {
        float r = std::sqrt(p.x * p.x + p.z * p.z);
        if (r <= fade_start) return 1.0f;
        if (r >= fade_end) return 0.0f;
        float t = (r - fade_start) / (fade_end - fade_start);
        return 1.0f - t * t * (3.0f - 2.0f * t);
    }
};

// envmap.rs:107-239 (tables taken from the descriptor)
struct EnvMap {
    int w, h;
    const float *weights, *marginal, *conditional;
    float rotation_deg;
    const float* basis;  // 3 x 471
    void uv(V3 dir, float* u, float* v) const {
        V3 d = rot_y(dir, -rotation_deg);
        float uu = 0.5f + std::atan2(d.x, -d.z) / (2.0f * PI);
        *v = std::acos(std::clamp(d.y, -1.0f, 1.0f)) / PI;
        float r = std::fmod(uu, 1.0f);
        *u = r < 0.0f ? r + 1.0f : r;
    }
    V3 dir(float u, float v) const {
        float phi = (u - 0.5f) * 2.0f * PI;
        float theta = v * PI;
        float st = std::sin(theta), ct = std::cos(theta);
        return rot_y(V3(st * std::sin(phi), ct, -st * std::cos(phi)), rotation_deg);
    }
    size_t texel(float u, float v) const {
        uint32_t x = std::min((uint32_t)(u * (float)w), (uint32_t)w - 1);
        uint32_t y = std::min((uint32_t)(v * (float)h), (uint32_t)h - 1);
        return (size_t)y * w + x;
    }
    float lin_eval(const float* wt, float l) const {
        float x = std::clamp(l - 360.0f, 0.0f, 469.999f);
        size_t i = (size_t)x;
        float f = x - (float)i;
        float v = 0.0f;
        for (int c = 0; c < 3; c++) {
            const float* b = basis + c * AG_SPEC_N;
            v += wt[c] * (b[i] * (1.0f - f) + b[i + 1] * f);
        }
        return std::fmax(v, 0.0f);
    }
    Spec eval(V3 d, const Wavelengths& wl) const {
        float u, v;
        uv(d, &u, &v);
        const float* wt = weights + 3 * texel(u, v);
        Spec r;
        for (int i = 0; i < N; i++) r.v[i] = lin_eval(wt, wl.lambda[i]);
        return r;
    }
    float pdf(V3 d) const {
        float u, v;
        uv(d, &u, &v);
        float st = std::sin(PI * v);
        if (st <= 1e-6f) return 0.0f;
        size_t x = std::min((size_t)(u * (float)w), (size_t)w - 1);
        size_t y = std::min((size_t)(v * (float)h), (size_t)h - 1);
        size_t row = y * (w + 1);
        float pc = (conditional[row + x + 1] - conditional[row + x]) * (float)w;
        float pm = (marginal[y + 1] - marginal[y]) * (float)h;
        return pc * pm / (2.0f * PI * PI * st);
    }
    V3 sample(float u1, float u2, float* pdf_out) const {
        size_t y = std::clamp((size_t)(std::upper_bound(marginal, marginal + h + 1, u1) - marginal), (size_t)1, (size_t)h) - 1;
        const float* cond = conditional + y * (w + 1);
        size_t x = std::clamp((size_t)(std::upper_bound(cond, cond + w + 1, u2) - cond), (size_t)1, (size_t)w) - 1;
        float fy = (u1 - marginal[y]) / std::fmax(marginal[y + 1] - marginal[y], 1e-12f);
        float fx = (u2 - cond[x]) / std::fmax(cond[x + 1] - cond[x], 1e-12f);
        float u = ((float)x + std::clamp(fx, 0.0f, 0.999f)) / (float)w;
        float v = ((float)y + std::clamp(fy, 0.0f, 0.999f)) / (float)h;
        V3 d = dir(u, v);
        *pdf_out = pdf(d);
        return d;
    }
};

// envmap.rs:15-105: RGB -> spectrum as three smooth bases under a 6504 K
// blackbody, corrected by the inverse of their linear-sRGB responses.
// spectrum.rs:104-108, 125-136 (Blackbody::new, planck)
static float planck_nm(float lambda_nm, float t) {
    if (t <= 0.0f) return 0.0f;
    const double c = 299792458.0, h = 6.62606957e-34, kb = 1.3806488e-23;
    double l = (double)lambda_nm * 1e-9;
    double l2 = l * l;
    double l5 = l * (l2 * l2);
    double le = (2.0 * h * c * c) / (l5 * (std::exp((h * c) / (l * kb * (double)t)) - 1.0));
    return (float)le;
}
static float smoothstep_f(float e0, float e1, float x) {
    float t = std::clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
struct LinRgb {
    float basis[3][471];
    float inv[3][3];
    LinRgb() {
        const float kelvin = 6504.0f;
        float peak = 2.897772e-3f / kelvin * 1e9f;
        float norm = 1.0f / planck_nm(peak, kelvin);
        for (int i = 0; i < 471; i++) {
            float l = 360.0f + (float)i;
            float wb = 1.0f - smoothstep_f(470.0f, 515.0f, l);
            float wr = smoothstep_f(565.0f, 610.0f, l);
            float wg = std::fmax(1.0f - wb - wr, 0.0f);
            float e = planck_nm(l, kelvin) * norm;
            basis[0][i] = wr * e;
            basis[1][i] = wg * e;
            basis[2][i] = wb * e;
        }
        float m[3][3];
        for (int j = 0; j < 3; j++) {
            float xyz[3] = {0.0f, 0.0f, 0.0f};
            for (int i = 0; i < 471; i++) {
                float cb[3];
                xyz_bar(360.0f + (float)i, cb);
                for (int k = 0; k < 3; k++) xyz[k] += cb[k] * basis[j][i];
            }
            float x = xyz[0] / CIE_Y_INTEGRAL, y = xyz[1] / CIE_Y_INTEGRAL, z = xyz[2] / CIE_Y_INTEGRAL;
            // cie.rs:74-81
            float rgb[3] = {3.2404542f * x - 1.5371385f * y - 0.4985314f * z,
                            -0.969266f * x + 1.8760108f * y + 0.041556f * z,
                            0.0556434f * x - 0.2040259f * y + 1.0572252f * z};
            for (int i = 0; i < 3; i++) m[i][j] = rgb[i];
        }
        // envmap.rs:94-103
        float a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0],
              hh = m[2][1], ii = m[2][2];
        float det = a * (e * ii - f * hh) - b * (d * ii - f * g) + c * (d * hh - e * g);
        float k = 1.0f / det;
        float r[3][3] = {{(e * ii - f * hh) * k, (c * hh - b * ii) * k, (b * f - c * e) * k},
                         {(f * g - d * ii) * k, (a * ii - c * g) * k, (c * d - a * f) * k},
                         {(d * hh - e * g) * k, (b * g - a * hh) * k, (a * e - b * d) * k}};
        std::memcpy(inv, r, sizeof(inv));
    }
    void weights(const float rgb[3], float out[3]) const {
        for (int i = 0; i < 3; i++) out[i] = inv[i][0] * rgb[0] + inv[i][1] * rgb[1] + inv[i][2] * rgb[2];
    }
};
static const LinRgb& lin_rgb() {
    static const LinRgb L;
    return L;
}

// envmap.rs:141-178 (EnvData::load after decoding): basis weights per texel,
// the luminance x sin(theta) + 1e-6 sampling function, conditional CDFs per
// row (w+1 each) and the marginal CDF (h+1).
static void env_build(const float* rgb, int w, int h, float exposure_ev, float* weights, float* marginal,
                      float* conditional, float* total) {
    float k = std::pow(2.0f, exposure_ev);
    const LinRgb& lin = lin_rgb();
    std::vector<float> row_sum((size_t)h);
    for (int y = 0; y < h; y++) {
        float s = std::sin(PI * ((float)y + 0.5f) / (float)h);
        size_t base = (size_t)y * (size_t)(w + 1);
        float acc = 0.0f;
        conditional[base] = 0.0f;
        for (int x = 0; x < w; x++) {
            const float* p = rgb + ((size_t)y * w + x) * 3;
            float c[3] = {p[0] * k, p[1] * k, p[2] * k};
            lin.weights(c, weights + ((size_t)y * w + x) * 3);
            acc += std::fmax(0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2], 0.0f) * s + 1e-6f;
            conditional[base + x + 1] = acc;
        }
        row_sum[y] = acc;
        for (int x = 0; x <= w; x++) conditional[base + x] /= acc;
    }
    float acc = 0.0f;
    marginal[0] = 0.0f;
    for (int y = 0; y < h; y++) {
        acc += row_sum[y];
        marginal[y + 1] = acc;
    }
    *total = acc / ((float)w * (float)h);
    for (int y = 0; y <= h; y++) marginal[y] /= acc;
}

// ------------------------------------------------------------------ scene.rs
enum HitKind { H_PIECE, H_PANEL, H_FLOOR };
struct Hit {
    HitKind kind;
    float t;
    V3 p, p_obj, ng, ns;
    uint16_t material;
    float cavity;
    int index;
};

struct Scene {
    std::vector<float> positions, normals, cavity;
    std::vector<uint32_t> triangles;
    std::vector<uint16_t> tri_material;
    Bvh bvh;
    std::vector<ag_material> materials;
    std::vector<Panel> panels;
    bool env_map = false;
    const float *top, *horizon, *bottom;
    float power;
    EnvMap map;
    std::vector<float> env_weights, env_marg, env_cond, env_basis;
    bool has_floor = false;
    Floor floor;
    float turntable_deg = 0.0f;
    bool has_white = false;
    float white[3];
    std::vector<int> lights;
    std::vector<float> light_cdf;
    float env_pick = 0.0f;
    std::vector<float> env_tabs;  // top, horizon, bottom copies
    // Estimator strategy, a checking knob of the oracle only (the reference
    // always combines NEE and BSDF sampling by the power heuristic,
    // integrator.rs:70-98, :130-178): 0 = that MIS; 1 = light sampling
    // alone (emitters reached by BSDF sampling weigh 0, NEE weighs 1);
    // 2 = BSDF sampling alone (no NEE contribution, emitters hit weigh 1);
    // 3 = both unweighted, a deliberate double count that the closed-form
    // tests must reject.
    int strategy = 0;
    float bsdf_hit_weight(float w) const { return strategy == 0 ? w : strategy == 1 ? 0.0f : 1.0f; }
    float nee_weight(float w) const { return strategy == 0 ? w : strategy == 2 ? 0.0f : 1.0f; }

    // scene.rs:88-110
    void rebuild_lights() {
        lights.clear();
        light_cdf.clear();
        float acc = 0.0f;
        for (size_t i = 0; i < panels.size(); i++) {
            if (!panels[i].emission) continue;
            float power_ = panels[i].area * std::fmax(luminance_y_tab(panels[i].emission), 1e-6f);
            acc += power_;
            lights.push_back((int)i);
            light_cdf.push_back(acc);
        }
        if (acc > 0.0f)
            for (auto& c : light_cdf) c /= acc;
        env_pick = env_map ? (lights.empty() ? 1.0f : 0.5f) : 0.0f;
    }
    float light_pick_pdf(int panel) const {
        for (size_t k = 0; k < lights.size(); k++)
            if (lights[k] == panel) return light_cdf[k] - (k == 0 ? 0.0f : light_cdf[k - 1]);
        return 0.0f;
    }
    bool pick_light(float u, int* li, float* pick) const {
        if (lights.empty()) return false;
        size_t k = (size_t)(std::lower_bound(light_cdf.begin(), light_cdf.end(), u) - light_cdf.begin());
        k = std::min(k, lights.size() - 1);
        float lo = k == 0 ? 0.0f : light_cdf[k - 1];
        *li = lights[k];
        *pick = light_cdf[k] - lo;
        return true;
    }
    V3 to_object(V3 v) const { return turntable_deg == 0.0f ? v : rot_y(v, -turntable_deg); }
    V3 to_world(V3 v) const { return turntable_deg == 0.0f ? v : rot_y(v, turntable_deg); }

    Spec env_eval(V3 dir, const Wavelengths& wl) const {
        if (env_map) return map.eval(dir, wl);
        float y = std::clamp(dir.y, -1.0f, 1.0f);
        if (y >= 0.0f) {
            float t = std::pow(y, power);
            return Spec::lerp(t, tab_sample(horizon, wl), tab_sample(top, wl));
        }
        float t = std::pow(-y, 0.5f);
        return Spec::lerp(t, tab_sample(horizon, wl), tab_sample(bottom, wl));
    }

    // scene.rs:120-140
    float sky_visibility(V3 from, V3 dir, bool robust) const {
        Ray oray{to_object(from), to_object(dir)};
        if (bvh.occluded(oray, INF, robust)) return 0.0f;
        Ray ray{from, dir};
        float t;
        for (auto& p : panels)
            if (p.intersect(ray, INF, &t)) return 0.0f;
        if (has_floor && floor.intersect(ray, INF, &t)) return 1.0f - floor.presence(ray.at(t));
        return 1.0f;
    }

    // scene.rs:169-218
    bool intersect(const Ray& ray, bool camera_ray, bool robust, Hit* best) const {
        bool any = false;
        float t_max = INF;
        Ray oray{to_object(ray.o), to_object(ray.d)};
        TriHit h;
        if (bvh.intersect(oray, t_max, robust, &h)) {
            t_max = h.t;
            const uint32_t* tri = &triangles[3 * (size_t)h.prim];
            V3 a = ld(&positions[3 * (size_t)tri[0]]), b = ld(&positions[3 * (size_t)tri[1]]), c = ld(&positions[3 * (size_t)tri[2]]);
            V3 ng = (b - a).cross(c - a).normalize();
            float w = 1.0f - h.u - h.v;
            uint16_t mat = tri_material[h.prim];
            bool flat = materials[mat].kind == AG_MAT_GEM;
            V3 ns = ng;
            if (!flat && !normals.empty()) {
                V3 na = ld(&normals[3 * (size_t)tri[0]]), nb = ld(&normals[3 * (size_t)tri[1]]), nc = ld(&normals[3 * (size_t)tri[2]]);
                V3 n = na * w + nb * h.u + nc * h.v;
                ns = n.length_squared() > 1e-12f ? n.normalize() : ng;
            }
            float cav = cavity.empty() ? 0.0f : cavity[tri[0]] * w + cavity[tri[1]] * h.u + cavity[tri[2]] * h.v;
            V3 p_obj = a * w + b * h.u + c * h.v;
            *best = Hit{H_PIECE, h.t, to_world(p_obj), p_obj, to_world(ng), to_world(ns), mat, cav, 0};
            any = true;
        }
        for (size_t i = 0; i < panels.size(); i++) {
            if (camera_ray && !panels[i].camera_visible) continue;
            float t;
            if (panels[i].intersect(ray, t_max, &t)) {
                t_max = t;
                best->kind = H_PANEL;
                best->t = t;
                best->index = (int)i;
                any = true;
            }
        }
        float t;
        if (has_floor && floor.intersect(ray, t_max, &t)) {
            best->kind = H_FLOOR;
            best->t = t;
            best->p = ray.at(t);
            any = true;
        }
        return any;
    }

    // scene.rs:222-243
    bool visible(V3 from, V3 to, int ignore_panel, bool robust) const {
        V3 d = to - from;
        float dist = d.length();
        V3 dir = d / dist;
        float t_max = dist * (1.0f - 1e-4f);
        Ray oray{to_object(from), to_object(dir)};
        if (bvh.occluded(oray, t_max, robust)) return false;
        Ray ray{from, dir};
        float t;
        for (size_t i = 0; i < panels.size(); i++)
            if ((int)i != ignore_panel && panels[i].intersect(ray, t_max, &t)) return false;
        if (has_floor && floor.intersect(ray, t_max, &t)) return false;
        return true;
    }
};

// ------------------------------------------------------------------ integrator.rs
struct Guides {
    float albedo[3] = {0, 0, 0}, normal[3] = {0, 0, 0};
};
struct Counters {
    uint64_t segments = 0, shadows = 0;
};

void add_clamped(Spec& l, const Spec& c, bool clamp, float m) {
    if (clamp) {
        float mx = c.max_value();
        if (mx > m) {
            l = l + c * (m / mx);
            return;
        }
    }
    l = l + c;
}

// integrator.rs:50-221
Spec trace(const Scene& scene, Ray ray, Wavelengths& wl, Sampler& s, const ag_path_settings& ps, Guides& guides,
           Counters& cnt) {
    const bool robust = true;
    Spec l = Spec::splat(0.0f), beta = Spec::splat(1.0f);
    bool specular = true, rough_seen = false;
    float prev_pdf = 1.0f;
    int medium = -1;
    const V3 floor_n(0.0f, 1.0f, 0.0f);
    for (uint32_t depth = 0; depth < ps.max_depth; depth++) {
        bool clamp = rough_seen;
        s.set_dim(8 + depth * 8);
        Hit hit;
        cnt.segments++;
        bool got = scene.intersect(ray, depth == 0, robust, &hit);
        if (medium >= 0 && got) {
            Spec sig = tab_sample(scene.materials[medium].spec_b, wl);
            for (int i = 0; i < N; i++) beta.v[i] = beta.v[i] * std::exp(-sig.v[i] * hit.t);
        }
        auto env_weight = [&](V3 dir) {
            if (scene.env_map && !specular) {
                float pl = scene.env_pick * scene.map.pdf(dir);
                return scene.bsdf_hit_weight(prev_pdf * prev_pdf / (prev_pdf * prev_pdf + pl * pl));
            }
            return 1.0f;
        };
        V3 p, p_obj, ng, ns;
        int mat_id;
        float cavity;
        if (!got) {
            add_clamped(l, beta * scene.env_eval(ray.d, wl) * env_weight(ray.d), clamp, ps.clamp_indirect);
            break;
        }
        if (hit.kind == H_PANEL) {
            const Panel& panel = scene.panels[hit.index];
            Spec le = panel.le(ray.at(hit.t), ray.d, wl);
            if (!le.is_black()) {
                if (specular)
                    add_clamped(l, beta * le, clamp, ps.clamp_indirect);
                else {
                    float cos_l = std::fmax(std::fabs((-ray.d).dot(panel.n)), 1e-8f);
                    float p_light = (1.0f - scene.env_pick) * scene.light_pick_pdf(hit.index) * hit.t * hit.t / (cos_l * panel.area);
                    float w = scene.bsdf_hit_weight(prev_pdf * prev_pdf / (prev_pdf * prev_pdf + p_light * p_light));
                    add_clamped(l, beta * le * w, clamp, ps.clamp_indirect);
                }
            }
            break;
        } else if (hit.kind == H_FLOOR) {
            if (s.next_random() >= scene.floor.presence(hit.p)) {
                add_clamped(l, beta * scene.env_eval(ray.d, wl) * env_weight(ray.d), clamp, ps.clamp_indirect);
                break;
            }
            V3 n = ray.d.y < 0.0f ? floor_n : -floor_n;
            p = hit.p;
            p_obj = hit.p;
            ng = n;
            ns = n;
            mat_id = scene.floor.material;
            cavity = 0.0f;
        } else {
            p = hit.p;
            p_obj = hit.p_obj;
            ng = hit.ng;
            ns = hit.ns;
            mat_id = hit.material;
            cavity = hit.cavity;
        }
        const ag_material& mat = scene.materials[mat_id];
        V3 wo_w = -ray.d;
        bool transmissive = mat.kind == AG_MAT_GEM;
        V3 nsu = ns, ngu = ng;
        if (!transmissive && ng.dot(wo_w) < 0.0f) {
            nsu = -ns;
            ngu = -ng;
        }
        if (!transmissive && nsu.dot(wo_w) <= 0.0f) nsu = ngu;
        Frame frame(nsu);
        V3 wo = frame.to_local(wo_w);
        HitCtx ctx{p_obj, nsu, &frame, scene.turntable_deg};
        Bsdf bsdf = material_bsdf(mat, wl, cavity, s.next(), ctx);
        if (depth == 0) {
            guide_rgb(mat, cavity, guides.albedo);
            guides.normal[0] = nsu.x;
            guides.normal[1] = nsu.y;
            guides.normal[2] = nsu.z;
        }
        if (!bsdf.is_delta()) {
            float ul = s.next();
            float u1 = s.next(), u2 = s.next();
            if (scene.env_map && ul < scene.env_pick) {
                float pdf_dir;
                V3 dir = scene.map.sample(u1, u2, &pdf_dir);
                if (pdf_dir > 0.0f && dir.dot(ngu) > 0.0f) {
                    V3 wi = frame.to_local(dir);
                    Spec f = bsdf.eval(wo, wi);
                    if (!f.is_black()) {
                        float p_light = scene.env_pick * pdf_dir;
                        float p_bsdf = bsdf.pdf(wo, wi);
                        float w = scene.nee_weight(p_light * p_light / (p_light * p_light + p_bsdf * p_bsdf));
                        cnt.shadows++;
                        float vis = scene.sky_visibility(offset_origin(p, ngu, dir), dir, robust);
                        if (vis > 0.0f) {
                            Spec le = scene.map.eval(dir, wl);
                            Spec c = beta * f * le * (std::fabs(wi.z) * w * vis / p_light);
                            add_clamped(l, c, rough_seen, ps.clamp_indirect);
                        }
                    }
                }
            } else {
                int li;
                float pick;
                if (scene.pick_light(scene.env_pick > 0.0f ? (ul - scene.env_pick) / (1.0f - scene.env_pick) : ul, &li, &pick)) {
                    pick = pick * (1.0f - scene.env_pick);
                    const Panel& panel = scene.panels[li];
                    V3 lp = panel.sample_point(u1, u2);
                    V3 to = lp - p;
                    float d2 = to.length_squared();
                    V3 dir = to / std::sqrt(d2);
                    float cos_l = (-dir).dot(panel.n);
                    if (cos_l > 0.0f && dir.dot(ngu) > 0.0f) {
                        V3 wi = frame.to_local(dir);
                        Spec f = bsdf.eval(wo, wi);
                        if (!f.is_black()) {
                            float p_light = pick * d2 / (cos_l * panel.area);
                            float p_bsdf = bsdf.pdf(wo, wi);
                            float w = scene.nee_weight(p_light * p_light / (p_light * p_light + p_bsdf * p_bsdf));
                            V3 origin = offset_origin(p, ngu, dir);
                            cnt.shadows++;
                            if (scene.visible(origin, lp, li, robust)) {
                                Spec le = panel.le(lp, dir, wl);
                                Spec c = beta * f * le * (std::fabs(wi.z) * w / p_light);
                                add_clamped(l, c, rough_seen, ps.clamp_indirect);
                            }
                        }
                    }
                }
            }
        } else {
            s.next();
            s.next();
            s.next();
        }
        float u0 = s.next();
        float u1 = s.next(), u2 = s.next();
        BsdfSample smp;
        if (!bsdf.sample(wo, u0, u1, u2, &smp)) break;
        if (smp.pdf <= 0.0f || smp.f.is_black()) break;
        V3 wi_w = frame.to_world(smp.wi);
        float side = wi_w.dot(ngu) * wo_w.dot(ngu);
        if ((!smp.transmit && side <= 0.0f) || (smp.transmit && side >= 0.0f)) break;
        if (smp.terminate_secondary) wl.terminate_secondary();
        beta = beta * (smp.f * (std::fabs(smp.wi.z) / smp.pdf));
        if (!beta.is_finite()) break;
        specular = smp.delta;
        if (!smp.delta) rough_seen = true;
        prev_pdf = smp.pdf;
        if (smp.transmit) medium = (wo.z > 0.0f && mat.kind == AG_MAT_GEM) ? mat_id : -1;
        ray = Ray{offset_origin(p, ngu, wi_w), wi_w};
        if (depth >= ps.rr_depth) {
            float survive = std::fmin(beta.max_value(), 0.95f);
            if (survive <= 0.0f || s.next_random() >= survive) break;
            beta = beta / survive;
        }
    }
    return l;
}

// camera.rs:57-70
Ray camera_ray(const ag_camera& c, float sx, float sy, float lu, float lv) {
    float x = (sx + c.shift[0]) * c.half_w;
    float y = (sy + c.shift[1]) * c.half_h;
    V3 fwd = ld(c.forward), right = ld(c.right), up = ld(c.up), pos = ld(c.pos);
    V3 dir = (fwd + right * x + up * y).normalize();
    if (c.lens_radius <= 0.0f) return Ray{pos, dir};
    float ft = c.focus_dist / dir.dot(fwd);
    V3 p_focus = pos + dir * ft;
    float dx, dy;
    disk(lu, lv, &dx, &dy);
    V3 o = pos + right * (dx * c.lens_radius) + up * (dy * c.lens_radius);
    return Ray{o, (p_focus - o).normalize()};
}

}  // namespace orc

using namespace orc;

struct or_scene {
    Scene sc;
};

static void parallel_for(uint32_t n, int threads, const std::function<void(uint32_t, uint32_t)>& fn) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<uint32_t>((uint32_t)threads, std::max(1u, n));
    if (threads == 1) {
        fn(0, n);
        return;
    }
    std::atomic<uint32_t> next{0};
    const uint32_t chunk = std::max(1u, std::min(4096u, n / (uint32_t)(threads * 8) + 1));
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++)
        th.emplace_back([&] {
            for (;;) {
                uint32_t a = next.fetch_add(chunk);
                if (a >= n) break;
                fn(a, std::min(n, a + chunk));
            }
        });
    for (auto& t : th) t.join();
}

extern "C" or_scene* or_scene_create(const ag_scene_desc* d) {
    if (!d) return nullptr;
    auto* s = new or_scene();
    Scene& S = s->sc;
    size_t nv = d->n_vertices, nt = d->n_triangles;
    S.positions.assign(d->positions, d->positions + 3 * nv);
    if (d->normals) S.normals.assign(d->normals, d->normals + 3 * nv);
    if (d->cavity) S.cavity.assign(d->cavity, d->cavity + nv);
    S.triangles.assign(d->triangles, d->triangles + 3 * nt);
    S.tri_material.assign(nt, 0);
    if (d->tri_material) S.tri_material.assign(d->tri_material, d->tri_material + nt);
    S.bvh.build(S.positions.data(), S.triangles.data(), (uint32_t)nt);
    S.materials.assign(d->materials, d->materials + d->n_materials);
    for (int i = 0; i < d->n_panels; i++) S.panels.emplace_back(d->panels[i]);
    S.env_tabs.assign(3 * AG_SPEC_N, 0.0f);
    std::memcpy(&S.env_tabs[0], d->env.top, AG_SPEC_N * 4);
    std::memcpy(&S.env_tabs[AG_SPEC_N], d->env.horizon, AG_SPEC_N * 4);
    std::memcpy(&S.env_tabs[2 * AG_SPEC_N], d->env.bottom, AG_SPEC_N * 4);
    S.top = &S.env_tabs[0];
    S.horizon = &S.env_tabs[AG_SPEC_N];
    S.bottom = &S.env_tabs[2 * AG_SPEC_N];
    S.power = d->env.power;
    if (d->env.kind == AG_ENV_MAP) {
        S.env_map = true;
        size_t w = d->env.width, h = d->env.height;
        S.env_weights.assign(d->env.weights, d->env.weights + 3 * w * h);
        S.env_marg.assign(d->env.marginal, d->env.marginal + h + 1);
        S.env_cond.assign(d->env.conditional, d->env.conditional + h * (w + 1));
        S.env_basis.assign(&d->env.basis[0][0], &d->env.basis[0][0] + 3 * AG_SPEC_N);
        S.map = EnvMap{(int)w, (int)h, S.env_weights.data(), S.env_marg.data(), S.env_cond.data(), d->env.rotation_deg,
                       S.env_basis.data()};
    }
    if (d->floor.present) {
        S.has_floor = true;
        S.floor = Floor{d->floor.y, d->floor.material, d->floor.fade_start, d->floor.fade_end};
    }
    S.turntable_deg = d->turntable_deg;
    S.has_white = d->has_white != 0;
    std::memcpy(S.white, d->white_xyz, sizeof(S.white));
    S.rebuild_lights();
    return s;
}

extern "C" void or_scene_destroy(or_scene* s) { delete s; }
extern "C" void or_scene_set_turntable(or_scene* s, float deg) {
    if (s) s->sc.turntable_deg = deg;
}
extern "C" int or_scene_set_material(or_scene* s, int32_t m, const ag_material* mat) {
    if (!s || !mat || m < 0 || (size_t)m >= s->sc.materials.size()) return 1;
    s->sc.materials[m] = *mat;
    return 0;
}
extern "C" int or_scene_set_strategy(or_scene* s, int32_t strategy) {
    if (!s || strategy < 0 || strategy > 3) return 1;
    s->sc.strategy = strategy;
    return 0;
}
extern "C" int or_scene_stats(const or_scene* s, uint32_t* bn, uint32_t* wn) {
    if (!s) return 1;
    if (bn) *bn = (uint32_t)s->sc.bvh.nodes.size();
    if (wn) *wn = (uint32_t)s->sc.bvh.wide.size();
    return 0;
}

extern "C" int or_intersect(const or_scene* s, const float* ox, const float* oy, const float* oz, const float* dx,
                            const float* dy, const float* dz, const float* tmax, uint32_t n, int32_t mode,
                            int32_t threads, float* t, float* u, float* v, uint32_t* prim) {
    if (!s) return 1;
    parallel_for(n, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; i++) {
            Ray r{V3(ox[i], oy[i], oz[i]), V3(dx[i], dy[i], dz[i])};
            TriHit h;
            if (s->sc.bvh.intersect(r, tmax ? tmax[i] : INF, mode == OR_MODE_ROBUST, &h)) {
                t[i] = h.t;
                u[i] = h.u;
                v[i] = h.v;
                prim[i] = h.prim;
            } else {
                t[i] = INF;
                u[i] = 0.0f;
                v[i] = 0.0f;
                prim[i] = 0xFFFFFFFFu;
            }
        }
    });
    return 0;
}

// Work of the reference's closest-hit query (bvh.rs:159-213) on its own
// 4-wide tree: totals[0] = Node4 visits, totals[1] = triangle tests.
extern "C" int or_intersect_stats(const or_scene* s, const float* ox, const float* oy, const float* oz,
                                  const float* dx, const float* dy, const float* dz, const float* tmax, uint32_t n,
                                  int32_t mode, uint64_t* totals) {
    if (!s) return 1;
    uint64_t c[2] = {0, 0};
    for (uint32_t i = 0; i < n; i++) {
        Ray r{V3(ox[i], oy[i], oz[i]), V3(dx[i], dy[i], dz[i])};
        TriHit h;
        s->sc.bvh.intersect(r, tmax ? tmax[i] : INF, mode == OR_MODE_ROBUST, &h, c);
    }
    totals[0] = c[0];
    totals[1] = c[1];
    return 0;
}

extern "C" int or_occluded(const or_scene* s, const float* ox, const float* oy, const float* oz, const float* dx,
                           const float* dy, const float* dz, const float* tmax, uint32_t n, int32_t mode,
                           int32_t threads, uint8_t* out) {
    if (!s) return 1;
    parallel_for(n, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; i++) {
            Ray r{V3(ox[i], oy[i], oz[i]), V3(dx[i], dy[i], dz[i])};
            out[i] = s->sc.bvh.occluded(r, tmax ? tmax[i] : INF, mode == OR_MODE_ROBUST) ? 1 : 0;
        }
    });
    return 0;
}

// argent-mesh cavity.rs:95-128 (ambient_occlusion), the occluder being the
// piece's BVH as the renderer supplies it (argent-core scene.rs:40-44:
// Bvh::occluded, bvh.rs:216-245).  Vertex normals are the scene's own (the
// reference's vertex_normals, cavity.rs:80-86, returns mesh.normals when
// present); bounds as Mesh::bounds (argent-mesh lib.rs:122-134).
extern "C" int or_ambient_occlusion(const or_scene* s, uint32_t samples, float distance, int32_t mode,
                                    int32_t threads, float* out) {
    if (!s) return 1;
    const Scene& S = s->sc;
    const uint32_t nv = (uint32_t)(S.positions.size() / 3);
    if (nv == 0) return 0;
    if (S.triangles.empty()) {  // cavity.rs:54-56
        std::fill(out, out + nv, 0.0f);
        return 0;
    }
    // cavity.rs:82-87 (vertex_normals): the mesh's own, else smooth_vertex_normals
    std::vector<float> smooth;
    const float* N = S.normals.data();
    if (S.normals.size() != S.positions.size()) {
        smooth.resize(3 * (size_t)nv);
        or_smooth_vertex_normals(S.positions.data(), nv, S.triangles.data(), (uint32_t)(S.triangles.size() / 3),
                                 smooth.data());
        N = smooth.data();
    }
    float lo[3] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                   std::numeric_limits<float>::max()};
    float hi[3] = {std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
                   std::numeric_limits<float>::lowest()};
    for (uint32_t v = 0; v < nv; v++)
        for (int a = 0; a < 3; a++) {
            lo[a] = std::fmin(lo[a], S.positions[3 * (size_t)v + a]);
            hi[a] = std::fmax(hi[a], S.positions[3 * (size_t)v + a]);
        }
    // cavity.rs:87-90 (diag: powi(2) is x * x) and :97
    const float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
    const float diag = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    const float eps = std::fmax(2e-4f * diag, 1e-3f);
    const uint32_t n = std::max(samples, 1u);
    const bool robust = mode == OR_MODE_ROBUST;
    parallel_for(nv, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t v = a; v < b; v++) {
            V3 nrm = ld(&N[3 * (size_t)v]);
            // Vec3::normalize_or_zero: times the reciprocal length when finite and positive
            const float rcp = 1.0f / nrm.length();
            if (!(std::isfinite(rcp) && rcp > 0.0f)) {
                out[v] = 0.0f;
                continue;
            }
            const V3 nn = nrm * rcp;
            const Frame F(nn);  // Vec3::any_orthonormal_pair
            const V3 o = ld(&S.positions[3 * (size_t)v]) + nn * eps;
            const uint32_t h = hash32(v ^ 0x5bd1e995u);
            const float rot = (float)(h >> 8) * (1.0f / 16777216.0f);   // argent-mesh util.rs:255-257
            const float jit = (float)(hash32(h) >> 8) * (1.0f / 16777216.0f);
            uint32_t hits = 0;
            for (uint32_t k = 0; k < n; k++) {
                const float u1 = std::fmin(((float)k + jit) / (float)n, 0.999999f);
                const float x2 = (float)k * 0.618034f + rot;
                const float u2 = x2 - std::trunc(x2);  // f32::fract
                const float r = std::sqrt(u1), phi = 6.28318530717958647692f * u2;
                const V3 d = (F.t * (r * std::cos(phi)) + F.b * (r * std::sin(phi))) + nn * std::sqrt(1.0f - u1);
                if (S.bvh.occluded(Ray{o, d}, distance, robust)) hits++;
            }
            out[v] = (float)hits / (float)n;
        }
    });
    return 0;
}

extern "C" int or_brute_intersect(const float* pos, const uint32_t* tri, uint32_t nt, const float* ox, const float* oy,
                                  const float* oz, const float* dx, const float* dy, const float* dz, const float* tmax,
                                  uint32_t n, int32_t threads, float* t, float* u, float* v, uint32_t* prim) {
    std::vector<Tri> tris(nt);
    for (uint32_t k = 0; k < nt; k++) {
        V3 v0 = ld(pos + 3 * (size_t)tri[3 * (size_t)k]);
        tris[k] = Tri{v0, ld(pos + 3 * (size_t)tri[3 * (size_t)k + 1]) - v0, ld(pos + 3 * (size_t)tri[3 * (size_t)k + 2]) - v0};
    }
    parallel_for(n, threads, [&](uint32_t a, uint32_t b) {
        for (uint32_t i = a; i < b; i++) {
            Ray r{V3(ox[i], oy[i], oz[i]), V3(dx[i], dy[i], dz[i])};
            float best = tmax ? tmax[i] : INF, bu = 0, bv = 0;
            uint32_t bp = 0xFFFFFFFFu;
            for (uint32_t k = 0; k < nt; k++) {
                float tt, uu, vv;
                if (tri_intersect(tris[k], r, &tt, &uu, &vv) && tt > 1e-7f && tt < best) {
                    best = tt;
                    bu = uu;
                    bv = vv;
                    bp = k;
                }
            }
            t[i] = bp == 0xFFFFFFFFu ? INF : best;
            u[i] = bu;
            v[i] = bv;
            prim[i] = bp;
        }
    });
    return 0;
}

extern "C" int or_render_pass(const or_scene* s, const ag_camera* cam, uint32_t w, uint32_t h, float* sums,
                              uint32_t first, const ag_path_settings* ps, uint32_t count, uint32_t seed,
                              int32_t threads, uint32_t row0, uint32_t row1, uint64_t* stats) {
    if (!s || !cam || !sums || !ps) return 1;
    row1 = std::min(row1, h);
    if (row0 >= row1) return 0;
    const Scene& S = s->sc;
    const float aspect_sx = 2.0f / (float)w, aspect_sy = 2.0f / (float)h;
    const uint32_t seed_h = hash32(seed ^ 0xa511e9b3u);
    M3 cam_m;
    camera_matrix(S.has_white, S.white, cam_m);
    const size_t plane = (size_t)w * h * 3;
    std::atomic<uint64_t> segs{0}, shadows{0};
    // pixels, not rows, are dealt to the threads, so that a pass over a few
    // rows still runs on every thread (each pixel's sum is independent)
    parallel_for((row1 - row0) * w, threads, [&](uint32_t pa, uint32_t pb) {
        Counters cnt;
        for (uint32_t pi = pa; pi < pb; pi++) {
            {
                const uint32_t y = row0 + pi / w, x = pi % w;
                uint32_t pix_seed = hash32((y * w + x) ^ seed_h);
                float acc[3] = {0, 0, 0}, aacc[3] = {0, 0, 0}, nacc[3] = {0, 0, 0};
                for (uint32_t si = first; si < first + count; si++) {
                    Sampler smp(pix_seed, si);
                    float jx = tent(smp.next(), 1.0f), jy = tent(smp.next(), 1.0f);
                    float lu = smp.next(), lv = smp.next();
                    Wavelengths wl = Wavelengths::sample_visible(smp.next());
                    float sx = (((float)x + 0.5f + jx) * aspect_sx) - 1.0f;
                    float sy = 1.0f - (((float)y + 0.5f + jy) * aspect_sy);
                    Ray ray = camera_ray(*cam, sx, sy, lu, lv);
                    Guides g;
                    Spec l = trace(S, ray, wl, smp, *ps, g, cnt);
                    float xyz[3];
                    to_xyz(l, wl, xyz);
                    float rgb[3];
                    for (int i = 0; i < 3; i++) rgb[i] = cam_m[i][0] * xyz[0] + cam_m[i][1] * xyz[1] + cam_m[i][2] * xyz[2];
                    if (std::isfinite(rgb[0]) && std::isfinite(rgb[1]) && std::isfinite(rgb[2]))
                        for (int c = 0; c < 3; c++) acc[c] += rgb[c];
                    for (int c = 0; c < 3; c++) {
                        aacc[c] += g.albedo[c];
                        nacc[c] += g.normal[c];
                    }
                }
                size_t o = ((size_t)y * w + x) * 3;
                for (int c = 0; c < 3; c++) {
                    sums[o + c] += acc[c];
                    sums[plane + o + c] += aacc[c];
                    sums[2 * plane + o + c] += nacc[c];
                }
            }
        }
        segs += cnt.segments;
        shadows += cnt.shadows;
    });
    if (stats) {
        stats[0] = segs.load();
        stats[1] = shadows.load();
        stats[2] = (uint64_t)(row1 - row0) * w * count;
    }
    return 0;
}

extern "C" int or_energy_table(float* out) {
    std::memcpy(out, energy_table(), 1024 * sizeof(float));
    return 0;
}
extern "C" float or_fresnel_dielectric(float c, float eta) { return fresnel_dielectric(c, eta); }
extern "C" float or_fresnel_conductor1(float c, float eta, float k) { return fresnel_conductor1(c, eta, k); }
extern "C" int or_sample_wavelengths(float u, float* lam, float* pdf) {
    Wavelengths w = Wavelengths::sample_visible(u);
    std::memcpy(lam, w.lambda, 16);
    std::memcpy(pdf, w.pdf, 16);
    return 0;
}
extern "C" int or_camera_matrix(int32_t has_white, const float* white, float* m9) {
    M3 m;
    camera_matrix(has_white != 0, white, m);
    std::memcpy(m9, m, 36);
    return 0;
}
extern "C" float or_rng_dim(uint32_t ps, uint32_t si, uint32_t dim) { return rng_dim(ps, si, dim); }
extern "C" int or_camera_ray(const ag_camera* cam, float sx, float sy, float lu, float lv, float* o3, float* d3) {
    Ray r = camera_ray(*cam, sx, sy, lu, lv);
    o3[0] = r.o.x; o3[1] = r.o.y; o3[2] = r.o.z;
    d3[0] = r.d.x; d3[1] = r.d.y; d3[2] = r.d.z;
    return 0;
}
extern "C" int or_bsdf_probe(const or_scene* s, int32_t m, const float* lam4, float mask, const float* u4,
                             const float* wo3, const float* wi3, float* out) {
    if (!s || m < 0 || (size_t)m >= s->sc.materials.size()) return 1;
    Wavelengths wl;
    for (int i = 0; i < N; i++) {
        wl.lambda[i] = lam4[i];
        wl.pdf[i] = 1.0f;
    }
    V3 n(0.0f, 0.0f, 1.0f);
    Frame fr(n);
    HitCtx ctx{V3(0.1f, 0.2f, 0.3f), n, &fr, s->sc.turntable_deg};
    Bsdf b = material_bsdf(s->sc.materials[m], wl, mask, u4[0], ctx);
    V3 wo = ld(wo3), wi = ld(wi3);
    Spec e = b.eval(wo, wi);
    for (int i = 0; i < 4; i++) out[i] = e.v[i];
    out[4] = b.pdf(wo, wi);
    BsdfSample smp{};
    bool ok = b.sample(wo, u4[1], u4[2], u4[3], &smp);
    out[5] = ok ? 1.0f : 0.0f;
    out[6] = smp.wi.x; out[7] = smp.wi.y; out[8] = smp.wi.z;
    for (int i = 0; i < 4; i++) out[9 + i] = ok ? smp.f.v[i] : 0.0f;
    out[13] = ok ? smp.pdf : 0.0f;
    out[14] = smp.delta ? 1.0f : 0.0f;
    out[15] = smp.transmit ? 1.0f : 0.0f;
    out[16] = smp.terminate_secondary ? 1.0f : 0.0f;
    out[17] = b.is_delta() ? 1.0f : 0.0f;
    return 0;
}

// ------------------------------------------------------------------ environment map probes
extern "C" int or_linear_rgb(float* basis3x471, float* inv9) {
    const orc::LinRgb& L = orc::lin_rgb();
    std::memcpy(basis3x471, L.basis, sizeof(L.basis));
    std::memcpy(inv9, L.inv, sizeof(L.inv));
    return 0;
}

extern "C" int or_env_build(const float* rgb, int32_t w, int32_t h, float exposure_ev, float* weights,
                            float* marginal, float* conditional, float* total) {
    if (!rgb || w <= 0 || h <= 0) return 1;
    orc::env_build(rgb, w, h, exposure_ev, weights, marginal, conditional, total);
    return 0;
}

// EnvMap::sample / pdf (envmap.rs:206-239) over the given tables: n draws
// from (u1, u2) pairs; out: dir (3) and pdf per draw.  With dirs_only, the
// pdf of the given directions instead.
extern "C" int or_env_sample(const float* marginal, const float* conditional, int32_t w, int32_t h, float rotation_deg,
                             const float* u, uint32_t n, float* dirs, float* pdfs) {
    orc::EnvMap E{w, h, nullptr, marginal, conditional, rotation_deg, nullptr};
    for (uint32_t i = 0; i < n; i++) {
        float p = 0.0f;
        orc::V3 d = E.sample(u[2 * i], u[2 * i + 1], &p);
        dirs[3 * i] = d.x;
        dirs[3 * i + 1] = d.y;
        dirs[3 * i + 2] = d.z;
        pdfs[i] = p;
    }
    return 0;
}
extern "C" int or_env_pdf(const float* marginal, const float* conditional, int32_t w, int32_t h, float rotation_deg,
                          const float* dirs, uint32_t n, float* pdfs) {
    orc::EnvMap E{w, h, nullptr, marginal, conditional, rotation_deg, nullptr};
    for (uint32_t i = 0; i < n; i++) pdfs[i] = E.pdf(orc::V3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
    return 0;
}

