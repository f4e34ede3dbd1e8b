// Device functions of the render core.  The whole translation unit is
// compiled with -ffp-contract=off so every expression rounds as written; the
// operation order follows the reference (glam Vec3A: dot = (x*x'+y*y')+z*z',
// normalize = v * (1/sqrt(dot))), so the GPU and a CPU evaluation of the same
// formula agree to the last bit wherever only +, -, *, / and sqrt are used.
#pragma once
#include <hip/hip_runtime.h>

#include "dev_types.hpp"

namespace ag {

#define AGD __device__ __forceinline__

constexpr float PI_F = 3.14159265358979323846f;
constexpr float FRAC_1_PI_F = 0.318309886183790671538f;
constexpr float FRAC_PI_2_F = 1.57079632679489661923f;
constexpr float FRAC_PI_4_F = 0.785398163397448309616f;
constexpr float INF_F = __builtin_huge_valf();

struct f3 {
    float x, y, z;
};
AGD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
AGD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
AGD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
AGD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
AGD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
AGD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
AGD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
AGD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
AGD float length_sq(f3 a) { return dot(a, a); }
AGD float length(f3 a) { return sqrtf(dot(a, a)); }
AGD f3 normalize(f3 a) { return a * (1.0f / length(a)); }
AGD float max_abs(f3 a) { return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fabsf(a.z)); }
AGD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }

// geom.rs:94-97 with precomputed sin/cos
AGD f3 rot_y(f3 v, float s, float c) { return f3{c * v.x + s * v.z, v.y, -s * v.x + c * v.z}; }

// geom.rs:26-31
AGD f3 offset_origin(f3 p, f3 ng, f3 dir) {
    float scale = 1e-4f * (1.0f + max_abs(p));
    f3 n = dot(ng, dir) >= 0.0f ? ng : -ng;
    return p + n * scale;
}

// geom.rs:33-66 (Duff et al. 2017)
struct Frame {
    f3 t, b, n;
};
AGD Frame make_frame(f3 n) {
    float sign = copysignf(1.0f, n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    Frame f;
    f.t = f3{1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    f.b = f3{b, sign + n.y * n.y * a, -n.y};
    f.n = n;
    return f;
}
AGD f3 to_local(const Frame& f, f3 v) { return f3{dot(v, f.t), dot(v, f.b), dot(v, f.n)}; }
AGD f3 to_world(const Frame& f, f3 v) { return f.t * v.x + f.b * v.y + f.n * v.z; }
AGD f3 reflect(f3 wo, f3 n) { return -wo + n * (2.0f * dot(wo, n)); }

// geom.rs:75-85
AGD bool refract(f3 wi, f3 n, float eta, f3* out) {
    float cos_i = dot(n, wi);
    float sin2_i = fmaxf(1.0f - cos_i * cos_i, 0.0f);
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return false;
    float cos_t = sqrtf(1.0f - sin2_t);
    *out = (-wi) / eta + n * (cos_i / eta - cos_t);
    return true;
}

// ------------------------------------------------------------------ Spec (4 wavelengths)
struct Spec {
    float v[4];
};
AGD Spec splat(float s) { return Spec{{s, s, s, s}}; }
AGD Spec from4(float4 a) { return Spec{{a.x, a.y, a.z, a.w}}; }
AGD float4 to4(const Spec& s) { return make_float4(s.v[0], s.v[1], s.v[2], s.v[3]); }
AGD Spec operator+(const Spec& a, const Spec& b) { return Spec{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}}; }
AGD Spec operator*(const Spec& a, const Spec& b) { return Spec{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}}; }
AGD Spec operator*(const Spec& a, float s) { return Spec{{a.v[0] * s, a.v[1] * s, a.v[2] * s, a.v[3] * s}}; }
AGD Spec operator/(const Spec& a, float s) { return Spec{{a.v[0] / s, a.v[1] / s, a.v[2] / s, a.v[3] / s}}; }
AGD float max_value(const Spec& a) { return fmaxf(fmaxf(fmaxf(a.v[0], a.v[1]), a.v[2]), a.v[3]); }
AGD float min_value(const Spec& a) { return fminf(fminf(fminf(a.v[0], a.v[1]), a.v[2]), a.v[3]); }
AGD float average(const Spec& a) { return (((a.v[0] + a.v[1]) + a.v[2]) + a.v[3]) / 4.0f; }
AGD bool is_black(const Spec& a) { return a.v[0] == 0.0f && a.v[1] == 0.0f && a.v[2] == 0.0f && a.v[3] == 0.0f; }
AGD bool is_finite(const Spec& a) {
    return isfinite(a.v[0]) && isfinite(a.v[1]) && isfinite(a.v[2]) && isfinite(a.v[3]);
}
AGD Spec lerp(float t, const Spec& a, const Spec& b) { return a * (1.0f - t) + b * t; }

// integrator.rs:34-46
AGD void add_clamped(Spec& l, const Spec& c, bool clamp, float m) {
    if (clamp) {
        float mx = max_value(c);
        if (mx > m) {
            l = l + c * (m / mx);
            return;
        }
    }
    l = l + c;
}

// ------------------------------------------------------------------ sampling
// sampler.rs:18-25
AGD uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
// The stratified dimensions of sampler.rs:42-55: Owen-scrambled Sobol in
// 4-dimension sets (Burley 2020, "Practical Hash-based Owen Scrambling":
// nested uniform scrambling by a Laine-Karras hash, shuffled index), with
// Joe-Kuo direction numbers for the first four Sobol dimensions.  A pure
// function of (pixel seed, sample index, dimension).  Past 256 dimensions or
// 2^16 samples it is a plain hash of the same three numbers.
__device__ __constant__ uint32_t SOBOL_DIRS[4][32] = {
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
AGD uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x += seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
AGD uint32_t nested_scramble(uint32_t x, uint32_t seed) { return __brev(lk_permute(__brev(x), seed)); }
// Byte tables for Sobol dimensions 1..3 of a set: word [k-1][j][v] is the XOR
// of the direction numbers SOBOL_DIRS[k][8j + b] over the set bits b of the
// byte value v, so a sample is four table reads instead of a 32-step loop of
// dependent loads.  A kernel fills them into LDS once per block.
constexpr int SOBOL_TAB_WORDS = 3 * 4 * 256;
AGD void sobol_tables_fill(AG_LDS uint32_t* tab) {
    for (int e = (int)threadIdx.x; e < SOBOL_TAB_WORDS; e += (int)blockDim.x) {
        const int k = 1 + e / 1024, j = (e >> 8) & 3, v = e & 255;
        uint32_t x = 0;
        for (int b = 0; b < 8; b++)
            if ((v >> b) & 1) x ^= SOBOL_DIRS[k][8 * j + b];
        tab[e] = x;
    }
}
AGD float rng_dim(const AG_LDS uint32_t* sob, uint32_t pix_seed, uint32_t si, uint32_t dim) {
    if (dim >= 256u || si >= 65536u) {
        uint32_t h = hash32(pix_seed ^ hash32(si ^ hash32(dim ^ 0x2f0b6e35u)));
        return (float)(h >> 8) * (1.0f / 16777216.0f);
    }
    uint32_t set = dim >> 2, k = dim & 3u;
    uint32_t s = hash32(pix_seed ^ hash32(set * 0x9e3779b9u + 0x632be5abu));
    const uint32_t i = nested_scramble(si, s);
    uint32_t x;  // XOR of SOBOL_DIRS[k][b] over the set bits b of i
    if (k == 0) {
        x = __brev(i);  // dimension 0: direction number b is 2^(31-b)
    } else {
        const AG_LDS uint32_t* t = sob + (k - 1u) * 1024u;
        x = t[i & 255u] ^ t[256u + ((i >> 8) & 255u)] ^ t[512u + ((i >> 16) & 255u)] ^ t[768u + (i >> 24)];
    }
    uint32_t hk = hash32(k + 1u);
    x = nested_scramble(x, s ^ (hk + (s << 6) + (s >> 2)));
    return (float)(x >> 8) * (1.0f / 16777216.0f);
}
// sampler.rs:64-70
AGD float next_random(uint64_t& s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t x = (uint32_t)(((s >> 18) ^ s) >> 27);
    uint32_t rot = (uint32_t)(s >> 59);
    uint32_t v = (x >> rot) | (x << ((32u - rot) & 31u));
    return (float)(v >> 8) * (1.0f / 16777216.0f);
}
AGD uint64_t rng_init(uint32_t pix_seed, uint32_t si) {
    return (((uint64_t)pix_seed << 32) | (uint64_t)si) ^ 0x9e3779b97f4a7c15ull;
}
// sampler.rs:81-108
AGD float tent(float u, float r) {
    u = 2.0f * u;
    return u < 1.0f ? r * (sqrtf(u) - 1.0f) : r * (1.0f - sqrtf(2.0f - u));
}
AGD void disk(float u, float v, float* x, float* y) {
    float a = 2.0f * u - 1.0f, b = 2.0f * v - 1.0f;
    if (a == 0.0f && b == 0.0f) {
        *x = 0.0f;
        *y = 0.0f;
        return;
    }
    float r, phi;
    if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = FRAC_PI_4_F * (b / a);
    } else {
        r = b;
        phi = FRAC_PI_2_F - FRAC_PI_4_F * (a / b);
    }
    *x = r * cosf(phi);
    *y = r * sinf(phi);
}
AGD f3 cosine_hemisphere(float u, float v) {
    float x, y;
    disk(u, v, &x, &y);
    float z = sqrtf(fmaxf(1.0f - x * x - y * y, 0.0f));
    return f3{x, y, z};
}

// spec.rs:157-220 (atanh as Rust's 0.5 * ln_1p(2x / (1 - x)))
AGD float visible_pdf(float l) {
    if (!(l >= 360.0f && l <= 830.0f)) return 0.0f;
    float c = coshf(0.0072f * (l - 538.0f));
    return 0.0039398042f / (c * c);
}
AGD float sample_visible_wavelength(float u) {
    float x = 0.8569106f - 1.827502f * u;
    float at = 0.5f * log1pf((2.0f * x) / (1.0f - x));
    return 538.0f - 138.88889f * at;
}
AGD void sample_wavelengths(float u, Spec& lam, Spec& pdf) {
    for (int i = 0; i < 4; i++) {
        float up = u + (float)i / 4.0f;
        if (up > 1.0f) up -= 1.0f;
        lam.v[i] = sample_visible_wavelength(up);
        pdf.v[i] = visible_pdf(lam.v[i]);
    }
}
AGD void terminate_secondary(Spec& pdf) {
    if (pdf.v[1] == 0.0f && pdf.v[2] == 0.0f && pdf.v[3] == 0.0f) return;
    pdf.v[1] = pdf.v[2] = pdf.v[3] = 0.0f;
    pdf.v[0] /= 4.0f;
}

// ------------------------------------------------------------------ tables
// tab.rs:44-52: 1 nm table
AGD float tab_eval(const float* v, float l) {
    float x = fminf(fmaxf(l - 360.0f, 0.0f), 470.0f);
    int i = min((int)x, 469);
    float f = x - (float)i;
    return v[i] * (1.0f - f) + v[i + 1] * f;
}
AGD Spec tab_sample(const float* v, const Spec& lam) {
    Spec r;
    for (int i = 0; i < 4; i++) r.v[i] = tab_eval(v, lam.v[i]);
    return r;
}
// spectrum.rs:80-93 on the uniform 5 nm grid 360 + 5 i (95 points): the
// same interval and weight as the binary search
AGD float grid_eval(const float* v, float l) {
    if (l <= 360.0f) return v[0];
    if (l >= 830.0f) return v[94];
    int i = (int)((l - 360.0f) / 5.0f);
    i = min(max(i, 0), 93);
    if (360.0f + 5.0f * (float)i > l) i--;
    else if (360.0f + 5.0f * (float)(i + 1) <= l) i++;
    float li = 360.0f + 5.0f * (float)i;
    float t = (l - li) / ((360.0f + 5.0f * (float)(i + 1)) - li);
    return v[i] * (1.0f - t) + v[i + 1] * t;
}
// cie.rs:38-49
AGD void xyz_bar(const float* cmf, float l, float* out) {
    float x = l - 360.0f;
    if (!(x >= 0.0f && x <= 470.0f)) {
        out[0] = out[1] = out[2] = 0.0f;
        return;
    }
    int i = min((int)x, 469);
    float t = x - (float)i;
    const float* a = cmf + 3 * i;
    const float* b = a + 3;
    out[0] = a[0] + (b[0] - a[0]) * t;
    out[1] = a[1] + (b[1] - a[1]) * t;
    out[2] = a[2] + (b[2] - a[2]) * t;
}

// ------------------------------------------------------------------ Fresnel
// fresnel.rs:9-25
AGD float fresnel_dielectric(float cos_theta_i, float eta) {
    float cos_i = fminf(fmaxf(cos_theta_i, -1.0f), 1.0f);
    if (cos_i < 0.0f) {
        eta = 1.0f / eta;
        cos_i = -cos_i;
    }
    float sin2_i = 1.0f - cos_i * cos_i;
    float sin2_t = sin2_i / (eta * eta);
    if (sin2_t >= 1.0f) return 1.0f;
    float cos_t = sqrtf(fmaxf(1.0f - sin2_t, 0.0f));
    float r_parl = (eta * cos_i - cos_t) / (eta * cos_i + cos_t);
    float r_perp = (cos_i - eta * cos_t) / (cos_i + eta * cos_t);
    return 0.5f * (r_parl * r_parl + r_perp * r_perp);
}

struct Cf {
    float re, im;
};
AGD Cf cadd(Cf a, Cf b) { return Cf{a.re + b.re, a.im + b.im}; }
AGD Cf csub(Cf a, Cf b) { return Cf{a.re - b.re, a.im - b.im}; }
AGD Cf cmul(Cf a, Cf b) { return Cf{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
AGD Cf cscale(Cf a, float s) { return Cf{a.re * s, a.im * s}; }
AGD Cf cdiv(Cf a, Cf o) {
    float d = o.re * o.re + o.im * o.im;
    return Cf{(a.re * o.re + a.im * o.im) / d, (a.im * o.re - a.re * o.im) / d};
}
AGD float cnorm(Cf a) { return a.re * a.re + a.im * a.im; }
AGD Cf csqrt(Cf a) {
    float n = sqrtf(cnorm(a));
    if (n == 0.0f) return Cf{0.0f, 0.0f};
    float t1 = sqrtf(0.5f * (n + fabsf(a.re)));
    float t2 = 0.5f * a.im / t1;
    if (a.re >= 0.0f) return Cf{t1, t2};
    return Cf{fabsf(t2), copysignf(t1, a.im)};
}
// fresnel.rs:81-93
AGD float fresnel_conductor1(float cos_theta_i, float eta, float k) {
    float cos_i = fminf(fmaxf(cos_theta_i, 0.0f), 1.0f);
    float sin2_i = 1.0f - cos_i * cos_i;
    Cf e{eta, k};
    Cf sin2_t = cdiv(Cf{sin2_i, 0.0f}, cmul(e, e));
    Cf cos_t = csqrt(csub(Cf{1.0f, 0.0f}, sin2_t));
    Cf ec = cscale(e, cos_i);
    Cf r_parl = cdiv(csub(ec, cos_t), cadd(ec, cos_t));
    Cf ci{cos_i, 0.0f};
    Cf ect = cmul(e, cos_t);
    Cf r_perp = cdiv(csub(ci, ect), cadd(ci, ect));
    return 0.5f * (cnorm(r_parl) + cnorm(r_perp));
}
AGD Spec fresnel_conductor(float c, const Spec& eta, const Spec& k) {
    Spec r;
    for (int i = 0; i < 4; i++) r.v[i] = fresnel_conductor1(c, eta.v[i], k.v[i]);
    return r;
}

// ------------------------------------------------------------------ GGX
// microfacet.rs:9-110
struct Ggx {
    float alpha;
    const float* energy;
};
AGD bool ggx_smooth(float alpha) { return alpha < 1.5e-3f; }
AGD float ggx_d(float alpha, f3 h) {
    if (h.z <= 0.0f) return 0.0f;
    float a2 = alpha * alpha;
    float c2 = h.z * h.z;
    float k = c2 * (a2 - 1.0f) + 1.0f;
    return a2 / (PI_F * k * k);
}
AGD float ggx_lambda(float alpha, f3 w) {
    float c2 = w.z * w.z;
    float t2 = fmaxf(1.0f - c2, 0.0f) / fmaxf(c2, 1e-12f);
    if (!isfinite(t2)) return 0.0f;
    return 0.5f * (-1.0f + sqrtf(1.0f + alpha * alpha * t2));
}
AGD float ggx_g1(float alpha, f3 w) { return 1.0f / (1.0f + ggx_lambda(alpha, w)); }
AGD float ggx_g2(float alpha, f3 wo, f3 wi) { return 1.0f / (1.0f + ggx_lambda(alpha, wo) + ggx_lambda(alpha, wi)); }
AGD float ggx_d_visible(float alpha, f3 wo, f3 h) {
    float c = fmaxf(fabsf(wo.z), 1e-8f);
    return ggx_g1(alpha, wo) * fmaxf(dot(wo, h), 0.0f) * ggx_d(alpha, h) / c;
}
AGD f3 ggx_sample_h(float a, f3 wo, float u1, float u2) {
    f3 vh = normalize(f3{a * wo.x, a * wo.y, wo.z});
    float lensq = vh.x * vh.x + vh.y * vh.y;
    f3 t1 = lensq > 1e-12f ? f3{-vh.y, vh.x, 0.0f} / sqrtf(lensq) : f3{1.0f, 0.0f, 0.0f};
    f3 t2 = cross(vh, t1);
    float r = sqrtf(u1);
    float phi = 2.0f * PI_F * u2;
    float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * sqrtf(fmaxf(1.0f - p1 * p1, 0.0f)) + s * p2;
    f3 nh = t1 * p1 + t2 * p2 + vh * sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f));
    return normalize(f3{a * nh.x, a * nh.y, fmaxf(nh.z, 1e-6f)});
}
AGD float ggx_e(const float* energy, float mu, float alpha) {
    const int EN = 32;
    float x = fminf(fminf(fmaxf(mu, 0.0f), 1.0f) * (float)(EN - 1), (float)(EN - 1) - 1e-4f);
    float y = fminf(fminf(fmaxf(sqrtf(alpha), 0.0f), 1.0f) * (float)(EN - 1), (float)(EN - 1) - 1e-4f);
    int xi = (int)x, yi = (int)y;
    float fx = x - (float)xi, fy = y - (float)yi;
    float a = energy[yi * EN + xi] * (1.0f - fx) + energy[yi * EN + xi + 1] * fx;
    float b = energy[(yi + 1) * EN + xi] * (1.0f - fx) + energy[(yi + 1) * EN + xi + 1] * fx;
    return fminf(fmaxf(a * (1.0f - fy) + b * fy, 0.05f), 1.0f);
}

// ------------------------------------------------------------------ BSDF
// bsdf.rs:10-27
enum { B_CONDUCTOR = 0, B_COATED = 1, B_DIFFUSE = 2, B_DIELECTRIC = 3, B_BLACK = 4, B_OPAL = 5 };
// The BSDF kinds a caller can meet, as a bit mask over B_*: a shading kernel
// specialised to one family of materials compiles only that family's lobes
// (a kind outside the mask evaluates as black).
constexpr uint32_t KIND_ALL = 0x3Fu;
constexpr uint32_t kind_bit(int k) { return 1u << k; }
constexpr bool has_kind(uint32_t km, int k) { return ((km >> k) & 1u) != 0u; }
struct Bsdf {
    int kind;
    float alpha;
    float eta1;      // coated eta
    Spec a, b;       // conductor eta/k, coated base, diffuse albedo, dielectric eta, opal body / lambda
    Spec f0;         // conductor: Fresnel at normal incidence
    f3 on;           // opal lattice normal (local)
    float spacing, strength, w;
};
struct BSample {
    f3 wi;
    Spec f;
    float pdf;
    bool delta, transmit, term2;
};

template <uint32_t KM = KIND_ALL>
AGD bool bsdf_is_delta(const Bsdf& b) {
    return (has_kind(KM, B_DIELECTRIC) && b.kind == B_DIELECTRIC) ||
           (has_kind(KM, B_CONDUCTOR) && b.kind == B_CONDUCTOR && ggx_smooth(b.alpha));
}

AGD float coat_spec_prob(float cos_o, float eta, const Spec& base) {
    float f = fresnel_dielectric(cos_o, eta);
    float d = average(base) * (1.0f - f);
    return fminf(fmaxf(f / fmaxf(f + d, 1e-6f), 0.15f), 0.95f);
}

AGD Spec coated_eval(float alpha, float eta, const Spec& base, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
    f3 h = normalize(wo + wi);
    float fs = fresnel_dielectric(dot(wo, h), eta);
    float spec = ggx_d(alpha, h) * ggx_g2(alpha, wo, wi) * fs / (4.0f * wo.z * wi.z);
    float fo = fresnel_dielectric(wo.z, eta);
    float fi = fresnel_dielectric(wi.z, eta);
    return base * ((1.0f - fo) * (1.0f - fi) * FRAC_1_PI_F) + splat(spec);
}
AGD float coated_pdf(float alpha, float eta, const Spec& base, f3 wo, f3 wi) {
    if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
    float ps = coat_spec_prob(wo.z, eta, base);
    f3 h = normalize(wo + wi);
    float p_spec = ggx_d_visible(alpha, wo, h) / (4.0f * fmaxf(dot(wo, h), 1e-8f));
    return ps * p_spec + (1.0f - ps) * wi.z * FRAC_1_PI_F;
}

// bsdf.rs:257-292
AGD float bragg(float spacing_nm, float cos_t, float l) {
    float peak = 2.0f * spacing_nm * 1.37f * fminf(fmaxf(cos_t, 0.0f), 1.0f);
    float sigma = 0.045f * fmaxf(peak, 1.0f);
    float x = (l - peak) / sigma;
    return expf(-0.5f * x * x);
}
AGD Spec opal_flake_eval(const Bsdf& b, f3 wo, f3 wi) {
    Frame fr = make_frame(b.on);
    f3 a = to_local(fr, wo), c = to_local(fr, wi);
    if (a.z <= 0.0f || c.z <= 0.0f) return splat(0.0f);
    f3 h = normalize(a + c);
    float g = ggx_d(b.alpha, h) * ggx_g2(b.alpha, a, c) / (4.0f * a.z * c.z);
    float cos_t = dot(a, h);
    Spec r;
    for (int i = 0; i < 4; i++) r.v[i] = b.strength * bragg(b.spacing, cos_t, b.b.v[i]);
    return r * (g * c.z / fmaxf(wi.z, 1e-6f));
}
AGD float opal_flake_pdf(const Bsdf& b, f3 wo, f3 wi) {
    Frame fr = make_frame(b.on);
    f3 a = to_local(fr, wo), c = to_local(fr, wi);
    if (a.z <= 0.0f || c.z <= 0.0f) return 0.0f;
    f3 h = normalize(a + c);
    return ggx_d_visible(b.alpha, a, h) / (4.0f * fmaxf(dot(a, h), 1e-8f));
}
constexpr float OPAL_BASE_ALPHA = 0.12f * 0.12f;  // Ggx::from_roughness(0.12)

// bsdf.rs:53-93
template <uint32_t KM = KIND_ALL>
AGD Spec bsdf_eval(const Bsdf& b, const float* energy, f3 wo, f3 wi) {
    switch (b.kind) {
        case B_CONDUCTOR: {
            if constexpr (!has_kind(KM, B_CONDUCTOR)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f || ggx_smooth(b.alpha)) return splat(0.0f);
            f3 h = normalize(wo + wi);
            Spec f = fresnel_conductor(dot(wo, h), b.a, b.b);
            float base = ggx_d(b.alpha, h) * ggx_g2(b.alpha, wo, wi) / (4.0f * wo.z * wi.z);
            float e = ggx_e(energy, wo.z, b.alpha);
            const Spec f0 = b.f0;
            float q = (1.0f - e) / e;
            Spec r;
            for (int i = 0; i < 4; i++) r.v[i] = (f.v[i] * base) * (1.0f + f0.v[i] * q);
            return r;
        }
        case B_COATED:
            if constexpr (!has_kind(KM, B_COATED)) break;
            return coated_eval(b.alpha, b.eta1, b.a, wo, wi);
        case B_DIFFUSE:
            if constexpr (!has_kind(KM, B_DIFFUSE)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
            return b.a * FRAC_1_PI_F;
        case B_OPAL: {
            if constexpr (!has_kind(KM, B_OPAL)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f) return splat(0.0f);
            Spec base = coated_eval(OPAL_BASE_ALPHA, 1.45f, b.a, wo, wi);
            Spec flake = opal_flake_eval(b, wo, wi);
            return base * (1.0f - b.w) + flake * b.w;
        }
        default: break;
    }
    return splat(0.0f);
}

// bsdf.rs:95-129
template <uint32_t KM = KIND_ALL>
AGD float bsdf_pdf(const Bsdf& b, f3 wo, f3 wi) {
    switch (b.kind) {
        case B_CONDUCTOR: {
            if constexpr (!has_kind(KM, B_CONDUCTOR)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f || ggx_smooth(b.alpha)) return 0.0f;
            f3 h = normalize(wo + wi);
            return ggx_d_visible(b.alpha, wo, h) / (4.0f * fmaxf(dot(wo, h), 1e-8f));
        }
        case B_COATED:
            if constexpr (!has_kind(KM, B_COATED)) break;
            return coated_pdf(b.alpha, b.eta1, b.a, wo, wi);
        case B_DIFFUSE:
            if constexpr (!has_kind(KM, B_DIFFUSE)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
            return wi.z * FRAC_1_PI_F;
        case B_OPAL: {
            if constexpr (!has_kind(KM, B_OPAL)) break;
            if (wo.z <= 0.0f || wi.z <= 0.0f) return 0.0f;
            float pb = coated_pdf(OPAL_BASE_ALPHA, 1.45f, b.a, wo, wi);
            return pb * (1.0f - b.w) + opal_flake_pdf(b, wo, wi) * b.w;
        }
        default: break;
    }
    return 0.0f;
}

AGD bool coated_sample(float alpha, float eta, const Spec& base, f3 wo, float u0, float u1, float u2, BSample* s) {
    if (wo.z <= 0.0f) return false;
    float ps = coat_spec_prob(wo.z, eta, base);
    f3 wi;
    if (u0 < ps)
        wi = reflect(wo, ggx_sample_h(alpha, wo, u1, u2));
    else
        wi = cosine_hemisphere(u1, u2);
    if (wi.z <= 0.0f) return false;
    float pdf = coated_pdf(alpha, eta, base, wo, wi);
    if (pdf <= 0.0f) return false;
    s->wi = wi;
    s->f = coated_eval(alpha, eta, base, wo, wi);
    s->pdf = pdf;
    s->delta = s->transmit = s->term2 = false;
    return true;
}

// bsdf.rs:132-239
template <uint32_t KM = KIND_ALL>
AGD bool bsdf_sample(const Bsdf& b, const float* energy, f3 wo, float u0, float u1, float u2, BSample* s) {
    switch (b.kind) {
        case B_CONDUCTOR: {
            if constexpr (!has_kind(KM, B_CONDUCTOR)) break;
            if (wo.z <= 0.0f) return false;
            if (ggx_smooth(b.alpha)) {
                f3 wi{-wo.x, -wo.y, wo.z};
                Spec f = fresnel_conductor(wo.z, b.a, b.b);
                s->f = f / fmaxf(wi.z, 1e-8f);
                s->wi = wi;
                s->pdf = 1.0f;
                s->delta = true;
                s->transmit = s->term2 = false;
                return true;
            }
            f3 h = ggx_sample_h(b.alpha, wo, u1, u2);
            f3 wi = reflect(wo, h);
            if (wi.z <= 0.0f) return false;
            s->pdf = ggx_d_visible(b.alpha, wo, h) / (4.0f * fmaxf(dot(wo, h), 1e-8f));
            s->f = bsdf_eval<KM>(b, energy, wo, wi);
            s->wi = wi;
            s->delta = s->transmit = s->term2 = false;
            return true;
        }
        case B_COATED:
            if constexpr (!has_kind(KM, B_COATED)) break;
            return coated_sample(b.alpha, b.eta1, b.a, wo, u0, u1, u2, s);
        case B_DIFFUSE: {
            if constexpr (!has_kind(KM, B_DIFFUSE)) break;
            if (wo.z <= 0.0f) return false;
            f3 wi = cosine_hemisphere(u1, u2);
            float pdf = wi.z * FRAC_1_PI_F;
            if (pdf <= 0.0f) return false;
            s->wi = wi;
            s->f = b.a * FRAC_1_PI_F;
            s->pdf = pdf;
            s->delta = s->transmit = s->term2 = false;
            return true;
        }
        case B_DIELECTRIC: {
            if constexpr (!has_kind(KM, B_DIELECTRIC)) break;
            bool dispersive = (max_value(b.a) - min_value(b.a)) > 1e-5f;
            float e0 = b.a.v[0];
            float cos_o = wo.z;
            float f0 = fresnel_dielectric(cos_o, e0);
            if (u0 < f0) {
                f3 wi{-wo.x, -wo.y, wo.z};
                float c = fmaxf(fabsf(wi.z), 1e-8f);
                Spec f;
                for (int i = 0; i < 4; i++) f.v[i] = fresnel_dielectric(cos_o, b.a.v[i]);
                s->f = f / c;
                s->wi = wi;
                s->pdf = f0;
                s->delta = true;
                s->transmit = s->term2 = false;
                return true;
            }
            bool entering = cos_o > 0.0f;
            float etap = entering ? e0 : 1.0f / e0;
            f3 n = entering ? f3{0.0f, 0.0f, 1.0f} : f3{-0.0f, -0.0f, -1.0f};
            f3 wi;
            if (!refract(wo, n, etap, &wi)) return false;
            float c = fmaxf(fabsf(wi.z), 1e-8f);
            float ft = (1.0f - f0) / (c * etap * etap);
            s->wi = wi;
            s->f = splat(ft);
            s->pdf = 1.0f - f0;
            s->delta = true;
            s->transmit = true;
            s->term2 = dispersive;
            return true;
        }
        case B_OPAL: {
            if constexpr (!has_kind(KM, B_OPAL)) break;
            if (wo.z <= 0.0f) return false;
            f3 wi;
            if (u0 < b.w) {
                Frame fr = make_frame(b.on);
                f3 wol = to_local(fr, wo);
                if (wol.z <= 0.0f) return false;
                f3 h = ggx_sample_h(b.alpha, wol, u1, u2);
                wi = to_world(fr, reflect(wol, h));
            } else {
                BSample bs;
                if (!coated_sample(OPAL_BASE_ALPHA, 1.45f, b.a, wo, (u0 - b.w) / fmaxf(1.0f - b.w, 1e-6f), u1, u2, &bs))
                    return false;
                wi = bs.wi;
            }
            if (wi.z <= 0.0f) return false;
            float pdf = bsdf_pdf<KM>(b, wo, wi);
            if (pdf <= 0.0f) return false;
            s->wi = wi;
            s->f = bsdf_eval<KM>(b, energy, wo, wi);
            s->pdf = pdf;
            s->delta = s->transmit = s->term2 = false;
            return true;
        }
        default: break;
    }
    return false;
}

// ------------------------------------------------------------------ opal domains
// material.rs:195-237
AGD uint32_t hash3i(int i, int j, int k, uint32_t salt) {
    uint32_t h = ((uint32_t)i * 0x8da6b343u) ^ ((uint32_t)j * 0xd8163841u) ^ ((uint32_t)k * 0xcb1ab31fu) ^ salt;
    return hash32(h);
}
AGD float unit_h(uint32_t h) { return (float)(h >> 8) / (float)(1u << 24); }
AGD void opal_domain(f3 p, f3* dir, float* spacing, float* brightness) {
    int fx = (int)floorf(p.x), fy = (int)floorf(p.y), fz = (int)floorf(p.z);
    float best = 3.4028235e38f;
    int bi = 0, bj = 0, bk = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int i = fx + dx, j = fy + dy, k = fz + dz;
                f3 c{(float)i + unit_h(hash3i(i, j, k, 1)), (float)j + unit_h(hash3i(i, j, k, 2)),
                     (float)k + unit_h(hash3i(i, j, k, 3))};
                float d2 = length_sq(c - p);
                if (d2 < best) {
                    best = d2;
                    bi = i;
                    bj = j;
                    bk = k;
                }
            }
    float z = 2.0f * unit_h(hash3i(bi, bj, bk, 4)) - 1.0f;
    float phi = (2.0f * PI_F) * unit_h(hash3i(bi, bj, bk, 5));
    float r = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    *dir = f3{r * cosf(phi), r * sinf(phi), z};
    *spacing = 190.0f + 75.0f * unit_h(hash3i(bi, bj, bk, 6));
    *brightness = 0.35f + 0.65f * powf(unit_h(hash3i(bi, bj, bk, 7)), 0.7f);
}

}  // namespace ag
