// TEST INFRASTRUCTURE ONLY — the oracle's preview development: a CPU
// restatement of argent-post's tone mapping (tonemap.rs, color.rs) and its
// built-in à-trous denoiser (denoise.rs:526-709), used by the tests as the
// checker of the HIP path.  Every function names the reference lines it
// follows.  Built with -ffp-contract=off and the host libm (powf, log2f,
// expf), as the reference's release build calls them.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace orp {

// Rust f32::clamp: NaN stays NaN (unlike fmin/fmax, which drop it)
static float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
static float bits_f(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
static uint32_t f_bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

const float LUMA[3] = {0.2126f, 0.7152f, 0.0722f};  // color.rs:7

// color.rs:11-19
float srgb_encode(float x) {
    if (x <= 0.0031308f) return 12.92f * std::fmax(x, 0.0f);
    if (x >= 1.0f) return 1.0f;
    return 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
}

// tonemap.rs:75-86
const float AGX_INSET[3][3] = {{0.84247906f, 0.0784336f, 0.079223745f},
                               {0.042328242f, 0.8784686f, 0.07916613f},
                               {0.042375655f, 0.0784336f, 0.879143f}};
const float AGX_OUTSET[3][3] = {{1.196879f, -0.09802088f, -0.09902974f},
                                {-0.05289685f, 1.1519031f, -0.09896118f},
                                {-0.052971635f, -0.09804345f, 1.1510737f}};
const float AGX_MIN_EV = -12.47393f, AGX_MAX_EV = 4.026069f;
const float AGX_OUT_MAX = 1.25f;  // tonemap.rs:225
const int LUT_SEGMENTS = 4096;    // tonemap.rs:223
const uint32_t CURVE_SHIFT = 23 - 9;

struct Cdl {
    float slope[3], offset[3], power[3], sat;
};
// tonemap.rs:96-102; look 0 Base, 1 Punchy, 2 Golden
Cdl look_cdl(int look) {
    if (look == 1) return Cdl{{1, 1, 1}, {0, 0, 0}, {1.35f, 1.35f, 1.35f}, 1.4f};
    if (look == 2) return Cdl{{1.0f, 0.9f, 0.5f}, {0, 0, 0}, {0.8f, 0.8f, 0.8f}, 1.3f};
    return Cdl{{1, 1, 1}, {0, 0, 0}, {1, 1, 1}, 1.0f};
}
// tonemap.rs:104-111
void mat3(const float m[3][3], const float v[3], float o[3]) {
    for (int r = 0; r < 3; r++) o[r] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2];
}
float luma(const float v[3]) { return LUMA[0] * v[0] + LUMA[1] * v[1] + LUMA[2] * v[2]; }
// tonemap.rs:118-122
float agx_log_encode(float v) {
    float ev = clampf(std::log2(std::fmax(v, 1e-10f)), AGX_MIN_EV, AGX_MAX_EV);
    return (ev - AGX_MIN_EV) / (AGX_MAX_EV - AGX_MIN_EV);
}
// tonemap.rs:125-128
float agx_sigmoid(float x) {
    return ((((((15.5f * x - 40.14f) * x + 31.96f) * x - 6.868f) * x + 0.4298f) * x + 0.1191f) * x) - 0.00232f;
}
// tonemap.rs:131-135
float cdl_sop(float x, float slope, float offset, float power) {
    float v = clampf(x * slope + offset, 0.0f, 1.0f);
    return power == 1.0f ? v : std::pow(v, power);
}
// tonemap.rs:138-145
void saturate(float v[3], float sat) {
    if (sat == 1.0f) {
        for (int c = 0; c < 3; c++) v[c] = clampf(v[c], 0.0f, 1.0f);
        return;
    }
    float l = luma(v);
    for (int c = 0; c < 3; c++) v[c] = clampf(l + sat * (v[c] - l), 0.0f, 1.0f);
}
// tonemap.rs:147-157
void agx_exact(const float x[3], int look, float sat_mul, float out[3]) {
    Cdl cdl = look_cdl(look);
    float v[3], e[3];
    mat3(AGX_INSET, x, v);
    for (int c = 0; c < 3; c++) e[c] = cdl_sop(agx_sigmoid(agx_log_encode(v[c])), cdl.slope[c], cdl.offset[c], cdl.power[c]);
    saturate(e, cdl.sat * sat_mul);
    mat3(AGX_OUTSET, e, out);
    for (int c = 0; c < 3; c++) out[c] = std::fmin(std::pow(std::fmax(out[c], 0.0f), 2.2f), 1.0f);
}
// tonemap.rs:162-176
const float ACES_IN[3][3] = {{0.59719f, 0.35458f, 0.04823f}, {0.07600f, 0.90834f, 0.01566f}, {0.02840f, 0.13383f, 0.83777f}};
const float ACES_OUT[3][3] = {{1.60475f, -0.53108f, -0.07367f}, {-0.10208f, 1.10813f, -0.00605f}, {-0.00327f, -0.07276f, 1.07602f}};
void aces_fitted(const float x[3], float out[3]) {
    float v[3];
    mat3(ACES_IN, x, v);
    for (int c = 0; c < 3; c++) {
        float a = v[c] * (v[c] + 0.0245786f) - 0.000090537f;
        float b = v[c] * (0.983729f * v[c] + 0.432951f) + 0.238081f;
        v[c] = a / b;
    }
    mat3(ACES_OUT, v, out);
    for (int c = 0; c < 3; c++) out[c] = clampf(out[c], 0.0f, 1.0f);
}
// tonemap.rs:179-194
void khronos_neutral(const float in[3], float c[3]) {
    const float START = 0.8f - 0.04f, DESAT = 0.15f;
    for (int k = 0; k < 3; k++) c[k] = in[k];
    float x = std::fmin(std::fmin(c[0], c[1]), c[2]);
    float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    for (int k = 0; k < 3; k++) c[k] = c[k] - offset;
    float peak = std::fmax(std::fmax(c[0], c[1]), c[2]);
    if (peak < START) return;
    float d = 1.0f - START;
    float new_peak = 1.0f - d * d / (peak + d - START);
    for (int k = 0; k < 3; k++) c[k] = c[k] * (new_peak / peak);
    float g = 1.0f - 1.0f / (DESAT * (peak - new_peak) + 1.0f);
    for (int k = 0; k < 3; k++) c[k] = c[k] + (new_peak - c[k]) * g;
}
// tonemap.rs:197-205
float sanitize(float v) { return std::isnan(v) ? 0.0f : clampf(v, 0.0f, 65504.0f); }
float exposure_scale(float ev) { return std::isfinite(ev) ? std::pow(2.0f, ev) : 1.0f; }
float grade_sat(float s) { return std::isfinite(s) ? std::fmax(s, 0.0f) : 1.0f; }

// tonemap.rs:207-218 (tonemap_pixel); op: 0-2 AgX Base/Punchy/Golden, 3 ACES
// fitted, 4 Khronos Neutral, 5 Reinhard, 6 Clamp
void tonemap_pixel(const float rgb[3], float ev, int op, float sat_in, float out[3]) {
    float k = exposure_scale(ev), x[3];
    for (int c = 0; c < 3; c++) x[c] = sanitize(rgb[c] * k);
    float sat = grade_sat(sat_in);
    switch (op) {
        case 0: case 1: case 2: agx_exact(x, op, sat, out); return;
        case 3: aces_fitted(x, out); break;
        case 4: khronos_neutral(x, out); break;
        case 5: for (int c = 0; c < 3; c++) out[c] = x[c] / (1.0f + x[c]); break;
        default: for (int c = 0; c < 3; c++) out[c] = x[c]; break;
    }
    saturate(out, sat);
}

// tonemap.rs:231-259
struct BitCurve {
    uint32_t base = 0;
    std::vector<float> t;
    BitCurve(float slope, float offset, float power) {
        float lo = std::pow(2.0f, AGX_MIN_EV) * 0.999f;
        float hi = std::pow(2.0f, AGX_MAX_EV) * 1.001f;
        base = f_bits(lo) >> CURVE_SHIFT;
        uint32_t top = (f_bits(hi) >> CURVE_SHIFT) + 1;
        for (uint32_t k = base; k <= top; k++)
            t.push_back(cdl_sop(agx_sigmoid(agx_log_encode(bits_f(k << CURVE_SHIFT))), slope, offset, power));
    }
    float eval(float v) const {
        float lo = bits_f(base << CURVE_SHIFT);
        float hi = bits_f((base + (uint32_t)t.size() - 1) << CURVE_SHIFT);
        uint32_t bits = f_bits(std::fmin(std::fmax(v, lo), hi));
        size_t k = (bits >> CURVE_SHIFT) - base;
        float f = (float)(bits & ((1u << CURVE_SHIFT) - 1)) * (1.0f / (float)(1u << CURVE_SHIFT));
        k = std::min(k, t.size() - 2);
        return t[k] * (1.0f - f) + t[k + 1] * f;
    }
};
// color.rs:41-58
struct Lut1 {
    float lo = 0, scale = 0;
    size_t last = 0;
    std::vector<float> t;
    template <class F>
    Lut1(float lo_, float hi, int segments, F f) : lo(lo_) {
        size_t n = (size_t)std::max(segments, 2);
        for (size_t i = 0; i <= n; i++) t.push_back(f(lo + (hi - lo) * (float)i / (float)n));
        scale = (float)n / (hi - lo);
        last = n;
    }
    float eval(float x) const {
        float u = (x - lo) * scale;
        u = u >= 0.0f ? std::fmin(u, (float)last) : 0.0f;
        size_t i = std::min((size_t)u, last - 1);
        float f = u - (float)i;
        return t[i] * (1.0f - f) + t[i + 1] * f;
    }
};

// tonemap.rs:262-338 (Prepared)
struct Prepared {
    float k;
    int op;
    float sat;
    std::vector<BitCurve> curves;
    int pick[3] = {0, 0, 0};
    Lut1 encode;
    static Lut1 make_encode(int op) {
        if (op <= 2)
            return Lut1(0.0f, AGX_OUT_MAX, LUT_SEGMENTS,
                        [](float o) { return srgb_encode(std::fmin(std::pow(std::fmax(o, 0.0f), 2.2f), 1.0f)); });
        return Lut1(0.0f, 1.0f, LUT_SEGMENTS, [](float o) { return srgb_encode(o); });
    }
    Prepared(float ev, int op_, float sat_in) : k(exposure_scale(ev)), op(op_), encode(make_encode(op_)) {
        float s = grade_sat(sat_in);
        if (op <= 2) {
            Cdl cdl = look_cdl(op);
            bool shared = true;
            for (int c = 0; c < 3; c++)
                shared = shared && cdl.slope[c] == cdl.slope[0] && cdl.offset[c] == cdl.offset[0] &&
                         cdl.power[c] == cdl.power[0];
            int n = shared ? 1 : 3;
            for (int c = 0; c < n; c++) curves.emplace_back(cdl.slope[c], cdl.offset[c], cdl.power[c]);
            if (!shared) pick[1] = 1, pick[2] = 2;
            sat = cdl.sat * s;
        } else {
            sat = s;
        }
    }
    void encoded(const float p[3], float out[3]) const {
        float x[3];
        for (int c = 0; c < 3; c++) x[c] = sanitize(p[c] * k);
        if (op <= 2) {
            float v[3], e[3], o[3];
            mat3(AGX_INSET, x, v);
            for (int c = 0; c < 3; c++) e[c] = curves[pick[c]].eval(v[c]);
            if (sat != 1.0f) saturate(e, sat);
            mat3(AGX_OUTSET, e, o);
            for (int c = 0; c < 3; c++) out[c] = encode.eval(o[c]);
            return;
        }
        float d[3];
        if (op == 3) {
            aces_fitted(x, d);
        } else if (op == 4) {
            khronos_neutral(x, d);
        } else if (op == 5) {
            for (int c = 0; c < 3; c++) d[c] = x[c] / (1.0f + x[c]);
        } else {
            for (int c = 0; c < 3; c++) d[c] = x[c];
        }
        saturate(d, sat);
        for (int c = 0; c < 3; c++) out[c] = encode.eval(d[c]);
    }
};

// tonemap.rs:341-367
uint64_t hash64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
void tpdf3(uint64_t i, float n[3]) {
    uint64_t h = hash64(i ^ 0x9e3779b97f4a7c15ull);
    const float S = 1.0f / 1024.0f;
    for (int c = 0; c < 3; c++) {
        float u0 = (float)((h >> (2 * c * 10)) & 0x3ff), u1 = (float)((h >> ((2 * c + 1) * 10)) & 0x3ff);
        n[c] = (u0 + u1 + 1.0f) * S - 1.0f;
    }
}
uint8_t quantize8(float v, float n) {
    if (v <= 0.0f) return 0;
    if (v >= 1.0f) return 255;
    float q = v * 255.0f + 0.5f + n;
    int i = std::isnan(q) ? 0 : (int)std::trunc(std::fmin(std::fmax(q, -1e9f), 1e9f));  // Rust `as i32`
    return (uint8_t)std::min(std::max(i, 0), 255);
}

// ------------------------------------------------------------------ à-trous (denoise.rs:526-709)
const int ATROUS_PASSES = 5;
const float SIGMA_LUM = 4.0f;
const float INV_SIGMA_ALBEDO2 = 1.0f / 0.01f;
const float DEMOD_FLOOR = 0.02f;
const float B3[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};

float lum(const float* p) { return 0.2126f * p[0] + 0.7152f * p[1] + 0.0722f * p[2]; }
float clean_radiance(float v) { return std::isfinite(v) ? std::fmax(v, 0.0f) : 0.0f; }     // denoise.rs:168-170
float clean_unit(float v, float lo) { return std::isfinite(v) ? clampf(v, lo, 1.0f) : 0.0f; }  // denoise.rs:173-175

// denoise.rs:541-581
std::vector<float> blur_sep(const std::vector<float>& src, int w, int h, const float* k, int len) {
    int r = len / 2;
    std::vector<float> tmp((size_t)w * h), out((size_t)w * h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.0f, ws = 0.0f;
            for (int i = 0; i < len; i++) {
                int xx = x + i - r;
                if (xx >= 0 && xx < w) {
                    acc += k[i] * src[(size_t)y * w + xx];
                    ws += k[i];
                }
            }
            tmp[(size_t)y * w + x] = acc / ws;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.0f, ws = 0.0f;
            for (int i = 0; i < len; i++) {
                int yy = y + i - r;
                if (yy >= 0 && yy < h) {
                    acc += k[i] * tmp[(size_t)yy * w + x];
                    ws += k[i];
                }
            }
            out[(size_t)y * w + x] = acc / ws;
        }
    return out;
}
// denoise.rs:583-589
std::vector<float> noise_variance(const std::vector<float>& l, int w, int h) {
    const float k3[3] = {1, 1, 1}, k5[5] = {1, 1, 1, 1, 1};
    std::vector<float> mean = blur_sep(l, w, h, k3, 3), resid(l.size());
    for (size_t i = 0; i < l.size(); i++) resid[i] = 1.125f * (l[i] - mean[i]) * (l[i] - mean[i]);
    return blur_sep(resid, w, h, k5, 5);
}

// denoise.rs:660-709 (atrous) with denoise.rs:596-657 (atrous_pass)
void atrous(int w, int h, const float* color, const float* albedo, const float* normal, float* out) {
    const size_t n = (size_t)w * h;
    std::vector<float> demod, ga, gn, sig(3 * n), l(n);
    if (albedo) {
        demod.resize(3 * n);
        ga.resize(3 * n);
        for (size_t i = 0; i < 3 * n; i++) {
            demod[i] = std::fmax(clean_unit(albedo[i], 0.0f), DEMOD_FLOOR);
            ga[i] = clean_unit(albedo[i], 0.0f);
        }
    }
    if (normal) {
        gn.resize(3 * n);
        for (size_t i = 0; i < n; i++) {
            float v[3];
            for (int c = 0; c < 3; c++) v[c] = std::isfinite(normal[3 * i + c]) ? normal[3 * i + c] : 0.0f;
            float len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            for (int c = 0; c < 3; c++) gn[3 * i + c] = len > 1e-6f ? v[c] / len : 0.0f;
        }
    }
    for (size_t i = 0; i < 3 * n; i++) sig[i] = albedo ? clean_radiance(color[i]) / demod[i] : clean_radiance(color[i]);
    for (size_t i = 0; i < n; i++) l[i] = lum(&sig[3 * i]);
    std::vector<float> var = noise_variance(l, w, h);
    const float kv[3] = {0.25f, 0.5f, 0.25f};
    std::vector<float> nsig(3 * n), nvar(n);
    for (int pass = 0; pass < ATROUS_PASSES; pass++) {
        std::vector<float> var_blur = blur_sep(var, w, h, kv, 3);
        const int step = 1 << pass;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const size_t p = (size_t)y * w + x;
                const float lp = l[p];
                const float inv_sl = 1.0f / (SIGMA_LUM * std::sqrt(std::fmax(var_blur[p], 0.0f)) + 1e-10f);
                float acc[3] = {0, 0, 0}, ws = 0.0f, wv = 0.0f;
                for (int ky = 0; ky < 5; ky++) {
                    int qy = y + (ky - 2) * step;
                    if (qy < 0 || qy >= h) continue;
                    for (int kx = 0; kx < 5; kx++) {
                        int qx = x + (kx - 2) * step;
                        if (qx < 0 || qx >= w) continue;
                        const size_t q = (size_t)qy * w + qx;
                        float e = std::fabs(l[q] - lp) * inv_sl;
                        if (albedo) {
                            float d0 = ga[3 * p] - ga[3 * q], d1 = ga[3 * p + 1] - ga[3 * q + 1], d2 = ga[3 * p + 2] - ga[3 * q + 2];
                            e += (d0 * d0 + d1 * d1 + d2 * d2) * INV_SIGMA_ALBEDO2;
                        }
                        float wgt = B3[kx] * B3[ky] * std::exp(-e);
                        if (normal) {
                            const float* np = &gn[3 * p];
                            const float* nq = &gn[3 * q];
                            bool pz = np[0] == 0.0f && np[1] == 0.0f && np[2] == 0.0f;
                            bool qz = nq[0] == 0.0f && nq[1] == 0.0f && nq[2] == 0.0f;
                            if (!pz && !qz) {
                                float d = std::fmax(np[0] * nq[0] + np[1] * nq[1] + np[2] * nq[2], 0.0f);
                                for (int s = 0; s < 7; s++) d = d * d;  // powi(128): seven squarings
                                wgt *= d;
                            }
                        }
                        for (int c = 0; c < 3; c++) acc[c] += wgt * sig[3 * q + c];
                        ws += wgt;
                        wv += wgt * wgt * var[q];
                    }
                }
                for (int c = 0; c < 3; c++) nsig[3 * p + c] = acc[c] / ws;
                nvar[p] = wv / (ws * ws);
            }
        sig.swap(nsig);
        var.swap(nvar);
        for (size_t i = 0; i < n; i++) l[i] = lum(&sig[3 * i]);
    }
    for (size_t i = 0; i < 3 * n; i++) out[i] = albedo ? sig[i] * demod[i] : sig[i];
}

}  // namespace orp

extern "C" int or_tonemap_pixel(const float* rgb, uint32_t n, float ev, int32_t op, float sat, float* out) {
    if (op < 0 || op > 6) return 1;
    for (uint32_t i = 0; i < n; i++) orp::tonemap_pixel(rgb + 3 * (size_t)i, ev, op, sat, out + 3 * (size_t)i);
    return 0;
}
extern "C" int or_develop_encoded(const float* rgb, uint32_t n, float ev, int32_t op, float sat, float* out) {
    if (op < 0 || op > 6) return 1;
    orp::Prepared prep(ev, op, sat);
    for (uint32_t i = 0; i < n; i++) prep.encoded(rgb + 3 * (size_t)i, out + 3 * (size_t)i);
    return 0;
}
extern "C" float or_srgb_encode(float x) { return orp::srgb_encode(x); }
// tonemap.rs:370-388
extern "C" int or_develop(const float* rgb, uint32_t w, uint32_t h, float ev, int32_t op, float sat, uint8_t* out) {
    if (op < 0 || op > 6) return 1;
    orp::Prepared prep(ev, op, sat);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const size_t i = (size_t)y * w + x;
            float s[3], nz[3];
            prep.encoded(rgb + 3 * i, s);
            orp::tpdf3((uint64_t)y * w + x, nz);
            for (int c = 0; c < 3; c++) out[3 * i + c] = orp::quantize8(s[c], nz[c]);
        }
    return 0;
}
extern "C" int or_atrous(uint32_t w, uint32_t h, const float* color, const float* albedo, const float* normal,
                         float* out) {
    if (w == 0 || h == 0) return 0;
    orp::atrous((int)w, (int)h, color, albedo, normal, out);
    return 0;
}

// live.rs:208-218 (relative_change): every 4th pixel, f64 luma sums in pixel order.
extern "C" double or_relative_change(const float* a, const float* b, uint32_t n) {
    double diff = 0.0, base = 0.0;
    for (uint32_t i = 0; i < n; i += 4) {
        const double la = 0.2126 * (double)a[3 * i] + 0.7152 * (double)a[3 * i + 1] + 0.0722 * (double)a[3 * i + 2];
        const double lb = 0.2126 * (double)b[3 * i] + 0.7152 * (double)b[3 * i + 1] + 0.0722 * (double)b[3 * i + 2];
        diff += std::fabs(la - lb);
        base += std::fabs(lb);
    }
    return base <= 1e-9 ? 0.0 : diff / base;
}
