// Host side of preview development: the tables of a grade (argent-post
// tonemap.rs:284-313, Prepared::new, with BitCurve::new tonemap.rs:239-247
// and Lut1::new color.rs:41-46), built in f32 with the host libm (powf,
// log2f) as the reference builds them, so the device's develop reproduces the
// reference's bytes.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "host_internal.hpp"
#include "post.hpp"

namespace ag {

namespace {

float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }
uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
float u2f(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// color.rs:11-19 (srgb_encode)
float srgb_encode(float x) {
    if (x <= 0.0031308f) return 12.92f * std::fmax(x, 0.0f);
    if (x >= 1.0f) return 1.0f;
    return 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
}
constexpr float MIN_EV = -12.47393f, MAX_EV = 4.026069f;  // tonemap.rs:87-88
// tonemap.rs:118-122
float log_encode(float v) {
    float ev = clampf(std::log2(std::fmax(v, 1e-10f)), MIN_EV, MAX_EV);
    return (ev - MIN_EV) / (MAX_EV - MIN_EV);
}
// tonemap.rs:125-128
float sigmoid(float x) {
    return ((((((15.5f * x - 40.14f) * x + 31.96f) * x - 6.868f) * x + 0.4298f) * x + 0.1191f) * x) - 0.00232f;
}
// tonemap.rs:131-135
float sop(float x, float slope, float offset, float power) {
    float v = clampf(x * slope + offset, 0.0f, 1.0f);
    return power == 1.0f ? v : std::pow(v, power);
}

const float INSET[9] = {0.84247906f, 0.0784336f, 0.079223745f, 0.042328242f, 0.8784686f,
                        0.07916613f, 0.042375655f, 0.0784336f, 0.879143f};
const float OUTSET[9] = {1.196879f, -0.09802088f, -0.09902974f, -0.05289685f, 1.1519031f,
                         -0.09896118f, -0.052971635f, -0.09804345f, 1.1510737f};
const float ACES_IN[9] = {0.59719f, 0.35458f, 0.04823f, 0.07600f, 0.90834f, 0.01566f, 0.02840f, 0.13383f, 0.83777f};
const float ACES_OUT[9] = {1.60475f, -0.53108f, -0.07367f, -0.10208f, 1.10813f, -0.00605f, -0.00327f, -0.07276f, 1.07602f};

}  // namespace

// tonemap.rs:284-313.  The blob holds the curves (curve_len floats each) then
// the encode table; the caller uploads it and sets the DevGrade pointers.
int build_grade(float exposure_ev, int32_t op, float saturation, std::vector<float>* blob, DevGrade* g, int* n_curves,
                size_t* curve_len, size_t* encode_len) {
    if (op < 0 || op > 6) return fail(AG_E_INVALID, "unknown tone operator");
    DevGrade d{};
    d.op = op;
    d.k = std::isfinite(exposure_ev) ? std::pow(2.0f, exposure_ev) : 1.0f;           // tonemap.rs:202-205
    const float sat = std::isfinite(saturation) ? std::fmax(saturation, 0.0f) : 1.0f;  // tonemap.rs:212
    blob->clear();
    *n_curves = 0;
    *curve_len = 0;
    if (op <= 2) {
        // tonemap.rs:96-102: Base, Punchy, Golden
        const float slope[3][3] = {{1, 1, 1}, {1, 1, 1}, {1.0f, 0.9f, 0.5f}};
        const float power[3][3] = {{1, 1, 1}, {1.35f, 1.35f, 1.35f}, {0.8f, 0.8f, 0.8f}};
        const float look_sat[3] = {1.0f, 1.4f, 1.3f};
        const bool shared = slope[op][1] == slope[op][0] && slope[op][2] == slope[op][0] &&
                            power[op][1] == power[op][0] && power[op][2] == power[op][0];
        const int n = shared ? 1 : 3;
        // BitCurve::new: segments from the bit pattern of lo to that of hi, +1
        const float lo = std::pow(2.0f, MIN_EV) * 0.999f, hi = std::pow(2.0f, MAX_EV) * 1.001f;
        const uint32_t base = f2u(lo) >> DEV_CURVE_SHIFT, top = (f2u(hi) >> DEV_CURVE_SHIFT) + 1;
        const size_t len = top - base + 1;
        for (int c = 0; c < n; c++)
            for (uint32_t k = base; k <= top; k++)
                blob->push_back(sop(sigmoid(log_encode(u2f(k << DEV_CURVE_SHIFT))), slope[op][c], 0.0f, power[op][c]));
        d.curve_base = base;
        d.curve_len = (uint32_t)len;
        for (int c = 0; c < 3; c++) d.pick[c] = shared ? 0 : c;
        d.sat = look_sat[op] * sat;
        *n_curves = n;
        *curve_len = len;
        std::memcpy(d.inset, INSET, sizeof INSET);
        std::memcpy(d.outset, OUTSET, sizeof OUTSET);
    } else {
        d.sat = sat;
        if (op == 3) {
            std::memcpy(d.inset, ACES_IN, sizeof ACES_IN);
            std::memcpy(d.outset, ACES_OUT, sizeof ACES_OUT);
        }
    }
    // Lut1::new(lo, hi, 4096, f): entries f(lo + (hi - lo) * i / n), i = 0..=n
    const float elo = 0.0f, ehi = op <= 2 ? 1.25f : 1.0f;  // AGX_OUT_MAX, tonemap.rs:225
    const int n = 4096;
    for (int i = 0; i <= n; i++) {
        const float x = elo + (ehi - elo) * (float)i / (float)n;
        blob->push_back(op <= 2 ? srgb_encode(std::fmin(std::pow(std::fmax(x, 0.0f), 2.2f), 1.0f)) : srgb_encode(x));
    }
    d.enc_lo = elo;
    d.enc_scale = (float)n / (ehi - elo);
    d.enc_last = (uint32_t)n;
    *encode_len = (size_t)n + 1;
    *g = d;
    return AG_OK;
}

}  // namespace ag
