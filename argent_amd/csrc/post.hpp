// Preview development: device-side grade and work buffers, and the launchers
// of post.hip.  The grade's tables are built on the host (host_post.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace ag {

constexpr uint32_t DEV_CURVE_SHIFT = 23 - 9;  // tonemap.rs:236-237: 9 mantissa bits per segment

// A grade ready for the device (tonemap.rs:262-338, Prepared): exposure
// scale, operator, saturation, the AgX channel curves (one shared or three)
// and the encode table, with the 3x3 matrices of the operator.
struct DevGrade {
    float k;        // 2^exposure_ev
    int32_t op;     // 0-2 AgX Base/Punchy/Golden, 3 ACES fitted, 4 Khronos Neutral, 5 Reinhard, 6 Clamp
    float sat;      // AgX: look saturation x grade saturation; others: grade saturation
    const float* curve[3];
    int32_t pick[3];
    uint32_t curve_base, curve_len;
    const float* encode;
    float enc_lo, enc_scale;
    uint32_t enc_last;
    float inset[9], outset[9];  // AgX inset/outset, or ACES in/out
};

// Work buffers of one à-trous run over w*h pixels.
struct AtrousWork {
    float4 *demod, *ga, *gn, *sig[2];
    float *l[2], *var[2], *t0, *t1;
};

void launch_pack4(hipStream_t s, const float* rgb, size_t npix, float4* out);    // (r, g, b) -> float4
void launch_unpack4(hipStream_t s, const float4* img, size_t npix, float* rgb);  // float4 -> (r, g, b)
void launch_film_images(hipStream_t s, const float* sums, size_t npix, float k, float4* color, float4* albedo,
                        float4* normal);
void launch_atrous(hipStream_t s, int w, int h, const float4* color, const float4* albedo, const float4* normal,
                   const AtrousWork& W, float4* out);
void launch_develop(hipStream_t s, const float4* img, int w, int h, const DevGrade& g, uint8_t* out);
// out2 = (sum |lum a - lum b|, sum |lum b|) over every 4th pixel, f64;
// part holds 2 doubles per block (at most 512 blocks).
void launch_relative_change(hipStream_t s, const float4* a, const float4* b, size_t npix, double* part, double* out2);

}  // namespace ag
