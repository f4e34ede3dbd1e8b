// Preview development on the GPU: the film's running sums -> colour, albedo
// and normal images (render.rs:35-64), the edge-avoiding à-trous denoiser
// (argent-post denoise.rs:526-709), and develop (tonemap.rs:262-388): exposure,
// tone operator, saturation and sRGB encode through per-grade tables built on
// the host, TPDF dither, 8-bit.  One thread per pixel; images are pixel-major
// float4 (xyz used) so a tap is one 16-byte load; every expression keeps the
// reference's operation order (no FMA contraction), so develop is exact and
// the denoiser differs from the reference only through expf.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dev_types.hpp"
#include "post.hpp"

namespace ag {

#define AGD __device__ __forceinline__

// Rust f32::clamp keeps NaN; fminf/fmaxf (f32::min/max) drop it
AGD float clampf(float x, float lo, float hi) { return x < lo ? lo : (x > hi ? hi : x); }

// render.rs:35-64: colour and albedo are sums times 1/samples; the normal is
// renormalised (zero below 1e-6)
__global__ void k_film_images(const float* __restrict__ sums, size_t npix, float k, float4* color, float4* albedo,
                              float4* normal) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const float* c = sums + 3 * i;
        const float* a = sums + 3 * (npix + i);
        const float* n = sums + 3 * (2 * npix + i);
        color[i] = make_float4(c[0] * k, c[1] * k, c[2] * k, 0.0f);
        albedo[i] = make_float4(a[0] * k, a[1] * k, a[2] * k, 0.0f);
        float v0 = n[0] * k, v1 = n[1] * k, v2 = n[2] * k;
        float l = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
        normal[i] = l > 1e-6f ? make_float4(v0 / l, v1 / l, v2 / l, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// ------------------------------------------------------------------ à-trous
constexpr float SIGMA_LUM = 4.0f;
constexpr float INV_SIGMA_ALBEDO2 = 1.0f / 0.01f;
constexpr float DEMOD_FLOOR = 0.02f;
__constant__ float B3[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};

AGD float lum3(float4 p) { return 0.2126f * p.x + 0.7152f * p.y + 0.0722f * p.z; }  // denoise.rs:535-538
AGD float clean_radiance(float v) { return isfinite(v) ? fmaxf(v, 0.0f) : 0.0f; }     // denoise.rs:168-170
AGD float clean_unit(float v) { return isfinite(v) ? clampf(v, 0.0f, 1.0f) : 0.0f; }   // denoise.rs:173-175

// denoise.rs:667-692: demodulation albedo, guides, demodulated signal and its luminance
__global__ void k_atrous_prep(const float4* __restrict__ color, const float4* __restrict__ albedo,
                              const float4* __restrict__ normal, size_t npix, float4* demod, float4* ga, float4* gn,
                              float4* sig, float* l) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        float4 c = color[i];
        float4 s = make_float4(clean_radiance(c.x), clean_radiance(c.y), clean_radiance(c.z), 0.0f);
        if (albedo) {
            float4 a = albedo[i];
            float4 g = make_float4(clean_unit(a.x), clean_unit(a.y), clean_unit(a.z), 0.0f);
            float4 d = make_float4(fmaxf(g.x, DEMOD_FLOOR), fmaxf(g.y, DEMOD_FLOOR), fmaxf(g.z, DEMOD_FLOOR), 0.0f);
            ga[i] = g;
            demod[i] = d;
            s = make_float4(s.x / d.x, s.y / d.y, s.z / d.z, 0.0f);
        }
        if (normal) {
            float4 n = normal[i];
            float v0 = isfinite(n.x) ? n.x : 0.0f, v1 = isfinite(n.y) ? n.y : 0.0f, v2 = isfinite(n.z) ? n.z : 0.0f;
            float len = sqrtf(v0 * v0 + v1 * v1 + v2 * v2);
            gn[i] = len > 1e-6f ? make_float4(v0 / len, v1 / len, v2 / len, 0.0f) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
        sig[i] = s;
        l[i] = lum3(s);
    }
}

// denoise.rs:541-581, one direction of blur_sep: taps in kernel order, borders
// renormalised by the weights that fell inside
template <int LEN>
__global__ void k_blur(const float* __restrict__ src, float* dst, int w, int h, bool vertical, float k0, float k1,
                       float k2, float k3, float k4) {
    const float kk[5] = {k0, k1, k2, k3, k4};
    const int r = LEN / 2;
    const size_t npix = (size_t)w * h;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % (size_t)w), y = (int)(i / (size_t)w);
        float acc = 0.0f, ws = 0.0f;
#pragma unroll
        for (int t = 0; t < LEN; t++) {
            const int q = (vertical ? y : x) + t - r;
            if (q >= 0 && q < (vertical ? h : w)) {
                acc += kk[t] * src[vertical ? (size_t)q * w + x : (size_t)y * w + q];
                ws += kk[t];
            }
        }
        dst[i] = acc / ws;
    }
}

// denoise.rs:583-589: squared residual against the 3x3 mean, x 9/8
__global__ void k_resid(const float* __restrict__ l, const float* __restrict__ mean, float* out, size_t npix) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x)
        out[i] = 1.125f * (l[i] - mean[i]) * (l[i] - mean[i]);
}

// denoise.rs:591-657 (one pass, holes of `step` px), writing the next signal,
// its variance and luminance (denoise.rs:700)
__global__ void k_atrous_pass(const float4* __restrict__ sig, const float* __restrict__ l, const float* __restrict__ var,
                              const float* __restrict__ var_blur, const float4* __restrict__ ga,
                              const float4* __restrict__ gn, int w, int h, int step, float4* out, float* out_var,
                              float* out_l) {
    const size_t npix = (size_t)w * h;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < npix; p += (size_t)gridDim.x * blockDim.x) {
        const int x = (int)(p % (size_t)w), y = (int)(p / (size_t)w);
        const float lp = l[p];
        const float inv_sl = 1.0f / (SIGMA_LUM * sqrtf(fmaxf(var_blur[p], 0.0f)) + 1e-10f);
        const float4 ap = ga ? ga[p] : make_float4(0, 0, 0, 0);
        const float4 np = gn ? gn[p] : make_float4(0, 0, 0, 0);
        const bool np_zero = np.x == 0.0f && np.y == 0.0f && np.z == 0.0f;
        float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, ws = 0.0f, wv = 0.0f;
        for (int ky = 0; ky < 5; ky++) {
            const int qy = y + (ky - 2) * step;
            if (qy < 0 || qy >= h) continue;
            const float hy = B3[ky];
            for (int kx = 0; kx < 5; kx++) {
                const int qx = x + (kx - 2) * step;
                if (qx < 0 || qx >= w) continue;
                const size_t q = (size_t)qy * w + qx;
                float e = fabsf(l[q] - lp) * inv_sl;
                if (ga) {
                    const float4 aq = ga[q];
                    const float d0 = ap.x - aq.x, d1 = ap.y - aq.y, d2 = ap.z - aq.z;
                    e += (d0 * d0 + d1 * d1 + d2 * d2) * INV_SIGMA_ALBEDO2;
                }
                float wgt = B3[kx] * hy * expf(-e);
                if (gn) {
                    const float4 nq = gn[q];
                    if (!np_zero && !(nq.x == 0.0f && nq.y == 0.0f && nq.z == 0.0f)) {
                        float d = fmaxf(np.x * nq.x + np.y * nq.y + np.z * nq.z, 0.0f);
#pragma unroll
                        for (int s = 0; s < 7; s++) d = d * d;  // powi(128): seven squarings
                        wgt *= d;
                    }
                }
                const float4 s = sig[q];
                a0 += wgt * s.x;
                a1 += wgt * s.y;
                a2 += wgt * s.z;
                ws += wgt;
                wv += wgt * wgt * var[q];
            }
        }
        const float4 o = make_float4(a0 / ws, a1 / ws, a2 / ws, 0.0f);
        out[p] = o;
        out_var[p] = wv / (ws * ws);
        out_l[p] = lum3(o);
    }
}

// denoise.rs:702-707: remodulate
__global__ void k_remod(const float4* __restrict__ sig, const float4* __restrict__ demod, size_t npix, float4* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        float4 s = sig[i];
        if (demod) {
            float4 d = demod[i];
            s = make_float4(s.x * d.x, s.y * d.y, s.z * d.z, 0.0f);
        }
        out[i] = s;
    }
}

// ------------------------------------------------------------------ develop
AGD void mat3(const float* m, float v0, float v1, float v2, float o[3]) {
    for (int r = 0; r < 3; r++) o[r] = m[3 * r] * v0 + m[3 * r + 1] * v1 + m[3 * r + 2] * v2;
}
AGD void saturate3(float v[3], float sat) {  // tonemap.rs:138-145
    if (sat == 1.0f) {
        for (int c = 0; c < 3; c++) v[c] = clampf(v[c], 0.0f, 1.0f);
        return;
    }
    const float l = 0.2126f * v[0] + 0.7152f * v[1] + 0.0722f * v[2];
    for (int c = 0; c < 3; c++) v[c] = clampf(l + sat * (v[c] - l), 0.0f, 1.0f);
}
AGD float sanitize(float v) { return isnan(v) ? 0.0f : clampf(v, 0.0f, 65504.0f); }  // tonemap.rs:197-200
// tonemap.rs:250-258 (BitCurve::eval)
AGD float curve_eval(const float* t, uint32_t base, uint32_t len, float v) {
    const float lo = __uint_as_float(base << DEV_CURVE_SHIFT);
    const float hi = __uint_as_float((base + len - 1u) << DEV_CURVE_SHIFT);
    const uint32_t bits = __float_as_uint(fminf(fmaxf(v, lo), hi));
    uint32_t k = (bits >> DEV_CURVE_SHIFT) - base;
    const float f = (float)(bits & ((1u << DEV_CURVE_SHIFT) - 1u)) * (1.0f / (float)(1u << DEV_CURVE_SHIFT));
    k = min(k, len - 2u);
    return t[k] * (1.0f - f) + t[k + 1] * f;
}
// color.rs:48-57 (Lut1::eval)
AGD float lut_eval(const float* t, float lo, float scale, uint32_t last, float x) {
    float u = (x - lo) * scale;
    u = u >= 0.0f ? fminf(u, (float)last) : 0.0f;
    const uint32_t i = min((uint32_t)u, last - 1u);
    const float f = u - (float)i;
    return t[i] * (1.0f - f) + t[i + 1] * f;
}
// tonemap.rs:162-176
AGD void aces_fitted(const float* in_m, const float* out_m, const float x[3], float o[3]) {
    float v[3];
    mat3(in_m, x[0], x[1], x[2], v);
    for (int c = 0; c < 3; c++) {
        const float a = v[c] * (v[c] + 0.0245786f) - 0.000090537f;
        const float b = v[c] * (0.983729f * v[c] + 0.432951f) + 0.238081f;
        v[c] = a / b;
    }
    mat3(out_m, v[0], v[1], v[2], o);
    for (int c = 0; c < 3; c++) o[c] = clampf(o[c], 0.0f, 1.0f);
}
// tonemap.rs:179-194
AGD void khronos_neutral(const float in[3], float c[3]) {
    const float START = 0.8f - 0.04f, DESAT = 0.15f;
    for (int k = 0; k < 3; k++) c[k] = in[k];
    const float x = fminf(fminf(c[0], c[1]), c[2]);
    const float offset = x < 0.08f ? x - 6.25f * x * x : 0.04f;
    for (int k = 0; k < 3; k++) c[k] = c[k] - offset;
    const float peak = fmaxf(fmaxf(c[0], c[1]), c[2]);
    if (peak < START) return;
    const float d = 1.0f - START;
    const float new_peak = 1.0f - d * d / (peak + d - START);
    for (int k = 0; k < 3; k++) c[k] = c[k] * (new_peak / peak);
    const float g = 1.0f - 1.0f / (DESAT * (peak - new_peak) + 1.0f);
    for (int k = 0; k < 3; k++) c[k] = c[k] + (new_peak - c[k]) * g;
}
// tonemap.rs:341-367
AGD uint64_t hash64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}
AGD uint8_t quantize8(float v, float n) {
    if (v <= 0.0f) return 0;
    if (v >= 1.0f) return 255;
    const float q = v * 255.0f + 0.5f + n;
    const int i = isnan(q) ? 0 : (int)truncf(fminf(fmaxf(q, -1e9f), 1e9f));  // Rust `as i32`
    return (uint8_t)min(max(i, 0), 255);
}

// tonemap.rs:316-338 (Prepared::encoded) + :377-386 (develop's row loop)
__global__ void k_develop(const float4* __restrict__ img, int w, int h, DevGrade g, uint8_t* out) {
    const size_t npix = (size_t)w * h;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const float4 p = img[i];
        const float x[3] = {sanitize(p.x * g.k), sanitize(p.y * g.k), sanitize(p.z * g.k)};
        float s[3];
        if (g.op <= 2) {
            float v[3], e[3], o[3];
            mat3(g.inset, x[0], x[1], x[2], v);
            for (int c = 0; c < 3; c++) e[c] = curve_eval(g.curve[g.pick[c]], g.curve_base, g.curve_len, v[c]);
            if (g.sat != 1.0f) saturate3(e, g.sat);
            mat3(g.outset, e[0], e[1], e[2], o);
            for (int c = 0; c < 3; c++) s[c] = lut_eval(g.encode, g.enc_lo, g.enc_scale, g.enc_last, o[c]);
        } else {
            float d[3];
            if (g.op == 3) {
                aces_fitted(g.inset, g.outset, x, d);
            } else if (g.op == 4) {
                khronos_neutral(x, d);
            } else if (g.op == 5) {
                for (int c = 0; c < 3; c++) d[c] = x[c] / (1.0f + x[c]);
            } else {
                for (int c = 0; c < 3; c++) d[c] = x[c];
            }
            saturate3(d, g.sat);
            for (int c = 0; c < 3; c++) s[c] = lut_eval(g.encode, g.enc_lo, g.enc_scale, g.enc_last, d[c]);
        }
        const uint64_t hh = hash64((uint64_t)i ^ 0x9e3779b97f4a7c15ull);  // pixel index y * w + x
        const float S = 1.0f / 1024.0f;
        for (int c = 0; c < 3; c++) {
            const float u0 = (float)((hh >> (2 * c * 10)) & 0x3ffu), u1 = (float)((hh >> ((2 * c + 1) * 10)) & 0x3ffu);
            out[3 * i + c] = quantize8(s[c], (u0 + u1 + 1.0f) * S - 1.0f);
        }
    }
}

__global__ void k_pack4(const float* __restrict__ rgb, size_t npix, float4* out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x)
        out[i] = make_float4(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2], 0.0f);
}
__global__ void k_unpack4(const float4* __restrict__ img, size_t npix, float* rgb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix; i += (size_t)gridDim.x * blockDim.x) {
        const float4 p = img[i];
        rgb[3 * i] = p.x;
        rgb[3 * i + 1] = p.y;
        rgb[3 * i + 2] = p.z;
    }
}

// ------------------------------------------------------------------ launchers
static int grid_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

void launch_pack4(hipStream_t s, const float* rgb, size_t npix, float4* out) {
    hipLaunchKernelGGL(k_pack4, dim3(grid_for(npix)), dim3(256), 0, s, rgb, npix, out);
}
void launch_unpack4(hipStream_t s, const float4* img, size_t npix, float* rgb) {
    hipLaunchKernelGGL(k_unpack4, dim3(grid_for(npix)), dim3(256), 0, s, img, npix, rgb);
}

void launch_film_images(hipStream_t s, const float* sums, size_t npix, float k, float4* color, float4* albedo,
                        float4* normal) {
    hipLaunchKernelGGL(k_film_images, dim3(grid_for(npix)), dim3(256), 0, s, sums, npix, k, color, albedo, normal);
}
void launch_atrous(hipStream_t s, int w, int h, const float4* color, const float4* albedo, const float4* normal,
                   const AtrousWork& W, float4* out) {
    const size_t n = (size_t)w * h;
    const int g = grid_for(n);
    hipLaunchKernelGGL(k_atrous_prep, dim3(g), dim3(256), 0, s, color, albedo, normal, n, W.demod, W.ga, W.gn, W.sig[0],
                       W.l[0]);
    // noise_variance: 3x3 mean, residual, 5x5 box
    hipLaunchKernelGGL(k_blur<3>, dim3(g), dim3(256), 0, s, W.l[0], W.t0, w, h, false, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f);
    hipLaunchKernelGGL(k_blur<3>, dim3(g), dim3(256), 0, s, W.t0, W.t1, w, h, true, 1.0f, 1.0f, 1.0f, 0.0f, 0.0f);
    hipLaunchKernelGGL(k_resid, dim3(g), dim3(256), 0, s, W.l[0], W.t1, W.t0, n);
    hipLaunchKernelGGL(k_blur<5>, dim3(g), dim3(256), 0, s, W.t0, W.t1, w, h, false, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f);
    hipLaunchKernelGGL(k_blur<5>, dim3(g), dim3(256), 0, s, W.t1, W.var[0], w, h, true, 1.0f, 1.0f, 1.0f, 1.0f, 1.0f);
    int cur = 0;
    for (int pass = 0; pass < 5; pass++) {
        hipLaunchKernelGGL(k_blur<3>, dim3(g), dim3(256), 0, s, W.var[cur], W.t0, w, h, false, 0.25f, 0.5f, 0.25f, 0.0f,
                           0.0f);
        hipLaunchKernelGGL(k_blur<3>, dim3(g), dim3(256), 0, s, W.t0, W.t1, w, h, true, 0.25f, 0.5f, 0.25f, 0.0f, 0.0f);
        hipLaunchKernelGGL(k_atrous_pass, dim3(g), dim3(256), 0, s, W.sig[cur], W.l[cur], W.var[cur], W.t1,
                           albedo ? W.ga : nullptr, normal ? W.gn : nullptr, w, h, 1 << pass, W.sig[1 - cur],
                           W.var[1 - cur], W.l[1 - cur]);
        cur = 1 - cur;
    }
    hipLaunchKernelGGL(k_remod, dim3(g), dim3(256), 0, s, W.sig[cur], albedo ? W.demod : nullptr, n, out);
}
// live.rs:208-218 (relative_change): over every 4th pixel, the sums of
// |lum(a) - lum(b)| and |lum(b)| in f64, lum = Rec. 709 luma of f32 RGB.
// Each block writes its two partial sums; k_relchange_fold adds them.
constexpr int RELCHANGE_BLOCKS = 512;

__device__ inline void block_sum2(double a, double b, double* out2) {
    __shared__ double red[256 / 64][2];
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        b += __shfl_xor(b, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][0] = a;
        red[threadIdx.x >> 6][1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
        for (int w = 0; w < 256 / 64; w++) {
            sa += red[w][0];
            sb += red[w][1];
        }
        out2[0] = sa;
        out2[1] = sb;
    }
}

__global__ void k_relchange(const float4* __restrict__ a, const float4* __restrict__ b, size_t npix, double* part) {
    double diff = 0.0, base = 0.0;
    for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; 4 * q < npix; q += (size_t)gridDim.x * blockDim.x) {
        const float4 pa = a[4 * q], pb = b[4 * q];
        const double la = 0.2126 * (double)pa.x + 0.7152 * (double)pa.y + 0.0722 * (double)pa.z;
        const double lb = 0.2126 * (double)pb.x + 0.7152 * (double)pb.y + 0.0722 * (double)pb.z;
        diff += fabs(la - lb);
        base += fabs(lb);
    }
    block_sum2(diff, base, part + 2 * blockIdx.x);
}

__global__ void k_relchange_fold(const double* part, int nb, double* out2) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    block_sum2(a, b, out2);
}

void launch_relative_change(hipStream_t s, const float4* a, const float4* b, size_t npix, double* part, double* out2) {
    const int nb = std::max(1, std::min(RELCHANGE_BLOCKS, (int)((npix / 4 + 255) / 256)));
    hipLaunchKernelGGL(k_relchange, dim3(nb), dim3(256), 0, s, a, b, npix, part);
    hipLaunchKernelGGL(k_relchange_fold, dim3(1), dim3(256), 0, s, part, nb, out2);
}

void launch_develop(hipStream_t s, const float4* img, int w, int h, const DevGrade& g, uint8_t* out) {
    hipLaunchKernelGGL(k_develop, dim3(grid_for((size_t)w * h)), dim3(256), 0, s, img, w, h, g, out);
}

}  // namespace ag
