// Host-side spectral library (scene build).  Each function cites the
// reference item whose behaviour it reproduces; the measured tables come from
// tables/spectral_data.h.
#include "host_spectra.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <mutex>

#include "tables/spectral_data.h"

namespace ag {

static const float PI_F = 3.14159265358979323846f;

// ------------------------------------------------------------------ CIE
// cie.rs:38-49: linear interpolation of the 1 nm table, zero outside.
std::array<float, 3> xyz_bar(float lambda) {
    float x = lambda - 360.0f;
    if (!(x >= 0.0f && x <= 470.0f)) return {0.0f, 0.0f, 0.0f};
    int i = std::min((int)x, 469);
    float t = x - (float)i;
    const float* a = &ag_data::CMF_1931_2DEG[i * 3];
    const float* b = &ag_data::CMF_1931_2DEG[(i + 1) * 3];
    return {a[0] + (b[0] - a[0]) * t, a[1] + (b[1] - a[1]) * t, a[2] + (b[2] - a[2]) * t};
}

std::array<float, 3> xyz_to_linear_srgb(std::array<float, 3> v) {
    float x = v[0], y = v[1], z = v[2];
    return {3.2404542f * x - 1.5371385f * y - 0.4985314f * z,
            -0.969266f * x + 1.8760108f * y + 0.041556f * z,
            0.0556434f * x - 0.2040259f * y + 1.0572252f * z};
}

std::array<float, 3> linear_srgb_to_xyz(std::array<float, 3> v) {
    float r = v[0], g = v[1], b = v[2];
    return {0.4124564f * r + 0.3575761f * g + 0.1804375f * b,
            0.2126729f * r + 0.7151522f * g + 0.072175f * b,
            0.0193339f * r + 0.119192f * g + 0.9503041f * b};
}

// spectrum.rs:80-93
float Piecewise::eval(float l) const {
    size_t n = lambda.size();
    if (l <= lambda[0]) return value[0];
    if (l >= lambda[n - 1]) return value[n - 1];
    size_t i = (size_t)(std::upper_bound(lambda.begin(), lambda.end(), l) - lambda.begin()) - 1;
    float t = (l - lambda[i]) / (lambda[i + 1] - lambda[i]);
    return value[i] * (1.0f - t) + value[i + 1] * t;
}

static const Piecewise& d65_table() {
    static Piecewise p = [] {
        Piecewise q;
        for (int i = 0; i < 107; i++) {
            q.lambda.push_back(300.0f + 5.0f * (float)i);
            q.value.push_back(ag_data::D65_300_830_5NM[i]);
        }
        return q;
    }();
    return p;
}

float d65(float l) { return d65_table().eval(l); }

// spectrum.rs:413-420
float luminance_y(const SpecFn& s) {
    double y = 0.0;
    for (int i = 0; i <= 470; i++) {
        float l = 360.0f + (float)i;
        y += (double)s(l) * (double)xyz_bar(l)[1];
    }
    return (float)(y / (double)CIE_Y_INTEGRAL);
}

// cie.rs:107-113: the D65 table scaled (in f32) by 1 / luminance.
static const Piecewise& d65n_table() {
    static Piecewise p = [] {
        Piecewise q = d65_table();
        float y = luminance_y([](float l) { return d65(l); });
        float s = 1.0f / y;
        for (auto& v : q.value) v *= s;
        return q;
    }();
    return p;
}

float d65_normalized(float l) { return d65n_table().eval(l); }

// cie.rs:132-152
static const std::vector<std::array<double, 3>>& d65_cmf_weights() {
    static std::vector<std::array<double, 3>> w = [] {
        std::vector<std::array<double, 3>> out(471);
        double ny = 0.0;
        for (int i = 0; i < 471; i++) {
            double e = (double)d65(360.0f + (float)i);
            for (int k = 0; k < 3; k++) out[i][k] = e * (double)ag_data::CMF_1931_2DEG[i * 3 + k];
            ny += out[i][1];
        }
        for (auto& r : out)
            for (auto& v : r) v /= ny;
        return out;
    }();
    return w;
}

std::array<float, 3> reflectance_xyz(const SpecFn& r) {
    const auto& w = d65_cmf_weights();
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < 471; i++) {
        double v = (double)r(360.0f + (float)i);
        for (int k = 0; k < 3; k++) acc[k] += v * w[i][k];
    }
    return {(float)acc[0], (float)acc[1], (float)acc[2]};
}

std::array<float, 3> reflectance_srgb(const SpecFn& r) { return xyz_to_linear_srgb(reflectance_xyz(r)); }

std::array<float, 3> emission_xyz(const SpecFn& s) {
    double acc[3] = {0, 0, 0};
    for (int i = 0; i < 471; i++) {
        double v = (double)s(360.0f + (float)i);
        for (int k = 0; k < 3; k++) acc[k] += v * (double)ag_data::CMF_1931_2DEG[i * 3 + k];
    }
    double n = (double)CIE_Y_INTEGRAL;
    return {(float)(acc[0] / n), (float)(acc[1] / n), (float)(acc[2] / n)};
}

// cie.rs:176-185
std::array<float, 3> xyz_to_lab(std::array<float, 3> xyz) {
    static std::array<float, 3> w = reflectance_xyz([](float) { return 1.0f; });
    auto f = [](float t) {
        const float D = 6.0f / 29.0f;
        return t > D * D * D ? std::cbrt(t) : t / (3.0f * D * D) + 4.0f / 29.0f;
    };
    float fx = f(xyz[0] / w[0]), fy = f(xyz[1] / w[1]), fz = f(xyz[2] / w[2]);
    return {116.0f * fy - 16.0f, 500.0f * (fx - fy), 200.0f * (fy - fz)};
}

// ------------------------------------------------------------------ Planck
// spectrum.rs:126-136 (f64 Planck, returned as f32)
float planck(float lambda_nm, float t) {
    if (t <= 0.0f) return 0.0f;
    const double c = 299792458.0, h = 6.62606957e-34, kb = 1.3806488e-23;
    double l = (double)lambda_nm * 1e-9;
    double l2 = l * l;
    double l5 = l * (l2 * l2);  // powi(5) lowered by repeated squaring
    double le = (2.0 * h * c * c) / (l5 * (std::exp((h * c) / (l * kb * (double)t)) - 1.0));
    return (float)le;
}

Blackbody Blackbody::peak_normalized(float kelvin) {
    float peak = 2.897772e-3f / kelvin * 1e9f;
    return Blackbody{kelvin, 1.0f / planck(peak, kelvin)};
}

Blackbody Blackbody::with_luminance(float kelvin, float y) {
    Blackbody raw{kelvin, 1.0f};
    float l = luminance_y([&](float x) { return raw.eval(x); });
    return Blackbody{kelvin, (l > 0.0f && std::isfinite(l)) ? y / l : 0.0f};
}

float Blackbody::eval(float l) const { return planck(l, kelvin) * norm; }

// ------------------------------------------------------------------ sigmoid albedo
// spectrum.rs:145-375: Jakob & Hanika's sigmoid solved by Newton with a
// continuation on 5 nm bins, polished on the 1 nm quadrature.
namespace {
struct Quadrature {
    std::vector<double> t;
    std::vector<std::array<double, 3>> w;
};

const Quadrature& fine_quadrature() {
    static Quadrature q = [] {
        const double M[3][3] = {{3.2404542, -1.5371385, -0.4985314},
                                {-0.969266, 1.8760108, 0.041556},
                                {0.0556434, -0.2040259, 1.0572252}};
        Quadrature o;
        const auto& xyz = d65_cmf_weights();
        for (size_t j = 0; j < xyz.size(); j++) {
            const auto& v = xyz[j];
            std::array<double, 3> r;
            for (int k = 0; k < 3; k++) r[k] = M[k][0] * v[0] + M[k][1] * v[1] + M[k][2] * v[2];
            o.w.push_back(r);
            o.t.push_back((double)j / (double)470.0f);
        }
        return o;
    }();
    return q;
}

const Quadrature& coarse_quadrature() {
    static Quadrature q = [] {
        const Quadrature& f = fine_quadrature();
        size_t bins = (f.t.size() - 1) / 5 + 1;
        Quadrature o;
        o.w.assign(bins, {0.0, 0.0, 0.0});
        for (size_t i = 0; i < f.w.size(); i++) {
            size_t b = std::min((i + 2) / 5, bins - 1);
            for (int k = 0; k < 3; k++) o.w[b][k] += f.w[i][k];
        }
        for (size_t b = 0; b < bins; b++) o.t.push_back((double)(5 * b) / (double)470.0f);
        return o;
    }();
    return q;
}

std::array<double, 3> sigmoid_rgb(const double c[3], const Quadrature& q) {
    std::array<double, 3> acc = {0, 0, 0};
    for (size_t j = 0; j < q.t.size(); j++) {
        double t = q.t[j];
        double x = (c[0] * t + c[1]) * t + c[2];
        double s = 0.5 + 0.5 * x / std::sqrt(1.0 + x * x);
        for (int k = 0; k < 3; k++) acc[k] += s * q.w[j][k];
    }
    return acc;
}

bool solve3(double a[3][3], double b[3], double x[3]) {
    for (int col = 0; col < 3; col++) {
        int piv = col;
        for (int i = col + 1; i < 3; i++)
            if (std::fabs(a[i][col]) > std::fabs(a[piv][col])) piv = i;
        if (std::fabs(a[piv][col]) < 1e-300) return false;
        for (int k = 0; k < 3; k++) std::swap(a[col][k], a[piv][k]);
        std::swap(b[col], b[piv]);
        for (int r = col + 1; r < 3; r++) {
            double f = a[r][col] / a[col][col];
            for (int k = col; k < 3; k++) a[r][k] -= f * a[col][k];
            b[r] -= f * b[col];
        }
    }
    for (int r = 2; r >= 0; r--) {
        double s = 0.0;
        for (int k = r + 1; k < 3; k++) s += a[r][k] * x[k];
        x[r] = (b[r] - s) / a[r][r];
    }
    return std::isfinite(x[0]) && std::isfinite(x[1]) && std::isfinite(x[2]);
}

void newton(double c[3], const double goal[3], const Quadrature& q, int max_iter, double tol) {
    for (int it = 0; it < max_iter; it++) {
        double f[3] = {-goal[0], -goal[1], -goal[2]};
        double jac[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (size_t j = 0; j < q.t.size(); j++) {
            double t = q.t[j];
            double x = (c[0] * t + c[1]) * t + c[2];
            double r = 1.0 / std::sqrt(1.0 + x * x);
            double s = 0.5 + 0.5 * x * r;
            double d = 0.5 * r * r * r;
            double dx[3] = {d * t * t, d * t, d};
            for (int k = 0; k < 3; k++) {
                f[k] += s * q.w[j][k];
                for (int m = 0; m < 3; m++) jac[k][m] += dx[m] * q.w[j][k];
            }
        }
        double f2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
        if (f2 < tol) return;
        double rhs[3] = {-f[0], -f[1], -f[2]};
        double step[3];
        if (!solve3(jac, rhs, step)) return;
        double alpha = 1.0;
        for (;;) {
            double cn[3] = {c[0] + alpha * step[0], c[1] + alpha * step[1], c[2] + alpha * step[2]};
            auto r = sigmoid_rgb(cn, q);
            double e[3] = {r[0] - goal[0], r[1] - goal[1], r[2] - goal[2]};
            if (e[0] * e[0] + e[1] * e[1] + e[2] * e[2] < f2) {
                c[0] = cn[0];
                c[1] = cn[1];
                c[2] = cn[2];
                break;
            }
            alpha *= 0.5;
            if (alpha < 1e-6) return;
        }
    }
}

bool solve_sigmoid(std::array<float, 3> target, float out[3]) {
    const Quadrature& coarse = coarse_quadrature();
    const Quadrature& fine = fine_quadrature();
    double tg[3] = {(double)target[0], (double)target[1], (double)target[2]};
    double g0 = std::clamp((tg[0] + tg[1] + tg[2]) / 3.0, 0.02, 0.98);
    double u = 2.0 * g0 - 1.0;
    double c[3] = {0.0, 0.0, u / std::sqrt(1.0 - u * u)};
    auto start = sigmoid_rgb(c, coarse);
    const int STEPS = 8;
    for (int k = 1; k <= STEPS; k++) {
        double a = (double)k / (double)STEPS;
        double goal[3];
        for (int i = 0; i < 3; i++) goal[i] = start[i] + (tg[i] - start[i]) * a;
        newton(c, goal, coarse, 30, 1e-10);
    }
    newton(c, tg, fine, 30, 1e-16);
    if (!(std::isfinite(c[0]) && std::isfinite(c[1]) && std::isfinite(c[2]))) return false;
    for (int i = 0; i < 3; i++) out[i] = (float)c[i];
    return true;
}

inline float sigmoid(float x) {
    if (std::fabs(x) > 1e8f) return x > 0.0f ? 1.0f : 0.0f;
    return 0.5f + 0.5f * x / std::sqrt(1.0f + x * x);
}
}  // namespace

RgbAlbedo RgbAlbedo::make(std::array<float, 3> rgb) {
    for (auto& v : rgb) v = std::isnan(v) ? 0.0f : std::clamp(v, 0.0f, 1.0f);
    RgbAlbedo a{};
    if (rgb[0] == rgb[1] && rgb[1] == rgb[2]) {
        a.constant = true;
        a.value = rgb[0];
        return a;
    }
    if (solve_sigmoid(rgb, a.c)) {
        a.constant = false;
    } else {
        a.constant = true;
        a.value = (rgb[0] + rgb[1] + rgb[2]) / 3.0f;
    }
    return a;
}

float RgbAlbedo::eval(float l) const {
    if (constant) return value;
    float t = (l - 360.0f) * (1.0f / 470.0f);
    return sigmoid((c[0] * t + c[1]) * t + c[2]);
}

RgbIlluminant RgbIlluminant::make(std::array<float, 3> rgb) {
    for (auto& v : rgb) v = std::isfinite(v) ? std::max(v, 0.0f) : 0.0f;
    float m = 0.0f;
    for (float v : rgb) m = std::fmax(m, v);
    RgbIlluminant r;
    r.scale = 2.0f * m;
    if (m > 0.0f)
        r.albedo = RgbAlbedo::make({rgb[0] / r.scale, rgb[1] / r.scale, rgb[2] / r.scale});
    else
        r.albedo = RgbAlbedo::make({0.0f, 0.0f, 0.0f});
    return r;
}

float RgbIlluminant::eval(float l) const { return scale * albedo.eval(l) * d65_normalized(l); }

// ------------------------------------------------------------------ tables
void bake(const SpecFn& s, float* out) {
    for (int i = 0; i < SPEC_N; i++) out[i] = s(360.0f + (float)i);
}

float tab_eval(const float* v, float l) {
    float x = std::clamp(l - 360.0f, 0.0f, (float)(SPEC_N - 1));
    int i = std::min((int)x, SPEC_N - 2);
    float f = x - (float)i;
    return v[i] * (1.0f - f) + v[i + 1] * f;
}

// ------------------------------------------------------------------ metals
ComplexIor ior_from_grid(const float* nk) {
    ComplexIor m;
    for (int i = 0; i < GRID_N; i++) {
        float l = 360.0f + 5.0f * (float)i;
        m.eta.lambda.push_back(l);
        m.k.lambda.push_back(l);
        m.eta.value.push_back(nk[2 * i]);
        m.k.value.push_back(nk[2 * i + 1]);
    }
    return m;
}

float ComplexIor::r0(float l) const {
    float n = eta.eval(l), kk = k.eval(l);
    return ((n - 1.0f) * (n - 1.0f) + kk * kk) / ((n + 1.0f) * (n + 1.0f) + kk * kk);
}

std::array<float, 3> ComplexIor::f0_srgb() const {
    return reflectance_srgb([this](float l) { return r0(l); });
}

static std::string lower(std::string s) {
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

static std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\n\r"), b = s.find_last_not_of(" \t\n\r");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// metals.rs:111-161
static int metal_index(const std::string& name) {
    std::string n = lower(trim(name));
    const char* names[13][3] = {{"ag", "silver", ""},  {"au", "gold", ""},      {"cu", "copper", ""},
                                {"pt", "platinum", ""}, {"pd", "palladium", ""}, {"rh", "rhodium", ""},
                                {"ni", "nickel", ""},   {"ti", "titanium", ""},  {"al", "aluminium", "aluminum"},
                                {"cr", "chromium", "chrome"}, {"fe", "iron", ""}, {"zn", "zinc", ""},
                                {"sn", "tin", ""}};
    for (int i = 0; i < 13; i++)
        for (int j = 0; j < 3; j++)
            if (names[i][j][0] && n == names[i][j]) return i;
    return -1;
}

bool metal_get(const std::string& name, ComplexIor* out) {
    int i = metal_index(name);
    if (i < 0) return false;
    *out = ior_from_grid(ag_data::METAL_NK[i]);
    return true;
}

static const float* metal_rows(const char* sym) { return ag_data::METAL_NK[metal_index(sym)]; }

// ------------------------------------------------------------------ alloys
// alloys.rs:283-489 (f64 model, f32 tables out)
namespace {
const float DENSITY_AU = 19.32f, DENSITY_AG = 10.49f, DENSITY_CU = 8.96f;
const double HC_EV_NM = 1239.84193;
const double EPS0_OVER_HBAR = 8.8541878128e-12 / 6.582119569e-16;
const double EDGE_AU = 2.40, EDGE_CU = 2.10;
const double RHO_AU_CU = 20e-8;
const double DRUDE_AU[2] = {8.849, 0.0735}, DRUDE_AG[2] = {9.146, 0.0255}, DRUDE_CU[2] = {9.087, 0.1180};

struct Cx {
    double re, im;
    Cx add(Cx o) const { return {re + o.re, im + o.im}; }
    Cx sub(Cx o) const { return {re - o.re, im - o.im}; }
    Cx scale(double s) const { return {re * s, im * s}; }
    static Cx from_nk(float n, float k) {
        double a = n, b = k;
        return {a * a - b * b, 2.0 * a * b};
    }
    Cx sqrt() const {
        double m = std::hypot(re, im);
        double n = std::sqrt(std::max(0.5 * (m + re), 0.0));
        double k = std::sqrt(std::max(0.5 * (m - re), 0.0));
        return {n, im < 0.0 ? -k : k};
    }
};
using Grid = std::array<Cx, GRID_N>;

double grid_lambda(int i) { return (double)(360.0f + 5.0f * (float)i); }

Grid eps_of(const float* rows) {
    Grid g;
    for (int i = 0; i < GRID_N; i++) g[i] = Cx::from_nk(rows[2 * i], rows[2 * i + 1]);
    return g;
}

Cx drude(double wp, double g, double e) {
    double d = e * e + g * g;
    return {-wp * wp / d, wp * wp * g / (e * d)};
}

Grid interband(const Grid& eps, const double* wg) {
    Grid out = eps;
    for (int i = 0; i < GRID_N; i++) {
        double e = HC_EV_NM / grid_lambda(i);
        out[i] = out[i].sub(drude(wg[0], wg[1], e)).sub(Cx{1.0, 0.0});
    }
    return out;
}

Cx at(const Grid& g, double lambda) {
    double x = std::clamp((lambda - 360.0) / 5.0, 0.0, (double)(GRID_N - 1));
    int i = std::min((int)x, GRID_N - 2);
    double t = x - (double)i;
    return g[i].scale(1.0 - t).add(g[i + 1].scale(t));
}

ComplexIor to_ior(const Grid& eps) {
    float rows[2 * GRID_N];
    for (int i = 0; i < GRID_N; i++) {
        Cx nk = eps[i].sqrt();
        rows[2 * i] = std::fmax((float)nk.re, 1e-4f);
        rows[2 * i + 1] = std::fmax((float)nk.im, 0.0f);
    }
    return ior_from_grid(rows);
}

template <int K>
std::array<double, K> volume_fractions(const float (&m)[K], const float (&rho)[K]) {
    std::array<double, K> v;
    double s = 0.0;
    for (int i = 0; i < K; i++) {
        v[i] = (double)(std::fmax(m[i], 0.0f) / rho[i]);
        s += v[i];
    }
    if (s > 0.0)
        for (auto& o : v) o /= s;
    return v;
}

// alloys.rs:407-418
ComplexIor mix_permittivity(const std::vector<std::pair<float, const ComplexIor*>>& parts) {
    double total = 0.0;
    for (auto& p : parts) total += (double)std::fmax(p.first, 0.0f);
    Grid eps;
    for (int i = 0; i < GRID_N; i++) {
        eps[i] = Cx{0.0, 0.0};
        float l = (float)grid_lambda(i);
        for (auto& p : parts) {
            double w = total > 0.0 ? (double)std::fmax(p.first, 0.0f) / total : 0.0;
            eps[i] = eps[i].add(Cx::from_nk(p.second->eta.eval(l), p.second->k.eval(l)).scale(w));
        }
    }
    return to_ior(eps);
}

// alloys.rs:464-489
ComplexIor bronze() {
    const double M_CU = 63.546, M_SN = 118.71, M_ZN = 65.38;
    double x_sn = (0.12 / M_SN) / (0.12 / M_SN + 0.88 / M_CU);
    double ea = 1.0 + 3.0 * x_sn;
    auto ea_brass = [&](double zn) { return 1.0 + (zn / M_ZN) / (zn / M_ZN + (1.0 - zn) / M_CU); };
    double ea15 = ea_brass(0.15), ea30 = ea_brass(0.30);
    double t = (ea - ea15) / (ea30 - ea15);
    Grid q_cu = eps_of(ag_data::CU_QUERRY), q15 = eps_of(ag_data::CU85ZN15_QUERRY),
         q30 = eps_of(ag_data::CU70ZN30_QUERRY), jc_cu = eps_of(metal_rows("Cu"));
    double drho = (x_sn * 100.0 * 2.88 - (ea - 1.0) * 100.0 * 0.30) * 1e-8;
    double wp = DRUDE_CU[0], g = DRUDE_CU[1];
    double dg = EPS0_OVER_HBAR * wp * wp * drho;
    Grid eps;
    for (int i = 0; i < GRID_N; i++) {
        Cx brass = q15[i].scale(1.0 - t).add(q30[i].scale(t));
        double en = HC_EV_NM / grid_lambda(i);
        eps[i] = jc_cu[i].add(brass.sub(q_cu[i])).add(drude(wp, g + dg, en)).sub(drude(wp, g, en));
    }
    return to_ior(eps);
}
}  // namespace

// alloys.rs:423-460
ComplexIor karat_gold(float au, float ag_, float cu) {
    if (std::fmax(au, 0.0f) + std::fmax(ag_, 0.0f) + std::fmax(cu, 0.0f) <= 0.0f) {
        ComplexIor g;
        metal_get("Au", &g);
        return g;
    }
    const float m[3] = {au, ag_, cu}, rho[3] = {DENSITY_AU, DENSITY_AG, DENSITY_CU};
    auto v = volume_fractions<3>(m, rho);
    double fau = v[0], fag = v[1], fcu = v[2];
    Grid e_au = eps_of(metal_rows("Au")), e_ag = eps_of(metal_rows("Ag")), e_cu = eps_of(metal_rows("Cu"));
    Grid eps;
    for (int i = 0; i < GRID_N; i++) {
        Cx excess = Cx{(double)ag_data::AU_AG_EXCESS_50[2 * i], (double)ag_data::AU_AG_EXCESS_50[2 * i + 1]}.scale(
            4.0 * fau * fag);
        eps[i] = e_au[i].scale(fau).add(e_ag[i].scale(fag)).add(e_cu[i].scale(fcu)).add(excess);
    }
    if (fau > 0.0 && fcu > 0.0) {
        double edge = (fau * EDGE_AU + fcu * EDGE_CU) / (fau + fcu);
        struct Part {
            double f;
            const Grid* e;
            const double* dr;
            double edge;
        } parts[2] = {{fau, &e_au, DRUDE_AU, EDGE_AU}, {fcu, &e_cu, DRUDE_CU, EDGE_CU}};
        for (auto& p : parts) {
            Grid ib = interband(*p.e, p.dr);
            double shift = edge - p.edge;
            for (int i = 0; i < GRID_N; i++) {
                double e = HC_EV_NM / grid_lambda(i);
                Cx moved = at(ib, HC_EV_NM / (e - shift));
                eps[i] = eps[i].add(moved.sub(ib[i]).scale(p.f));
            }
        }
    }
    double wp2 = fau * (DRUDE_AU[0] * DRUDE_AU[0]) + fag * (DRUDE_AG[0] * DRUDE_AG[0]) + fcu * (DRUDE_CU[0] * DRUDE_CU[0]);
    double g0 = fau * DRUDE_AU[1] + fag * DRUDE_AG[1] + fcu * DRUDE_CU[1];
    double dg = EPS0_OVER_HBAR * wp2 * 4.0 * RHO_AU_CU * fau * fcu;
    if (dg > 0.0) {
        double wp = std::sqrt(wp2);
        for (int i = 0; i < GRID_N; i++) {
            double en = HC_EV_NM / grid_lambda(i);
            eps[i] = eps[i].add(drude(wp, g0 + dg, en)).sub(drude(wp, g0, en));
        }
    }
    return to_ior(eps);
}

// alloys.rs:260-280
bool alloy_from_name(const std::string& name, int* alloy) {
    std::string s = lower(trim(name));
    for (auto& c : s)
        if (c == '_' || c == ' ') c = '-';
    struct E {
        const char* k;
        int a;
    } t[] = {{"silver", Sterling},   {"sterling", Sterling},     {"925", Sterling},         {"sterling-silver", Sterling},
             {"fine-silver", FineSilver}, {"999", FineSilver},    {"gold", Gold14kYellow},   {"yellow", Gold14kYellow},
             {"gold14k", Gold14kYellow}, {"14k", Gold14kYellow}, {"gold-filled", Gold14kYellow}, {"gf", Gold14kYellow},
             {"yellow-gold", Gold14kYellow}, {"gold24k", Gold24k}, {"24k", Gold24k},       {"gold18k", Gold18kYellow},
             {"18k", Gold18kYellow},     {"gold10k", Gold10kYellow}, {"10k", Gold10kYellow}, {"rose", Gold14kRose},
             {"rose-gold", Gold14kRose}, {"rose14k", Gold14kRose}, {"white", Gold14kWhite},   {"white-gold", Gold14kWhite},
             {"white14k", Gold14kWhite}, {"platinum", Platinum},   {"pt", Platinum},         {"palladium", Palladium},
             {"pd", Palladium},          {"rhodium", Rhodium},     {"rh", Rhodium},          {"copper", Copper},
             {"cu", Copper},             {"bronze", Bronze},       {"brass", Brass}};
    for (auto& e : t)
        if (s == e.k) {
            *alloy = e.a;
            return true;
        }
    return false;
}

// alloys.rs:142-171
ComplexIor alloy_ior(int alloy) {
    auto pure = [](const char* s) {
        ComplexIor m;
        metal_get(s, &m);
        return m;
    };
    switch (alloy) {
        case FineSilver: return pure("Ag");
        case Gold24k: return pure("Au");
        case Gold14kWhite:
        case Rhodium: return pure("Rh");
        case Platinum: return pure("Pt");
        case Palladium: return pure("Pd");
        case Copper: return pure("Cu");
        case Sterling: {
            ComplexIor a = pure("Ag"), c = pure("Cu");
            const float m[2] = {0.925f, 0.075f}, rho[2] = {DENSITY_AG, DENSITY_CU};
            auto v = volume_fractions<2>(m, rho);
            return mix_permittivity({{(float)v[0], &a}, {(float)v[1], &c}});
        }
        case Gold18kYellow: return karat_gold(0.75f, 0.125f, 0.125f);
        case Gold14kYellow: return karat_gold(0.585f, 0.2075f, 0.2075f);
        case Gold10kYellow: return karat_gold(0.417f, 0.2915f, 0.2915f);
        case Gold14kRose: return karat_gold(0.585f, 0.09f, 0.325f);
        case Brass: return ior_from_grid(ag_data::CU70ZN30_QUERRY);
        case Bronze: return bronze();
        default: return pure("Ag");
    }
}

// ------------------------------------------------------------------ gems
// gems.rs:149-257
namespace {
const float HC_F = 1239.8419f;
const float FWHM_TO_SIGMA = 1.0f / 2.35482f;
const float LINE_F = 486.13f, LINE_D = 587.56f, LINE_C = 656.27f, LINE_B = 686.7f, LINE_G = 430.8f, LINE_NA_D = 589.3f;

struct Sellmeier {
    float b[3], c[3];
    float eval(float l_nm) const {
        float x = l_nm * 1e-3f;
        float l2 = x * x;
        float n2 = 1.0f;
        for (int i = 0; i < 3; i++) n2 += b[i] * l2 / (l2 - c[i]);
        return std::sqrt(std::fmax(n2, 1.0f));
    }
};

struct Cauchy {
    float a, b;
    static Cauchy gemological(float n_d, float bg) {
        auto inv2 = [](float nm) {
            float x = nm * 1e-3f;
            return 1.0f / (x * x);
        };
        float b = bg / (inv2(LINE_G) - inv2(LINE_B));
        return Cauchy{n_d - b * inv2(LINE_NA_D), b};
    }
    float eval(float l_nm) const {
        float l = l_nm * 1e-3f;
        return a + b / (l * l);
    }
};

struct Band {
    float center, lo, hi, peak;
    float eval_ev(float e) const {
        float e0 = HC_F / center;
        float w = (e < e0 ? lo : hi) * FWHM_TO_SIGMA;
        float z = (e - e0) / w;
        return peak * std::exp(-0.5f * z * z);
    }
};
struct Urbach {
    bool on;
    float edge_nm, at_edge, slope_ev;
};
struct Bands {
    float floor;
    std::vector<Band> bands;
    Urbach urbach;
    float eval(float l_nm) const {
        float e = HC_F / std::fmax(l_nm, 1.0f);
        float s = floor;
        for (auto& b : bands) s += b.eval_ev(e);
        if (urbach.on) s += urbach.at_edge * std::exp(std::fmin((e - HC_F / urbach.edge_nm) / urbach.slope_ev, 60.0f));
        return std::fmax(s, 0.0f);
    }
};
Band sym(float c, float f, float p) { return Band{c, f, f, p}; }
Band split(float c, float lo, float hi, float p) { return Band{c, lo, hi, p}; }

const Sellmeier DIAMOND{{4.3356f, 0.3306f, 0.0f}, {0.011236f, 0.030625f, 0.0f}};
const Sellmeier CUBIC_ZIRCONIA{{1.347091f, 2.117788f, 9.452943f}, {0.003911627f, 0.027801894f, 591.4901f}};
const Sellmeier MOISSANITE{{0.20075f, 5.54861f, 35.65066f}, {-12.07224f, 0.02641f, 1268.247f}};
const Sellmeier CORUNDUM_O{{1.4313493f, 0.65054713f, 5.3414021f}, {0.005279926f, 0.014238265f, 325.0178f}};
const Sellmeier QUARTZ_O{{0.28604141f, 1.0704408f, 1.1020224f}, {0.0f, 0.0100586f, 100.0f}};
const Sellmeier SPINEL{{1.8938f, 3.0755f, 0.0f}, {0.009884336f, 250.46228f, 0.0f}};
const Sellmeier N_BK7{{1.0396121f, 0.23179234f, 1.0104695f}, {0.0060006987f, 0.020017914f, 103.56065f}};

Urbach none() { return Urbach{false, 0, 0, 0}; }
Urbach urb(float e, float a, float s) { return Urbach{true, e, a, s}; }

Bands SAPPHIRE() {
    return {0.0706f, {sym(565, 0.62f, 0.5615f), sym(700, 0.50f, 0.9784f), sym(450, 0.08f, 0.15f), sym(388, 0.06f, 0.20f), sym(377, 0.06f, 0.25f)},
            urb(360, 0.3f, 0.12f)};
}
Bands RUBY() { return {0.005f, {split(556, 0.24f, 0.50f, 1.9751f), split(405, 0.40f, 0.60f, 2.9627f)}, urb(360, 0.5f, 0.15f)}; }
Bands PINK_SAPPHIRE() { return {0.003f, {split(556, 0.24f, 0.50f, 0.2785f), split(405, 0.40f, 0.60f, 0.4177f)}, none()}; }
Bands EMERALD() {
    return {0.1395f, {split(610, 0.36f, 0.44f, 0.40f), split(430, 0.40f, 0.50f, 0.6177f), sym(830, 0.55f, 0.171f)}, none()};
}
Bands AQUAMARINE() { return {0.0324f, {sym(820, 0.65f, 0.764f), sym(427, 0.10f, 0.03f), sym(370, 0.12f, 0.06f)}, none()}; }
Bands MORGANITE() { return {0.0067f, {sym(495, 0.45f, 0.0976f), sym(560, 0.35f, 0.0329f)}, urb(400, 0.18f, 0.25f)}; }
Bands AMETHYST() { return {0.0471f, {sym(545, 0.65f, 0.3908f), sym(357, 0.50f, 0.035f)}, none()}; }
Bands CITRINE() { return {0.0008f, {}, urb(420, 0.8236f, 0.36f)}; }
Bands SMOKY_QUARTZ() { return {0.0791f, {sym(460, 0.90f, 0.4317f), sym(620, 0.75f, 0.30f)}, urb(380, 0.6f, 0.25f)}; }
Bands PERIDOT() {
    return {0.035f,
            {sym(850, 0.55f, 1.0f), sym(1050, 0.55f, 6.0f), sym(1250, 0.55f, 3.0f), sym(453, 0.10f, 0.20f), sym(473, 0.10f, 0.12f),
             sym(493, 0.10f, 0.18f)},
            urb(400, 0.9f, 0.25f)};
}
Bands BLUE_TOPAZ() { return {0.0083f, {sym(640, 0.65f, 0.2855f)}, none()}; }
Bands GARNET() {
    return {0.0815f,
            {sym(505, 0.10f, 0.6f), sym(527, 0.10f, 1.6124f), sym(576, 0.12f, 2.25f), sym(460, 0.35f, 0.6986f), sym(615, 0.15f, 0.30f),
             sym(430, 0.12f, 0.20f)},
            urb(400, 1.0f, 0.25f)};
}
Bands TSAVORITE() { return {0.1083f, {split(605, 0.42f, 0.48f, 0.478f), split(430, 0.42f, 0.48f, 1.192f)}, none()}; }
Bands SPINEL_RED() { return {0.008f, {split(545, 0.24f, 0.50f, 2.5937f), split(395, 0.40f, 0.60f, 3.8905f)}, none()}; }
Bands TANZANITE() { return {0.146f, {sym(595, 0.38f, 0.7046f), sym(530, 0.35f, 0.208f)}, none()}; }
Bands TOURMALINE() { return {0.0365f, {sym(720, 0.55f, 1.5f), sym(470, 0.45f, 0.72f)}, none()}; }
Bands MOISSANITE_ABS() { return {0.0005f, {}, urb(380, 0.05f, 0.08f)}; }

Piecewise bk7_absorption() {
    const float t[][2] = {{350, 3.335e-3f}, {365, 1.177e-3f}, {370, 9.308e-4f}, {380, 6.859e-4f}, {390, 4.424e-4f},
                          {400, 3.213e-4f}, {405, 2.810e-4f}, {420, 2.810e-4f}, {436, 3.213e-4f}, {460, 2.810e-4f},
                          {500, 2.407e-4f}, {546, 1.603e-4f}, {580, 2.005e-4f}, {620, 2.407e-4f}, {660, 2.407e-4f},
                          {700, 1.603e-4f}, {1060, 1.202e-4f}};
    Piecewise p;
    for (auto& r : t) {
        p.lambda.push_back(r[0]);
        p.value.push_back(r[1]);
    }
    return p;
}

template <class I, class A>
GemOptics make_gem(const char* name, int kind, I ior, A abs) {
    GemOptics g;
    g.name = name;
    g.kind = kind;
    g.ior = [ior](float l) { return ior.eval(l); };
    g.absorption = [abs](float l) { return abs.eval(l); };
    g.has_body = false;
    g.body = {0, 0, 0};
    return g;
}
struct Const {
    float v;
    float eval(float) const { return v; }
};
}  // namespace

// gems.rs:395-440
bool gem_get(const std::string& name, GemOptics* out) {
    std::string n = lower(trim(name));
    for (auto& c : n)
        if (c == '-' || c == ' ') c = '_';
    auto g = [](float nd, float bg) { return Cauchy::gemological(nd, bg); };
    auto T = [](const char* nm, auto ior, auto abs) { return make_gem(nm, 0, ior, abs); };
    auto opaque = [&](const char* nm, float nd, float bg, std::array<float, 3> alb) {
        GemOptics o = make_gem(nm, 2, Cauchy::gemological(nd, bg), Const{20.0f});
        o.has_body = true;
        o.body = alb;
        return o;
    };
    if (n == "diamond") *out = T("diamond", DIAMOND, Const{0.0f});
    else if (n == "cubic_zirconia" || n == "cz" || n == "zirconia") *out = T("cubic_zirconia", CUBIC_ZIRCONIA, Const{0.0005f});
    else if (n == "moissanite") *out = T("moissanite", MOISSANITE, MOISSANITE_ABS());
    else if (n == "sapphire" || n == "blue_sapphire") *out = T("sapphire", CORUNDUM_O, SAPPHIRE());
    else if (n == "ruby") *out = T("ruby", CORUNDUM_O, RUBY());
    else if (n == "pink_sapphire") *out = T("pink_sapphire", CORUNDUM_O, PINK_SAPPHIRE());
    else if (n == "emerald") *out = T("emerald", g(1.580f, 0.014f), EMERALD());
    else if (n == "aquamarine") *out = T("aquamarine", g(1.577f, 0.014f), AQUAMARINE());
    else if (n == "morganite") *out = T("morganite", g(1.585f, 0.014f), MORGANITE());
    else if (n == "amethyst") *out = T("amethyst", QUARTZ_O, AMETHYST());
    else if (n == "citrine") *out = T("citrine", QUARTZ_O, CITRINE());
    else if (n == "smoky_quartz" || n == "smoky") *out = T("smoky_quartz", QUARTZ_O, SMOKY_QUARTZ());
    else if (n == "peridot" || n == "olivine") *out = T("peridot", g(1.672f, 0.020f), PERIDOT());
    else if (n == "blue_topaz") *out = T("blue_topaz", g(1.613f, 0.014f), BLUE_TOPAZ());
    else if (n == "white_topaz" || n == "topaz") *out = T("white_topaz", g(1.613f, 0.014f), Const{0.0005f});
    else if (n == "garnet" || n == "almandine" || n == "almandite") *out = T("garnet", g(1.790f, 0.027f), GARNET());
    else if (n == "tsavorite") *out = T("tsavorite", g(1.740f, 0.027f), TSAVORITE());
    else if (n == "spinel" || n == "red_spinel") *out = T("spinel", SPINEL, SPINEL_RED());
    else if (n == "tanzanite") *out = T("tanzanite", g(1.694f, 0.030f), TANZANITE());
    else if (n == "tourmaline" || n == "green_tourmaline") *out = T("tourmaline", g(1.635f, 0.017f), TOURMALINE());
    else if (n == "onyx" || n == "black_onyx") *out = opaque("onyx", 1.537f, 0.013f, {0.012f, 0.012f, 0.013f});
    else if (n == "obsidian") *out = opaque("obsidian", 1.490f, 0.010f, {0.016f, 0.015f, 0.015f});
    else if (n == "turquoise") *out = opaque("turquoise", 1.620f, 0.010f, {0.080f, 0.456f, 0.429f});
    else if (n == "lapis" || n == "lapis_lazuli") *out = opaque("lapis", 1.500f, 0.010f, {0.019f, 0.120f, 0.332f});
    else if (n == "opal" || n == "white_opal") {
        *out = make_gem("opal", 1, g(1.450f, 0.0115f), Const{0.05f});
        out->has_body = true;
        out->body = {0.55f, 0.58f, 0.62f};
    } else if (n == "black_opal") {
        *out = make_gem("black_opal", 1, g(1.450f, 0.0115f), Const{1.5f});
        out->has_body = true;
        out->body = {0.012f, 0.012f, 0.018f};
    } else if (n == "glass" || n == "bk7" || n == "n_bk7" || n == "crown_glass") {
        *out = T("glass", N_BK7, bk7_absorption());
    } else {
        return false;
    }
    return true;
}

// ------------------------------------------------------------------ colour
// color.rs:32-55
std::array<float, 3> blackbody_white(float kelvin) {
    Blackbody bb = Blackbody::peak_normalized(kelvin);
    float xyz[3] = {0, 0, 0};
    float l = 360.0f;
    while (l <= 830.0f) {
        auto c = xyz_bar(l);
        float e = bb.eval(l);
        for (int i = 0; i < 3; i++) xyz[i] += c[i] * e;
        l += 1.0f;
    }
    return {xyz[0] / xyz[1], 1.0f, xyz[2] / xyz[1]};
}

namespace {
using M3 = std::array<std::array<float, 3>, 3>;
const M3 BRADFORD = {{{0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}}};
const M3 BRADFORD_INV = {{{0.986993f, -0.147054f, 0.159963f}, {0.432305f, 0.51836f, 0.049291f}, {-0.008529f, 0.040043f, 0.968487f}}};
const M3 XYZ_TO_SRGB = {{{3.2404542f, -1.5371385f, -0.4985314f}, {-0.969266f, 1.8760108f, 0.041556f}, {0.0556434f, -0.2040259f, 1.0572252f}}};
const float D65W[3] = {0.95047f, 1.0f, 1.08883f};
M3 mul(const M3& a, const M3& b) {
    M3 r{};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            float s = 0.0f;  // Iterator::sum starts from 0
            for (int k = 0; k < 3; k++) s += a[i][k] * b[k][j];
            r[i][j] = s;
        }
    return r;
}
std::array<float, 3> apply(const M3& m, const float* v) {
    return {m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2], m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
            m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]};
}
}  // namespace

void camera_matrix(bool has_white, const float* w, float out[9]) {
    M3 r = XYZ_TO_SRGB;
    if (has_white) {
        auto src = apply(BRADFORD, w);
        auto dst = apply(BRADFORD, D65W);
        M3 d = {{{dst[0] / src[0], 0, 0}, {0, dst[1] / src[1], 0}, {0, 0, dst[2] / src[2]}}};
        M3 cat = mul(BRADFORD_INV, mul(d, BRADFORD));
        r = mul(XYZ_TO_SRGB, cat);
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[i * 3 + j] = r[i][j];
}

// ------------------------------------------------------------------ fresnel
// fresnel.rs:9-25
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

namespace {
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
}  // namespace

// fresnel.rs:81-93
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

// fresnel.rs:105-113
std::array<float, 2> gulbrandsen(float r, float g) {
    r = std::clamp(r, 0.0f, 0.9999f);
    float sr = std::sqrt(r);
    float n_min = (1.0f - r) / (1.0f + r);
    float n_max = (1.0f + sr) / (1.0f - sr);
    float n = g * n_min + (1.0f - g) * n_max;
    float k2 = ((n + 1.0f) * (n + 1.0f) * r - (n - 1.0f) * (n - 1.0f)) / (1.0f - r);
    return {n, std::sqrt(std::fmax(k2, 0.0f))};
}

// ------------------------------------------------------------------ GGX energy table
// microfacet.rs:9-140
namespace {
struct Ggx {
    float alpha;
    float lambda(V3 w) const {
        float c2 = w.z * w.z;
        float t2 = std::fmax(1.0f - c2, 0.0f) / std::fmax(c2, 1e-12f);
        if (!std::isfinite(t2)) return 0.0f;
        return 0.5f * (-1.0f + std::sqrt(1.0f + alpha * alpha * t2));
    }
    float g1(V3 w) const { return 1.0f / (1.0f + lambda(w)); }
    float g2(V3 wo, V3 wi) const { return 1.0f / (1.0f + lambda(wo) + lambda(wi)); }
    V3 sample_h(V3 wo, float u1, float u2) const {
        float a = alpha;
        V3 vh = normalize(V3(a * wo.x, a * wo.y, wo.z));
        float lensq = vh.x * vh.x + vh.y * vh.y;
        V3 t1 = lensq > 1e-12f ? V3(-vh.y, vh.x, 0.0f) / std::sqrt(lensq) : V3(1.0f, 0.0f, 0.0f);
        V3 t2 = cross(vh, t1);
        float r = std::sqrt(u1);
        float phi = 2.0f * PI_F * u2;
        float p1 = r * std::cos(phi);
        float p2 = r * std::sin(phi);
        float s = 0.5f * (1.0f + vh.z);
        p2 = (1.0f - s) * std::sqrt(std::fmax(1.0f - p1 * p1, 0.0f)) + s * p2;
        V3 nh = t1 * p1 + t2 * p2 + vh * std::sqrt(std::fmax(1.0f - p1 * p1 - p2 * p2, 0.0f));
        return normalize(V3(a * nh.x, a * nh.y, std::fmax(nh.z, 1e-6f)));
    }
};
}  // namespace

void energy_table(float* out) {
    static std::once_flag once;
    static float table[32 * 32];
    std::call_once(once, [] {
        const int EN = 32;
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
                        float u1 = ((float)a + 0.5f) / (float)n;
                        float u2 = ((float)b + 0.5f) / (float)n;
                        V3 h = g.sample_h(wo, u1, u2);
                        V3 wi = reflect(wo, h);
                        if (wi.z > 0.0f) sum += g.g2(wo, wi) / g.g1(wo);
                    }
                table[j * EN + i] = sum / (float)(n * n);
            }
        }
    });
    for (int i = 0; i < 1024; i++) out[i] = table[i];
}

// ------------------------------------------------------------------ LinearRgb
// envmap.rs:15-105
static float smoothstep(float e0, float e1, float x) {
    float t = std::clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

const LinearRgb& LinearRgb::get() {
    static LinearRgb L = [] {
        LinearRgb o;
        Blackbody bb = Blackbody::peak_normalized(6504.0f);
        for (int i = 0; i < SPEC_N; i++) {
            float l = 360.0f + (float)i;
            float wb = 1.0f - smoothstep(470.0f, 515.0f, l);
            float wr = smoothstep(565.0f, 610.0f, l);
            float wg = std::fmax(1.0f - wb - wr, 0.0f);
            float e = bb.eval(l);
            o.basis[0][i] = wr * e;
            o.basis[1][i] = wg * e;
            o.basis[2][i] = wb * e;
        }
        float m[3][3];
        for (int j = 0; j < 3; j++) {
            std::array<float, 3> xyz = {0, 0, 0};
            for (int i = 0; i < SPEC_N; i++) {
                auto c = xyz_bar(360.0f + (float)i);
                for (int k = 0; k < 3; k++) xyz[k] += c[k] * o.basis[j][i];
            }
            for (auto& v : xyz) v = v / CIE_Y_INTEGRAL;
            auto rgb = xyz_to_linear_srgb(xyz);
            for (int i = 0; i < 3; i++) m[i][j] = rgb[i];
        }
        float a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1],
              ii = m[2][2];
        float det = a * (e * ii - f * h) - b * (d * ii - f * g) + c * (d * h - e * g);
        float k = 1.0f / det;
        float inv[3][3] = {{(e * ii - f * h) * k, (c * h - b * ii) * k, (b * f - c * e) * k},
                           {(f * g - d * ii) * k, (a * ii - c * g) * k, (c * d - a * f) * k},
                           {(d * h - e * g) * k, (b * g - a * h) * k, (a * e - b * d) * k}};
        for (int r = 0; r < 3; r++)
            for (int s = 0; s < 3; s++) o.inv[r][s] = inv[r][s];
        return o;
    }();
    return L;
}

std::array<float, 3> LinearRgb::weights(std::array<float, 3> rgb) const {
    const auto& m = inv;
    return {m[0][0] * rgb[0] + m[0][1] * rgb[1] + m[0][2] * rgb[2], m[1][0] * rgb[0] + m[1][1] * rgb[1] + m[1][2] * rgb[2],
            m[2][0] * rgb[0] + m[2][1] * rgb[1] + m[2][2] * rgb[2]};
}

}  // namespace ag
