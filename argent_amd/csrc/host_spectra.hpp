// Host-side spectral library: the scene-build half of argent-spectra and the
// spectrum baking of argent-core (tab.rs), producing the tables the device
// kernels read.  All of this runs once per scene, never per sample.
#pragma once
#include <array>
#include <functional>
#include <string>
#include <vector>

#include "host_math.hpp"

namespace ag {

constexpr int SPEC_N = 471;
constexpr int GRID_N = 95;
constexpr float CIE_Y_INTEGRAL = 106.85692f;  // cie.rs:29

using SpecFn = std::function<float(float)>;

// cie.rs:38-49
std::array<float, 3> xyz_bar(float lambda);
// cie.rs:74-92
std::array<float, 3> xyz_to_linear_srgb(std::array<float, 3> xyz);
std::array<float, 3> linear_srgb_to_xyz(std::array<float, 3> rgb);
// cie.rs:154-185
std::array<float, 3> reflectance_xyz(const SpecFn& r);
std::array<float, 3> reflectance_srgb(const SpecFn& r);
std::array<float, 3> xyz_to_lab(std::array<float, 3> xyz);
std::array<float, 3> emission_xyz(const SpecFn& s);
// spectrum.rs:413-420
float luminance_y(const SpecFn& s);
float d65(float l);             // cie.rs:96-103
float d65_normalized(float l);  // cie.rs:107-113

// spectrum.rs:43-93
struct Piecewise {
    std::vector<float> lambda, value;
    float eval(float l) const;
};

// spectrum.rs:98-143
float planck(float lambda_nm, float t);
struct Blackbody {
    float kelvin, norm;
    static Blackbody peak_normalized(float kelvin);
    static Blackbody with_luminance(float kelvin, float y);
    float eval(float l) const;
};

// spectrum.rs:181-233 (sigmoid model, solved on the host)
struct RgbAlbedo {
    bool constant;
    float value;
    float c[3];
    static RgbAlbedo make(std::array<float, 3> rgb);
    float eval(float l) const;
};

// spectrum.rs:387-409
struct RgbIlluminant {
    RgbAlbedo albedo;
    float scale;
    static RgbIlluminant make(std::array<float, 3> rgb);
    float eval(float l) const;
};

// tab.rs:30-32: a 1 nm table 360..830.
void bake(const SpecFn& s, float* out471);
float tab_eval(const float* tab471, float l);

// metals.rs / alloys.rs
struct ComplexIor {
    Piecewise eta, k;
    float r0(float l) const;
    std::array<float, 3> f0_srgb() const;
};
ComplexIor ior_from_grid(const float* nk190);
bool metal_get(const std::string& name, ComplexIor* out);
bool alloy_from_name(const std::string& name, int* alloy);
ComplexIor alloy_ior(int alloy);
ComplexIor karat_gold(float au, float ag, float cu);
enum Alloy {
    FineSilver, Sterling, Gold24k, Gold18kYellow, Gold14kYellow, Gold10kYellow, Gold14kRose, Gold14kWhite,
    Platinum, Palladium, Rhodium, Copper, Bronze, Brass, ALLOY_COUNT
};

// gems.rs
struct GemOptics {
    std::string name;
    int kind;  // 0 transparent, 1 opal, 2 opaque
    SpecFn ior, absorption;
    bool has_body;
    std::array<float, 3> body;
};
bool gem_get(const std::string& name, GemOptics* out);

// color.rs
std::array<float, 3> blackbody_white(float kelvin);
void camera_matrix(bool has_white, const float* white, float m[9]);

// microfacet.rs:112-140
void energy_table(float* out1024);

// fresnel.rs
float fresnel_dielectric(float cos_i, float eta);
float fresnel_conductor1(float cos_i, float eta, float k);
std::array<float, 2> gulbrandsen(float r, float g);

// envmap.rs:15-105 LinearRgb
struct LinearRgb {
    float basis[3][SPEC_N];
    float inv[3][3];
    static const LinearRgb& get();
    std::array<float, 3> weights(std::array<float, 3> rgb) const;
};

}  // namespace ag
