// C-ABI scene-description builders: the reference's constructors
// (Material::metal/gem/plastic/diffuse, Studio::build, floor_at,
// EnvData::load, Scene::rebuild_lights) producing the flat ag_scene_desc
// that ag_scene_create uploads.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/argent.h"
#include "host_internal.hpp"
#include "host_spectra.hpp"

using namespace ag;

namespace ag {
thread_local std::string g_last_error;
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}
}  // namespace ag

extern "C" const char* ag_last_error(void) { return ag::g_last_error.c_str(); }

static float alpha_from_roughness(float r) {  // microfacet.rs:22-25
    r = std::clamp(r, 0.0f, 1.0f);
    return std::fmax(r * r, 1e-4f);
}

static void bake_rgb_albedo(std::array<float, 3> rgb, float* out) {
    RgbAlbedo a = RgbAlbedo::make(rgb);
    bake([&](float l) { return a.eval(l); }, out);
}

extern "C" int ag_alloy_ior(const char* name, float* n_out, float* k_out) {
    int a;
    if (!name || !alloy_from_name(name, &a)) return fail(AG_E_NOTFOUND, std::string("unknown alloy: ") + (name ? name : "(null)"));
    ComplexIor m = alloy_ior(a);
    for (int i = 0; i < GRID_N; i++) {
        n_out[i] = m.eta.value[i];
        k_out[i] = m.k.value[i];
    }
    return AG_OK;
}

extern "C" int ag_alloy_f0_srgb(const char* name, float* rgb) {
    int a;
    if (!name || !alloy_from_name(name, &a)) return fail(AG_E_NOTFOUND, std::string("unknown alloy: ") + (name ? name : "(null)"));
    auto c = alloy_ior(a).f0_srgb();
    for (int i = 0; i < 3; i++) rgb[i] = c[i];
    return AG_OK;
}

extern "C" int ag_karat_gold(float au, float ag_, float cu, float* n_out, float* k_out) {
    ComplexIor m = karat_gold(au, ag_, cu);
    for (int i = 0; i < GRID_N; i++) {
        n_out[i] = m.eta.value[i];
        k_out[i] = m.k.value[i];
    }
    return AG_OK;
}

// material.rs:142-149
static void default_patina(float strength, int alloy, ag_material* m) {
    std::array<float, 3> base;
    if (alloy == Sterling || alloy == FineSilver)
        base = {0.018f, 0.018f, 0.02f};
    else if (alloy == Bronze || alloy == Brass || alloy == Copper)
        base = {0.03f, 0.016f, 0.008f};
    else
        base = {0.03f, 0.02f, 0.01f};
    m->has_patina = 1;
    m->patina_strength = strength;
    m->patina_lo = 0.18f;
    m->patina_hi = 0.62f;
    m->patina_alpha = alpha_from_roughness(0.42f);
    m->patina_eta = 1.9f;
    bake_rgb_albedo(base, m->patina_base);
}

extern "C" int ag_material_metal(const char* alloy, float roughness, float patina_strength, ag_material* out) {
    int a;
    if (!out) return fail(AG_E_INVALID, "null material");
    if (!alloy || !alloy_from_name(alloy, &a)) return fail(AG_E_NOTFOUND, std::string("unknown alloy: ") + (alloy ? alloy : "(null)"));
    std::memset(out, 0, sizeof(*out));
    out->kind = AG_MAT_METAL;
    out->alpha = alpha_from_roughness(roughness);
    ComplexIor m = alloy_ior(a);
    for (int i = 0; i < GRID_N; i++) {
        out->metal_n[i] = m.eta.value[i];
        out->metal_k[i] = m.k.value[i];
    }
    if (patina_strength >= 0.0f) default_patina(patina_strength, a, out);
    return AG_OK;
}

extern "C" int ag_gem_optics(const char* name, int32_t* kind, float* ior_out, float* abs_out, float* body, int32_t* has_body) {
    GemOptics g;
    if (!name || !gem_get(name, &g)) return fail(AG_E_NOTFOUND, std::string("unknown gem: ") + (name ? name : "(null)"));
    if (kind) *kind = g.kind;
    if (ior_out) bake(g.ior, ior_out);
    if (abs_out) bake(g.absorption, abs_out);
    if (body)
        for (int i = 0; i < 3; i++) body[i] = g.body[i];
    if (has_body) *has_body = g.has_body ? 1 : 0;
    return AG_OK;
}

// material.rs:50-66
extern "C" int ag_material_gem(const char* name, ag_material* out) {
    if (!out || !name) return fail(AG_E_INVALID, "null argument");
    std::string lname(name);
    for (auto& c : lname) c = (char)std::tolower((unsigned char)c);
    std::memset(out, 0, sizeof(*out));
    if (lname.find("opal") != std::string::npos) {
        bool black = lname.find("black") != std::string::npos || lname == "opal";
        out->kind = AG_MAT_OPAL;
        std::array<float, 3> body = black ? std::array<float, 3>{0.012f, 0.012f, 0.018f} : std::array<float, 3>{0.55f, 0.58f, 0.62f};
        bake_rgb_albedo(body, out->spec_a);
        out->opal_cell_mm = 0.35f;
        out->opal_strength = black ? 1.0f : 0.7f;
        return AG_OK;
    }
    GemOptics g;
    if (!gem_get(name, &g)) return fail(AG_E_NOTFOUND, std::string("unknown gem: ") + name);
    float ior[SPEC_N];
    bake(g.ior, ior);
    bake(g.absorption, out->spec_b);
    if (g.kind == 2) {  // opaque: Coated{ggx r 0.06, eta n(550), base} (material.rs:127-131)
        out->kind = AG_MAT_GEM_OPAQUE;
        out->alpha = alpha_from_roughness(0.06f);
        out->eta = tab_eval(ior, 550.0f);
        if (g.has_body)
            bake_rgb_albedo(g.body, out->spec_a);
        else
            for (int i = 0; i < SPEC_N; i++) out->spec_a[i] = 0.01f;
    } else {
        out->kind = AG_MAT_GEM;
        std::memcpy(out->spec_a, ior, sizeof(ior));
    }
    return AG_OK;
}

extern "C" int ag_material_plastic(const float* rgb, float eta, float roughness, ag_material* out) {
    if (!out || !rgb) return fail(AG_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->kind = AG_MAT_PLASTIC;
    out->eta = eta;
    out->alpha = alpha_from_roughness(roughness);
    bake_rgb_albedo({rgb[0], rgb[1], rgb[2]}, out->spec_a);
    return AG_OK;
}

extern "C" int ag_material_diffuse(const float* rgb, ag_material* out) {
    if (!out || !rgb) return fail(AG_E_INVALID, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->kind = AG_MAT_DIFFUSE;
    bake_rgb_albedo({rgb[0], rgb[1], rgb[2]}, out->spec_a);
    return AG_OK;
}

// studio.rs:15-25
extern "C" int ag_spectrum_emitter(float kelvin, float intensity, float* out) {
    Blackbody bb = Blackbody::peak_normalized(kelvin);
    float y = std::fmax(luminance_y([&](float l) { return bb.eval(l); }), 1e-9f);
    float s = intensity / y;
    bake([&](float l) { return bb.eval(l) * s; }, out);
    return AG_OK;
}

extern "C" int ag_spectrum_rgb_albedo(const float* rgb, float* out) {
    bake_rgb_albedo({rgb[0], rgb[1], rgb[2]}, out);
    return AG_OK;
}

extern "C" int ag_spectrum_rgb_illuminant(const float* rgb, float* out) {
    RgbIlluminant il = RgbIlluminant::make({rgb[0], rgb[1], rgb[2]});
    bake([&](float l) { return il.eval(l); }, out);
    return AG_OK;
}

extern "C" float ag_luminance_y(const float* spec) {
    return luminance_y([&](float l) { return tab_eval(spec, l); });
}

extern "C" int ag_blackbody_white(float kelvin, float* xyz) {
    auto w = blackbody_white(kelvin);
    for (int i = 0; i < 3; i++) xyz[i] = w[i];
    return AG_OK;
}

extern "C" int ag_camera_matrix(int32_t has_white, const float* white, float* m) {
    camera_matrix(has_white != 0, white, m);
    return AG_OK;
}

extern "C" int ag_cie_xyz_bar(float lambda, float* out) {
    auto c = xyz_bar(lambda);
    for (int i = 0; i < 3; i++) out[i] = c[i];
    return AG_OK;
}

extern "C" int ag_reflectance_srgb(const float* spec, float* rgb) {
    auto c = reflectance_srgb([&](float l) { return tab_eval(spec, l); });
    for (int i = 0; i < 3; i++) rgb[i] = c[i];
    return AG_OK;
}

extern "C" float ag_fresnel_conductor1(float c, float eta, float k) { return fresnel_conductor1(c, eta, k); }
extern "C" float ag_fresnel_dielectric(float c, float eta) { return fresnel_dielectric(c, eta); }

extern "C" int ag_energy_table(float* out) {
    energy_table(out);
    return AG_OK;
}

extern "C" int ag_desc_init(ag_scene_desc* d) {
    if (!d) return fail(AG_E_INVALID, "null descriptor");
    std::memset(d, 0, sizeof(*d));
    d->env.kind = AG_ENV_GRADIENT;
    d->env.power = 1.0f;
    energy_table(d->energy);
    return AG_OK;
}

// light.rs:20-37 (Panel::new / Panel::facing)
static void panel_facing(ag_panel* p, V3 pos, V3 target, float w, float h, V3 up_hint, const float* emission) {
    V3 n = normalize(target - pos);
    V3 right = cross(up_hint, n);
    if (length_sq(right) < 1e-6f) right = cross(V3(1, 0, 0), n);
    right = normalize(right);
    V3 up = cross(n, right);
    V3 hu = right * (0.5f * w), hv = up * (0.5f * h);
    std::memset(p, 0, sizeof(*p));
    p->center[0] = pos.x; p->center[1] = pos.y; p->center[2] = pos.z;
    p->hu[0] = hu.x; p->hu[1] = hu.y; p->hu[2] = hu.z;
    p->hv[0] = hv.x; p->hv[1] = hv.y; p->hv[2] = hv.z;
    p->camera_visible = 0;
    p->feather = 0.35f;
    if (emission) {
        p->has_emission = 1;
        std::memcpy(p->emission, emission, sizeof(p->emission));
    }
}

// studio.rs:87-197
extern "C" int ag_desc_apply_studio(ag_scene_desc* d, const char* studio, float radius, float floor_y) {
    if (!d || !studio) return fail(AG_E_INVALID, "null argument");
    std::string s(studio);
    for (auto& c : s) c = (char)std::tolower((unsigned char)c);
    struct Box {
        const char* name;
        float x, y, z, w, h, kelvin, intensity;  // intensity < 0: flag
    };
    std::vector<Box> boxes;
    float top[3], hor[3], bot[3], power, fade0, fade1;
    ag_material floor_mat;
    float white_kelvin = -1.0f;
    if (s == "noir" || s == "brand" || s == "dark" || s == "black") {
        boxes = {{"key-top", 0, 270, 110, 340, 240, 5600, 3.2f},
                 {"strip-left", -230, 70, -50, 70, 540, 5600, 7.5f},
                 {"strip-right", 230, 70, -50, 70, 540, 5600, 7.5f},
                 {"fill-front", 30, 90, 330, 520, 300, 5400, 0.9f},
                 {"kicker-back", 0, 260, -230, 300, 90, 6000, 1.4f}};
        float t[3] = {0.010f, 0.010f, 0.011f}, h[3] = {0.006f, 0.006f, 0.0065f}, b[3] = {0.002f, 0.002f, 0.002f};
        std::memcpy(top, t, 12); std::memcpy(hor, h, 12); std::memcpy(bot, b, 12);
        power = 0.6f;
        float c[3] = {0.003f, 0.003f, 0.0032f};
        ag_material_plastic(c, 1.49f, 0.09f, &floor_mat);
        fade0 = 40.0f; fade1 = 130.0f;
        white_kelvin = 5600.0f;
    } else if (s == "white" || s == "listing" || s == "etsy" || s == "seamless") {
        boxes = {{"tent-top", 0, 300, 60, 460, 420, 5800, 2.6f},
                 {"tent-left", -270, 110, 90, 260, 380, 5800, 1.7f},
                 {"tent-right", 270, 110, 90, 260, 380, 5800, 1.7f},
                 {"tent-front", 0, 70, 340, 460, 220, 5800, 1.2f},
                 {"flag-left", -170, 60, -170, 170, 420, 0, -1},
                 {"flag-right", 170, 60, -170, 170, 420, 0, -1},
                 {"flag-back", 0, 120, -260, 300, 160, 0, -1},
                 {"flag-front-low", 0, 10, 260, 420, 60, 0, -1}};
        float t[3] = {1.8f, 1.8f, 1.8f}, h[3] = {0.55f, 0.55f, 0.56f}, b[3] = {0.9f, 0.9f, 0.9f};
        std::memcpy(top, t, 12); std::memcpy(hor, h, 12); std::memcpy(bot, b, 12);
        power = 0.45f;
        float c[3] = {0.86f, 0.86f, 0.86f};
        ag_material_diffuse(c, &floor_mat);
        fade0 = 80.0f; fade1 = 260.0f;
        white_kelvin = 5800.0f;
    } else if (s == "velvet" || s == "box") {
        boxes = {{"key-warm", -160, 220, 160, 220, 220, 3400, 3.2f},
                 {"strip-right", 240, 70, -40, 70, 460, 4200, 4.0f},
                 {"back-rim", -60, 120, -280, 260, 80, 3800, 2.0f}};
        float t[3] = {0.006f, 0.004f, 0.004f}, h[3] = {0.012f, 0.005f, 0.005f}, b[3] = {0.004f, 0.001f, 0.001f};
        std::memcpy(top, t, 12); std::memcpy(hor, h, 12); std::memcpy(bot, b, 12);
        power = 0.7f;
        float c[3] = {0.10f, 0.012f, 0.02f};
        ag_material_diffuse(c, &floor_mat);
        fade0 = 60.0f; fade1 = 200.0f;
        white_kelvin = 4300.0f;
    } else if (s == "ember" || s == "underworld" || s == "forge" || s == "fire") {
        boxes = {{"ember-left", -210, 40, -150, 80, 420, 1900, 9.0f},
                 {"ember-right", 210, 40, -150, 80, 420, 2100, 9.0f},
                 {"cold-key", 60, 220, 220, 220, 180, 9000, 1.4f},
                 {"floor-glow", 0, -5, -220, 260, 60, 1700, 2.5f}};
        float t[3] = {0.002f, 0.002f, 0.003f}, h[3] = {0.010f, 0.004f, 0.002f}, b[3] = {0.001f, 0.0005f, 0.0003f};
        std::memcpy(top, t, 12); std::memcpy(hor, h, 12); std::memcpy(bot, b, 12);
        power = 0.8f;
        float c[3] = {0.004f, 0.003f, 0.003f};
        ag_material_plastic(c, 1.5f, 0.12f, &floor_mat);
        fade0 = 40.0f; fade1 = 150.0f;
        white_kelvin = -1.0f;
    } else if (s == "window" || s == "daylight" || s == "natural") {
        boxes = {{"window", -300, 160, 80, 300, 380, 6500, 3.0f},
                 {"bounce", 260, 80, 120, 300, 300, 5800, 0.35f},
                 {"flag-back", 160, 60, -220, 160, 300, 0, -1}};
        float t[3] = {0.30f, 0.32f, 0.36f}, h[3] = {0.20f, 0.20f, 0.21f}, b[3] = {0.08f, 0.08f, 0.08f};
        std::memcpy(top, t, 12); std::memcpy(hor, h, 12); std::memcpy(bot, b, 12);
        power = 0.8f;
        float c[3] = {0.30f, 0.29f, 0.28f};
        ag_material_plastic(c, 1.5f, 0.35f, &floor_mat);
        fade0 = 70.0f; fade1 = 220.0f;
        white_kelvin = 6500.0f;
    } else {
        return fail(AG_E_NOTFOUND, "unknown studio: " + s);
    }
    if (d->n_panels + (int)boxes.size() > AG_MAX_PANELS) return fail(AG_E_INVALID, "too many panels");
    if (d->n_materials + 1 > AG_MAX_MATERIALS) return fail(AG_E_INVALID, "too many materials");
    float k = std::fmax(radius / 15.0f, 0.2f);
    for (auto& b : boxes) {
        V3 pos = V3(b.x, b.y, b.z) * k;
        float em[SPEC_N];
        if (b.intensity >= 0.0f) ag_spectrum_emitter(b.kelvin, b.intensity, em);
        panel_facing(&d->panels[d->n_panels++], pos, V3(0, 0, 0), b.w * k, b.h * k, V3(0, 1, 0),
                     b.intensity >= 0.0f ? em : nullptr);
    }
    d->env.kind = AG_ENV_GRADIENT;
    ag_spectrum_rgb_illuminant(top, d->env.top);
    ag_spectrum_rgb_illuminant(hor, d->env.horizon);
    ag_spectrum_rgb_illuminant(bot, d->env.bottom);
    d->env.power = power;
    d->materials[d->n_materials] = floor_mat;
    d->floor.present = 1;
    d->floor.material = d->n_materials;
    d->n_materials++;
    d->floor.y = floor_y - 0.005f;
    d->floor.fade_start = fade0 * k;
    d->floor.fade_end = fade1 * k;
    if (white_kelvin > 0.0f) {
        d->has_white = 1;
        auto w = blackbody_white(white_kelvin);
        for (int i = 0; i < 3; i++) d->white_xyz[i] = w[i];
    } else {
        d->has_white = 0;
    }
    return AG_OK;
}

// envmap.rs:141-178 without the file decode
extern "C" int ag_env_from_rgb(const float* rgb, int32_t w, int32_t h, float exposure_ev, float rotation_deg, ag_env* out) {
    if (!rgb || !out || w <= 0 || h <= 0) return fail(AG_E_INVALID, "bad environment image");
    const float PI_F = 3.14159265358979323846f;
    const LinearRgb& lin = LinearRgb::get();
    float k = std::pow(2.0f, exposure_ev);
    size_t n = (size_t)w * (size_t)h;
    float* weights = (float*)std::malloc(n * 3 * sizeof(float));
    float* cond = (float*)std::malloc((size_t)(w + 1) * h * sizeof(float));
    float* marg = (float*)std::malloc((size_t)(h + 1) * sizeof(float));
    if (!weights || !cond || !marg) {
        std::free(weights); std::free(cond); std::free(marg);
        return fail(AG_E_NOMEM, "environment tables");
    }
    std::vector<float> row_sum(h);
    for (int y = 0; y < h; y++) {
        float s = std::sin(PI_F * ((float)y + 0.5f) / (float)h);
        size_t base = (size_t)y * (w + 1);
        float acc = 0.0f;
        cond[base] = 0.0f;
        for (int x = 0; x < w; x++) {
            const float* c0 = rgb + ((size_t)y * w + x) * 3;
            std::array<float, 3> c = {c0[0] * k, c0[1] * k, c0[2] * k};
            auto wt = lin.weights(c);
            float* wo = weights + ((size_t)y * w + x) * 3;
            wo[0] = wt[0]; wo[1] = wt[1]; wo[2] = wt[2];
            acc += std::fmax(0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2], 0.0f) * s + 1e-6f;
            cond[base + x + 1] = acc;
        }
        row_sum[y] = acc;
        for (int x = 0; x <= w; x++) cond[base + x] /= acc;
    }
    float acc = 0.0f;
    marg[0] = 0.0f;
    for (int y = 0; y < h; y++) {
        acc += row_sum[y];
        marg[y + 1] = acc;
    }
    float total = acc / ((float)w * (float)h);
    for (int y = 0; y <= h; y++) marg[y] /= acc;
    std::memset(out, 0, sizeof(*out));
    out->kind = AG_ENV_MAP;
    out->width = w;
    out->height = h;
    out->weights = weights;
    out->marginal = marg;
    out->conditional = cond;
    out->total = total;
    out->rotation_deg = rotation_deg;
    out->power = 1.0f;
    for (int c = 0; c < 3; c++) std::memcpy(out->basis[c], lin.basis[c], sizeof(out->basis[c]));
    return AG_OK;
}

extern "C" void ag_env_free(ag_env* env) {
    if (!env || env->kind != AG_ENV_MAP) return;
    std::free((void*)env->weights);
    std::free((void*)env->marginal);
    std::free((void*)env->conditional);
    env->weights = env->marginal = env->conditional = nullptr;
}

// scene.rs:88-110
namespace ag {
void light_cdf(const ag_scene_desc* d, int* n_lights, int* lights, float* cdf, float* env_pick) {
    float acc = 0.0f;
    int n = 0;
    for (int i = 0; i < d->n_panels; i++) {
        const ag_panel& p = d->panels[i];
        if (!p.has_emission) continue;
        V3 hu(p.hu[0], p.hu[1], p.hu[2]), hv(p.hv[0], p.hv[1], p.hv[2]);
        float area = 4.0f * length(cross(hu, hv));
        float power = area * std::fmax(luminance_y([&](float l) { return tab_eval(p.emission, l); }), 1e-6f);
        acc += power;
        lights[n] = i;
        cdf[n] = acc;
        n++;
    }
    if (acc > 0.0f)
        for (int i = 0; i < n; i++) cdf[i] /= acc;
    *n_lights = n;
    if (d->env.kind == AG_ENV_MAP)
        *env_pick = n == 0 ? 1.0f : 0.5f;
    else
        *env_pick = 0.0f;
}
}  // namespace ag

extern "C" int ag_desc_light_cdf(const ag_scene_desc* d, int32_t* n_lights, int32_t* lights, float* cdf, float* env_pick) {
    if (!d) return fail(AG_E_INVALID, "null descriptor");
    int n, l[AG_MAX_PANELS];
    light_cdf(d, &n, l, cdf, env_pick);
    *n_lights = n;
    for (int i = 0; i < n; i++) lights[i] = l[i];
    return AG_OK;
}

// ------------------------------------------------------------------ camera
// camera.rs:38-55 (Camera::look_at)
static void look_at(V3 pos, V3 target, V3 up_hint, float focal_mm, float f_stop, float aspect, ag_camera* c) {
    V3 forward = normalize(target - pos);
    V3 right = cross(forward, up_hint);
    if (length_sq(right) < 1e-8f) right = cross(forward, V3(0, 0, 1));
    right = normalize(right);
    V3 up = cross(right, forward);
    float long_half = 18.0f / focal_mm;
    float half_w, half_h;
    if (aspect >= 1.0f) {
        half_w = long_half;
        half_h = long_half / aspect;
    } else {
        half_w = long_half * aspect;
        half_h = long_half;
    }
    std::memset(c, 0, sizeof(*c));
    c->pos[0] = pos.x; c->pos[1] = pos.y; c->pos[2] = pos.z;
    c->forward[0] = forward.x; c->forward[1] = forward.y; c->forward[2] = forward.z;
    c->right[0] = right.x; c->right[1] = right.y; c->right[2] = right.z;
    c->up[0] = up.x; c->up[1] = up.y; c->up[2] = up.z;
    c->half_w = half_w;
    c->half_h = half_h;
    c->focus_dist = length(target - pos);
    c->lens_radius = f_stop > 0.0f ? 0.5f * focal_mm / f_stop : 0.0f;
}

extern "C" int ag_camera_look_at(const float* pos, const float* target, const float* up_hint, float focal_mm,
                                 float f_stop, float aspect, ag_camera* out) {
    if (!pos || !target || !up_hint || !out) return fail(AG_E_INVALID, "null argument");
    look_at(V3(pos[0], pos[1], pos[2]), V3(target[0], target[1], target[2]), V3(up_hint[0], up_hint[1], up_hint[2]),
            focal_mm, f_stop, aspect, out);
    return AG_OK;
}

// camera.rs:73-80 (Camera::project)
static bool project(const ag_camera* c, V3 p, float* x, float* y) {
    V3 d = p - V3(c->pos[0], c->pos[1], c->pos[2]);
    float z = dot(d, V3(c->forward[0], c->forward[1], c->forward[2]));
    if (z <= 1e-6f) return false;
    *x = dot(d, V3(c->right[0], c->right[1], c->right[2])) / z / c->half_w - c->shift[0];
    *y = dot(d, V3(c->up[0], c->up[1], c->up[2])) / z / c->half_h - c->shift[1];
    return true;
}

// views.rs:45-83 (frame): place a camera along the view so the points fill
// `fill` of the frame, centred.
extern "C" int ag_camera_frame(const float* points, uint32_t n, float azimuth_deg, float elevation_deg, float fill,
                               float focal_mm, float f_stop, float aspect, ag_camera* out) {
    if (!points || !out || n == 0) return fail(AG_E_INVALID, "no points to frame");
    const float k = 3.14159265358979323846f / 180.0f;
    V3 lo(3.4028235e38f, 3.4028235e38f, 3.4028235e38f), hi(-3.4028235e38f, -3.4028235e38f, -3.4028235e38f);
    for (uint32_t i = 0; i < n; i++) {
        V3 p(points[3 * i], points[3 * i + 1], points[3 * i + 2]);
        lo = vmin(lo, p);
        hi = vmax(hi, p);
    }
    V3 center = (lo + hi) * 0.5f;
    float radius = std::fmax(length((hi - lo) * 0.5f), 1e-3f);
    float sa = std::sin(azimuth_deg * k), ca = std::cos(azimuth_deg * k);
    float se = std::sin(elevation_deg * k), ce = std::cos(elevation_deg * k);
    V3 dir(sa * ce, se, ca * ce);
    V3 up = std::fabs(dir.y) > 0.98f ? V3(0, 0, 1) : V3(0, 1, 0);
    float half = std::fmin(18.0f / focal_mm, 18.0f / focal_mm * (aspect < 1.0f ? aspect : 1.0f / aspect));
    float fl = std::fmax(fill, 0.05f);
    float dist = radius / std::sin(std::atan(half)) / fl;
    V3 target = center;
    ag_camera cam;
    look_at(target + dir * dist, target, up, focal_mm, f_stop, aspect, &cam);
    for (int it = 0; it < 6; it++) {
        float mnx = 3.4028235e38f, mxx = -3.4028235e38f, mny = 3.4028235e38f, mxy = -3.4028235e38f;
        for (uint32_t i = 0; i < n; i++) {
            float x, y;
            if (project(&cam, V3(points[3 * i], points[3 * i + 1], points[3 * i + 2]), &x, &y)) {
                mnx = std::fmin(mnx, x);
                mxx = std::fmax(mxx, x);
                mny = std::fmin(mny, y);
                mxy = std::fmax(mxy, y);
            }
        }
        if (mnx > mxx) break;
        float ext = std::fmax((mxx - mnx) * 0.5f, (mxy - mny) * 0.5f);
        float cx = (mxx + mnx) * 0.5f * cam.half_w;
        float cy = (mxy + mny) * 0.5f * cam.half_h;
        V3 cpos(cam.pos[0], cam.pos[1], cam.pos[2]);
        float pd = dot(center - cpos, V3(cam.forward[0], cam.forward[1], cam.forward[2]));
        target = target + V3(cam.right[0], cam.right[1], cam.right[2]) * (cx * pd) + V3(cam.up[0], cam.up[1], cam.up[2]) * (cy * pd);
        dist *= ext / fl;
        look_at(target + dir * dist, target, up, focal_mm, f_stop, aspect, &cam);
    }
    *out = cam;
    return AG_OK;
}
