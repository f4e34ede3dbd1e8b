// Autotuner: picks a context's launch shape (ag_tuning) for a workload by
// measurement.  Written over the public C ABI only: it changes nothing a
// caller could not, and every setting it tries renders the same film bit for
// bit (path results do not depend on which wave, queue slot or lane a sample
// runs in), so tuning trades only time.
//
// Search: coordinate descent from the context's current tuning over
//   paths_per_wave      queue length: path slots in flight per wave
//   blocks_per_cu       persistent traversal blocks per CU (occupancy)
//   refill_min          idle lanes before a wavefront takes new extension rays
//   shadow_refill_min   the same for shadow rays
//   spread_min_lanes    how thinly a short queue is spread over the grid
//   sort_by_material    sort granularity: shade in material-sorted order, in
//                       arrival order, in runs ordered by material, with
//                       misses and panel hits first, or by shading family
//                       with one kernel per family
//   ordered_emission    continuation rays in the hits' order or as appended
//   cursor_slices       how many queue slices (each with a cursor) the grid pulls from
//   shade_waves         the shading kernel's occupancy (2 or 3 waves per SIMD)
//   raygen_bands        camera rays queued by image band (one per XCD) or in path order
// (for ray batches: batch_blocks_per_cu, batch_refill_min, spread_min_lanes),
// each candidate timed as the median of `repeats` runs; a change is kept when
// it is at least 1 % faster, and rounds repeat until one changes nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <functional>
#include <limits>
#include <vector>

#include "../../include/argent.h"
#include "host_internal.hpp"

namespace {

double median(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    return v.empty() ? std::numeric_limits<double>::infinity() : v[v.size() / 2];
}

struct Knob {
    uint32_t ag_tuning::*field;
    std::vector<uint32_t> values;
};

// Coordinate descent; measure() returns the median time in ms (infinity for a
// setting the context refuses).  Returns the best tuning found.
ag_tuning descend(const ag_tuning& start, const std::vector<Knob>& knobs,
                  const std::function<double(const ag_tuning&)>& measure, double* ms_start, double* ms_best,
                  uint32_t* trials) {
    ag_tuning best = start;
    double t_best = measure(best);
    *ms_start = t_best;
    *trials = 1;
    for (int round = 0; round < 4; round++) {
        bool changed = false;
        for (const Knob& k : knobs) {
            for (uint32_t v : k.values) {
                if (best.*(k.field) == v) continue;
                ag_tuning t = best;
                t.*(k.field) = v;
                double ms = measure(t);
                (*trials)++;
                if (ms < t_best * 0.99) {
                    best = t;
                    t_best = ms;
                    changed = true;
                }
            }
        }
        if (!changed) break;
    }
    *ms_best = t_best;
    return best;
}

// Queue lengths worth trying: powers of two from 2^22 up to the first that
// holds the whole pass in one wave, as far as the device's free memory allows
// (about 284 bytes per path slot, with a quarter of the free memory kept back).
std::vector<uint32_t> wave_sizes(uint32_t w, uint32_t h, uint32_t spp) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    const double need = (double)w * h * spp;
    std::vector<uint32_t> out;
    for (uint32_t p = 1u << 22; p <= AG_MAX_PATHS_PER_WAVE; p <<= 1) {
        if (p > (1u << 26) && (double)p * 284.0 > 0.75 * (double)free_b) break;
        out.push_back(p);
        if ((double)p >= need) break;
        if (p == AG_MAX_PATHS_PER_WAVE) break;
    }
    return out;
}

}  // namespace

extern "C" int ag_autotune_render(ag_ctx* ctx, const ag_scene* scene, const ag_camera* cam, uint32_t width,
                                  uint32_t height, uint32_t spp, const ag_path_settings* ps, uint32_t repeats,
                                  ag_autotune_result* out) {
    if (!ctx || !scene || !cam || !ps || !out || width == 0 || height == 0 || spp == 0)
        return ag::fail(AG_E_INVALID, "autotune: null argument or empty workload");
    ag_tuning start{};
    int rc = ag_ctx_get_tuning(ctx, &start);
    if (rc) return rc;
    ag_film* film = nullptr;
    if ((rc = ag_film_create(ctx, width, height, &film))) return rc;
    const uint32_t reps = std::max(1u, repeats);
    int err = AG_OK;
    auto measure = [&](const ag_tuning& t) -> double {
        if (ag_ctx_set_tuning(ctx, &t) != AG_OK) return std::numeric_limits<double>::infinity();
        std::vector<double> ms;
        // one untimed pass: buffers for a new queue length are allocated there
        ag_film_clear(film);
        if (ag_render_pass(ctx, scene, cam, film, ps, spp, 1) != AG_OK) {
            err = AG_E_DEVICE;
            return std::numeric_limits<double>::infinity();
        }
        for (uint32_t r = 0; r < reps; r++) {
            ag_film_clear(film);
            ag_ctx_synchronize(ctx);
            auto t0 = std::chrono::steady_clock::now();
            if (ag_render_pass(ctx, scene, cam, film, ps, spp, 1) != AG_OK) {
                err = AG_E_DEVICE;
                return std::numeric_limits<double>::infinity();
            }
            ms.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        return median(ms);
    };
    std::vector<Knob> knobs = {
        {&ag_tuning::paths_per_wave, wave_sizes(width, height, spp)},
        {&ag_tuning::blocks_per_cu, {2, 3, 4, 5, 6, 7, 8}},
        {&ag_tuning::refill_min, {16, 32, 48, 64}},
        {&ag_tuning::shadow_refill_min, {8, 16, 32, 48, 64}},
        {&ag_tuning::spread_min_lanes, {1, 4, 16, 64}},
        {&ag_tuning::sort_by_material, {0, 1, 2, 3, 4}},
        {&ag_tuning::ordered_emission, {0, 1}},
        {&ag_tuning::cursor_slices, {1, 2, 4, 8}},
        {&ag_tuning::shade_waves, {2, 3}},
        {&ag_tuning::raygen_bands, {1, 8}},
    };
    out->best = descend(start, knobs, measure, &out->ms_start, &out->ms_best, &out->trials);
    ag_film_destroy(film);
    rc = ag_ctx_set_tuning(ctx, &out->best);
    if (err && !std::isfinite(out->ms_best)) return ag::fail(err, "autotune: every render pass failed");
    return rc;
}

extern "C" int ag_autotune_rays(ag_ctx* ctx, const ag_scene* scene, const ag_rays_soa* rays_dev, uint32_t n,
                                ag_hits_soa* hits_dev, uint32_t repeats, ag_autotune_result* out) {
    if (!ctx || !scene || !rays_dev || !hits_dev || !out || n == 0)
        return ag::fail(AG_E_INVALID, "autotune: null argument or empty batch");
    ag_tuning start{};
    int rc = ag_ctx_get_tuning(ctx, &start);
    if (rc) return rc;
    const uint32_t reps = std::max(1u, repeats);
    auto measure = [&](const ag_tuning& t) -> double {
        if (ag_ctx_set_tuning(ctx, &t) != AG_OK) return std::numeric_limits<double>::infinity();
        std::vector<double> ms;
        if (ag_intersect_device(ctx, scene, rays_dev, n, hits_dev) != AG_OK)
            return std::numeric_limits<double>::infinity();
        for (uint32_t r = 0; r < reps; r++) {
            if (ag_intersect_device(ctx, scene, rays_dev, n, hits_dev) != AG_OK)
                return std::numeric_limits<double>::infinity();
            ms.push_back(ag_last_query_ms(ctx));  // kernel time (HIP events)
        }
        return median(ms);
    };
    std::vector<Knob> knobs = {
        {&ag_tuning::batch_blocks_per_cu, {2, 3, 4, 5, 6, 7, 8}},
        {&ag_tuning::batch_refill_min, {8, 16, 32, 48, 64}},
        {&ag_tuning::spread_min_lanes, {1, 4, 16, 64}},
    };
    out->best = descend(start, knobs, measure, &out->ms_start, &out->ms_best, &out->trials);
    return ag_ctx_set_tuning(ctx, &out->best);
}
