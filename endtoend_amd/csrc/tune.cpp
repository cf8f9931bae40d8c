// tune.cpp -- the schedule autotuner: coordinate descent over the launch-schedule knobs, timing
// the caller's own workload on the device.  Host code; the workload runs through the library's C
// entry points.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "context.hpp"

using namespace cccs;

namespace {

struct Knob {
    const char* key;
    std::vector<int> values;
};

// The workload of one tuning trial; its result stays on the device.
struct Trial {
    cccs_ctx* ctx;
    int kind;
    const cccs_beam* beams;
    int nbeams;
    const float* const* fluence;
    void* d_fl = nullptr;
    void* d_dose = nullptr;
    size_t dose_bytes = 0;

    int run(double& ms) {
        const auto t0 = std::chrono::steady_clock::now();
        if (kind == CCCS_TUNE_BEAM) {
            CCCS_TRY(cccs_dev_zero(ctx, d_dose, dose_bytes));
            CCCS_TRY(cccs_beam_dose_device(ctx, beams, nbeams, static_cast<const float*>(d_fl),
                                           static_cast<float*>(d_dose)));
        } else {
            CCCS_TRY(cccs_beamlet_dose(ctx, beams, nbeams, fluence, 1e-4f, CCCS_BEAMLET_FAITHFUL, nullptr, nullptr));
        }
        CCCS_TRY(cccs_ctx_sync(ctx));
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return CCCS_OK;
    }
    // the faster of two calls, so that a one-off stall does not decide
    int time(double& ms, int& trials) {
        double a = 0.0, b = 0.0;
        CCCS_TRY(run(a));
        CCCS_TRY(run(b));
        trials += 2;
        ms = std::min(a, b);
        return CCCS_OK;
    }
};

int tune(cccs_ctx* ctx, Trial& t, const std::vector<Knob>& knobs, cccs_tune_report* rep) {
    CCCS_TRY(t.run(rep->default_ms));  // warm: buffers grow to the workload's size
    CCCS_TRY(t.time(rep->default_ms, rep->trials));
    double best = rep->default_ms;
    for (int pass = 0; pass < 2; pass++) {
        bool moved = false;
        for (const Knob& k : knobs) {
            int cur = 0;
            CCCS_TRY(cccs_ctx_get_schedule(ctx, k.key, &cur));
            int pick = cur;
            for (int v : k.values) {
                if (v == cur) continue;
                CCCS_TRY(cccs_ctx_set_schedule(ctx, k.key, v));
                double ms = 0.0;
                CCCS_TRY(t.time(ms, rep->trials));
                if (ms < 0.98 * best) {  // a move must beat the incumbent by 2 %
                    best = ms;
                    pick = v;
                }
            }
            CCCS_TRY(cccs_ctx_set_schedule(ctx, k.key, pick));
            moved = moved || pick != cur;
        }
        if (!moved) break;
    }
    rep->best_ms = best;
    rep->nkeys = (int)knobs.size();
    for (size_t i = 0; i < knobs.size() && i < 16; i++) {
        std::strncpy(rep->key[i], knobs[i].key, 31);
        CCCS_TRY(cccs_ctx_get_schedule(ctx, knobs[i].key, &rep->value[i]));
    }
    return CCCS_OK;
}

}  // namespace

extern "C" int cccs_autotune(cccs_ctx* ctx, int kind, const cccs_beam* beams, int nbeams, const float* const* fluence,
                             cccs_tune_report* report) {
    clear_error();
    if (!ctx || !beams || nbeams < 1 || !fluence || !report) return set_error(CCCS_EINVAL, "null argument");
    if (kind != CCCS_TUNE_BEAM && kind != CCCS_TUNE_BEAMLET) return set_error(CCCS_EINVAL, "unknown tuning kind");
    for (int i = 0; i < nbeams; i++)
        if (!fluence[i]) return set_error(CCCS_EINVAL, "null fluence map");
    std::memset(report, 0, sizeof(*report));
    std::vector<Knob> knobs;
    if (kind == CCCS_TUNE_BEAM) {
        knobs = {{"cone_ramp_min", {32, 48, 64, 96, 128}},
                 {"cone_class_len0", {48, 64, 96}},
                 {"cone_class_len1", {128, 160, 192}},
                 {"cone_class_len2", {256, 320, 384}},
                 {"cone_class_threads1", {64, 128}},
                 {"cone_class_threads2", {64, 128, 256}},
                 {"cone_class_threads3", {64, 128, 256}},
                 {"pvcs_pass_cones", {0, 2, 4, 8, 48}}};
    } else {
        knobs = {{"bl_sample_waves", {1024, 2048, 3072, 4096, 65536, 262144}},
                 {"bl_walk_waves", {16384, 32768, 65536}},
                 {"bl_walk_split", {0, 128, 256, 384}},
                 {"bl_streams", {1, 2, 3, 4}},
                 {"bl_job_gb", {8, 12, 24, 48}}};
    }
    Trial t{ctx, kind, beams, nbeams, fluence};
    int st = CCCS_OK;
    if (kind == CCCS_TUNE_BEAM) {
        std::vector<float> fl;
        for (int i = 0; i < nbeams; i++) {
            const size_t n = (size_t)beams[i].fmap_size[0] * beams[i].fmap_size[1];
            fl.insert(fl.end(), fluence[i], fluence[i] + n);
        }
        const cccs_problem& p = ctx->prob;
        t.dose_bytes = sizeof(float) * (size_t)p.size[0] * p.size[1] * p.size[2];
        st = cccs_dev_alloc(ctx, std::max<size_t>(sizeof(float) * fl.size(), 4), &t.d_fl);
        if (st == CCCS_OK) st = cccs_dev_upload(ctx, t.d_fl, fl.data(), sizeof(float) * fl.size());
        if (st == CCCS_OK) st = cccs_dev_alloc(ctx, t.dose_bytes, &t.d_dose);
    }
    if (st == CCCS_OK) st = tune(ctx, t, knobs, report);
    const std::string err = st == CCCS_OK ? std::string() : std::string(cccs_last_error());
    if (t.d_fl) cccs_dev_free(ctx, t.d_fl);
    if (t.d_dose) cccs_dev_free(ctx, t.d_dose);
    return st == CCCS_OK ? CCCS_OK : set_error(st, err);
}
