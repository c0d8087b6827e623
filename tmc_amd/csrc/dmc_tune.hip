// dmc_tune.hip -- autotuner of the launch shape (host code).
//
// The launch shape (grid-stride block cap, sweep waves) changes only the
// time of an extraction, never its result, so the tuner simply runs the
// caller's workload under a few candidates and keeps the fastest: first the
// sweep's wave count (it sets the planes per wave and so the halo re-reads
// against the waves in flight), then the block cap of the grid-stride
// kernels at that wave count.  Two one-dimensional passes instead of the
// full product: the two knobs act on different kernels.
//
// Built only on the public C ABI of include/tmc_dmc.h; no counterpart in the
// reference, whose CUDA kernels launch with fixed block sizes
// (dmc/DualMarchingCubes.cu).
#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/tmc_dmc.h"

namespace {

constexpr int32_t kSweepCandidates[] = {8192, 16384, 32768, 65536};
constexpr int32_t kCapCandidates[] = {4096, 8192, 16384, 32768, 65536};

struct Tuner {
    dmc_context* ctx;
    const dmc_volume* vols;
    int32_t nvol;
    int field_type;
    uint32_t flags;
    int reps;
    std::vector<dmc_tune_trial> trials;

    // median wall time (ms) of `reps` extractions under `l`, after one
    // untimed run
    dmc_status time(const dmc_launch& l, double& ms) {
        dmc_status st = dmc_set_launch(ctx, &l);
        if (st != DMC_OK) return st;
        std::vector<dmc_mesh> out(nvol);
        std::vector<double> t;
        for (int r = -1; r < reps; r++) {
            const auto t0 = std::chrono::steady_clock::now();
            st = dmc_extract_batch(ctx, vols, nvol, field_type, flags, out.data());
            const auto t1 = std::chrono::steady_clock::now();
            if (st != DMC_OK) return st;
            if (r >= 0) t.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
        }
        std::sort(t.begin(), t.end());
        ms = t[t.size() / 2];
        trials.push_back(dmc_tune_trial{l, ms});
        return DMC_OK;
    }
};

}  // namespace

dmc_status dmc_autotune(dmc_context* ctx, const dmc_volume* vols, int32_t nvol, int field_type, uint32_t flags,
                        int32_t reps, dmc_tune_trial* trials, int32_t* ntrials, dmc_tune_trial* best) {
    if (!ctx || !vols || nvol < 1 || reps < 1 || reps > 64 || (trials && !ntrials)) return DMC_ERR_INVALID_ARG;
    dmc_launch start{};
    dmc_status st = dmc_get_launch(ctx, &start);
    if (st != DMC_OK) return st;
    Tuner t{ctx, vols, nvol, field_type, flags, reps, {}};
    dmc_launch cur = start;
    double best_ms = 0;
    bool have = false;
    for (int32_t w : kSweepCandidates) {
        dmc_launch l{cur.grid_cap, w};
        double ms = 0;
        if ((st = t.time(l, ms)) != DMC_OK) break;
        if (!have || ms < best_ms) {
            best_ms = ms;
            cur = l;
            have = true;
        }
    }
    const int32_t best_waves = cur.sweep_waves;
    for (int32_t c : kCapCandidates) {
        if (st != DMC_OK) break;
        if (c == start.grid_cap) continue;  // timed in the first pass
        dmc_launch l{c, best_waves};
        double ms = 0;
        if ((st = t.time(l, ms)) != DMC_OK) break;
        if (ms < best_ms) {
            best_ms = ms;
            cur = l;
        }
    }
    if (st != DMC_OK) {
        (void)dmc_set_launch(ctx, &start);
        return st;
    }
    st = dmc_set_launch(ctx, &cur);
    if (trials) {
        const int32_t n = std::min<int32_t>(*ntrials, (int32_t)t.trials.size());
        std::copy(t.trials.begin(), t.trials.begin() + n, trials);
        *ntrials = n;
    } else if (ntrials) {
        *ntrials = (int32_t)t.trials.size();
    }
    if (best) *best = dmc_tune_trial{cur, best_ms};
    return st;
}
