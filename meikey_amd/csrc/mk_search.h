// The search driver above the Engine (the reference's search_with, gpu.rs:917-968): launch
// the engine over consecutive index ranges, adapting the iterations per launch so each
// launch takes about `target` seconds, until a hit, a cancel, or the time limit.
//
// Header-only and independent of HIP: the C ABI instantiates it with the device engine
// (mk_prepare / mk_launch_run), and the CPU test build with a scripted engine and a
// scripted clock, so the control loop is tested the way the reference tests it with a
// mock Engine (gpu.rs:3219-3364).
#pragma once

#include <stdint.h>

#include <algorithm>

// Engine: keys_per_iter(); launch(base, iters, found, index) -> status (0 = ok).
// Clock: now() -> seconds. Progress: report(tries, elapsed). Cancel: cancelled() -> bool.
struct mk_search_result {
    int status;       // 0 ok, else the engine's failing status
    int found;        // 1 when `index` is a hit
    uint64_t index;
    uint64_t tries;   // keys searched
    double elapsed;   // seconds
    uint64_t launches;
};

// Launch-size controller constants (gpu.rs:427, 964-966): scale the iterations by
// target / took, by a factor in [0.5, 4], between 1 and 2^16.
#define MK_MAX_ITERS (1u << 16)

template <class Engine, class Clock, class Progress, class Cancel>
mk_search_result mk_search_with(Engine& engine, Clock& clock, Progress& progress, Cancel& cancel, bool bench,
                                double target, double time_limit, double progress_interval) {
    mk_search_result r = {0, 0, 0, 0, 0.0, 0};
    const double start = clock.now();
    uint32_t iters = 1;
    double next_report = start + progress_interval;
    for (;;) {
        const double launch_start = clock.now();
        int found = 0;
        uint64_t index = 0;
        const int st = engine.launch(r.tries, iters, found, index);
        if (st != 0) {
            r.status = st;
            r.elapsed = clock.now() - start;
            return r;
        }
        const double took = clock.now() - launch_start;
        r.launches++;
        r.tries += engine.keys_per_iter() * (uint64_t)iters;
        r.elapsed = clock.now() - start;
        if (found && !bench) {
            r.found = 1;
            r.index = index;
            return r;
        }
        if (cancel.cancelled() || (time_limit > 0 && r.elapsed >= time_limit)) return r;
        if (clock.now() >= next_report) {
            progress.report(r.tries, r.elapsed);
            next_report += progress_interval;
        }
        const double scale = std::min(4.0, std::max(0.5, target / std::max(took, 1e-6)));
        const double next = (double)iters * scale;
        iters = (uint32_t)std::min<double>(MK_MAX_ITERS, std::max(1.0, next));
    }
}
