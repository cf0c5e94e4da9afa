// Internal definitions shared by the C-ABI translation units (mk_host.cpp, mk_serve.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/meikey_hip.h"
#include "mk_hitsort.h"
#include "mk_launch.h"

// Records the text of a failure for mk_last_error() and returns st.
mk_status mk_fail(mk_status st, const char* fmt, ...);

// One step's fixed device lists: candidates of the search kernel (16 B each) and exact
// hits of the confirm kernel (80 B each).
constexpr u32 kCandCap = 1u << 20;
constexpr u32 kHitCap = 1u << 20;

// Device buffer that grows on demand.
struct dbuf {
    void* p = nullptr;
    size_t cap = 0;
    mk_status ensure(size_t bytes) {
        if (bytes <= cap) return MK_OK;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            cap = 0;
            if (e != hipSuccess) return mk_fail(MK_ERR_DRIVER, "hipFree: %s", hipGetErrorString(e));
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return mk_fail(MK_ERR_DRIVER, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
        cap = want;
        return MK_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct mk_gpu {
    int device = 0;
    hipStream_t stream = nullptr;
    mk_options opt{};
    const mk_kernel_set* ks = nullptr;  // the kernels of opt.limbs
    u32 cus = 0, grid = 0, rows = 0, cols = 0;
    u32 lock_waves = 0;  // waves per workgroup of the lockstep search kernel (0: one-wave workgroups)
    u64 lanes = 0, hbm = 0;
    char name[64] = {};
    u32* table = nullptr;
    u64 table_bytes = 0;
    u32* spill = nullptr;  // per-lane buffer of the batched inversion
    u32* noalloc = nullptr;  // this device's staging array of mk_derive_noalloc (not owned)
    u64 spill_bytes = 0;
    mk_counters* ctr = nullptr;
    mk_cand* cands = nullptr;
    mk_hit* hits = nullptr;
    mk_counters* h_ctr = nullptr;  // pinned
    mk_hit* h_hits = nullptr;      // pinned
    // per-call pools of the batched path
    dbuf descs, masks, dfas, segs, stop;
    std::vector<u8> staging;
    mk_sort_pool sort;  // host buffers of the hit sort, kept between calls
    // events around the last step's search and confirm kernels (on `stream`)
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    float last_search_ms = 0.f, last_confirm_ms = 0.f;
};


// One step: clear the counters, run the search kernel over the segments (or the serving
// slots), confirm the candidates, and copy counters and exact hits back (unsorted).
mk_status mk_run_step(mk_gpu* g, const mk_search_desc* descs, const u64* masks, const u16* dfas, const mk_seg* segs,
                      u32 nsegs, u64 total_tiles, u32* stop, u32 nstop, const mk_hit*& hits, u32& n_cands,
                      u32& n_hits, u32& overflow);

void mk_words_to_bytes(const u32 w[8], uint8_t out[32]);
void mk_key_to_words(const uint8_t key[32], u32 w[8]);
