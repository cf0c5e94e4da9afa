// Serving: continuous batching of search requests on one device (no reference
// counterpart; the reference runs one search per GPU, src/service/search.rs:129-192).
//
// Requests are queued on submit and admitted into a device slot when a slot is free and
// their DFA table fits the DFA pool (the HBM budget). Each slot's descriptor, bitap
// table, DFA table, next key index and stop flag live on the device. A step is one
// launch of a fixed number of tiles dealt round-robin over the active slots: the search
// kernel claims each tile's index range from the slot's device counter, stops a slot
// once it has a candidate, and the confirm kernel turns candidates into exact hits.
// After the step the host reads the counters, retires every slot with an exact hit
// (its lowest-index hit of the step is the result) or with its candidate budget spent,
// and admits waiting requests. Each step serves at most active_per_step resident requests
// (16 by default), those with the fewest candidates searched so far (least attained
// service, which favours the short requests of a heavy-tailed mix without knowing their
// sizes). Against 64 per step, 16 found 2516 instead of 2502 of 3000 requests in the 60 s
// Poisson mix at 1.94 instead of 1.93 G keys/s, with a found-only p99 of 2.3 s instead of
// 8.9 s and the same p50 (18 ms). The step -- slot-list upload, counter reset, search,
// confirm, counter and next-index download -- is captured once as a hipGraph and
// replayed; its kernels clear their own counters (no memset nodes).
//
// With `overlap`, two steps are in flight: each of two step buffers has its own stream,
// slot list, counters, candidate and hit lists, inversion spill and graph. A call
// launches the next step and only then waits for the previous one and retires its
// requests, so the next step's waves start on the SIMDs the previous step's last units
// free, and the host's retire / admit work runs while the device searches. The two steps
// serve disjoint requests, so each request's indices are still claimed by one step at a
// time, in order; with every claimed index searched, a request's result stays the first
// exact hit at or after its first index. A retired slot (and its DFA memory) is reused
// only once every step that included it has completed.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "mk_claims.h"
#include "mk_internal.h"

#define MK_MATCHER_FAIL(st, ...) (mk_fail((mk_status)(st), __VA_ARGS__), (st))
#include "mk_matcher_host.h"

#define MK_HIPS(call)                                                                                        \
    do {                                                                                                     \
        hipError_t e_ = (call);                                                                              \
        if (e_ != hipSuccess)                                                                                \
            return mk_fail(MK_ERR_DRIVER, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

namespace {

struct pending {
    u64 id;
    mk_search_spec spec;
    std::vector<u8> needle;  // owned copy (spec.needle points into it)
    double submit;
};

struct slot_state {
    bool used = false;
    u64 last_step = 0;  // the last step that included the slot (it is reused only once that step is done)
    u64 request = 0;
    u64 first = 0, count = 0, tries = 0;
    size_t dfa_off = 0, dfa_n = 0;  // u16 units in the pool
    double submit = 0, admit = 0;
};

// First-fit allocator over the DFA pool (u16 units).
struct pool_alloc {
    std::map<size_t, size_t> free_;  // offset -> size
    void init(size_t n) {
        free_.clear();
        if (n) free_[0] = n;
    }
    bool take(size_t n, size_t& off) {
        if (n == 0) {
            off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void give(size_t off, size_t n) {
        if (n == 0) return;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace

// One of the two step buffers: what a step in flight reads and writes.
struct step_buf {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    u32* spill = nullptr;  // per-lane buffer of the batched inversion
    bool own_spill = false;
    mk_serve_state* d_serve = nullptr;
    mk_counters* ctr = nullptr;
    mk_cand* cands = nullptr;
    mk_hit* hits = nullptr;
    // pinned host
    mk_serve_state* h_serve = nullptr;
    mk_counters* h_ctr = nullptr;
    u64* h_next = nullptr;
    mk_hit* h_hits = nullptr;
    u32 h_hits_cap = 0;
    // graph of one step on this buffer
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t done = nullptr;
    bool in_flight = false;
    u64 id = 0;                                   // step number
    std::vector<std::pair<u32, u64>> active;      // (slot, request) of the step
};

struct mk_server {
    mk_gpu* g = nullptr;
    mk_server_options opt{};
    u64 step_tiles = 0;
    // device
    mk_search_desc* d_descs = nullptr;
    u64* d_masks = nullptr;
    u16* d_dfas = nullptr;
    u32* d_stop = nullptr;
    u64* d_next = nullptr;
    step_buf buf[2];
    u32 cur = 0;           // buffer of the next step
    u64 launched = 0;      // steps launched
    u64 completed = 0;     // steps whose results have been retired
    // host state
    std::vector<slot_state> slots;
    std::deque<pending> queue;
    std::vector<mk_server_result> done;
    pool_alloc pool;
    struct deferred_dfa {
        size_t off, n;
        u64 step;
    };
    std::vector<deferred_dfa> deferred;  // DFA memory of retired requests, freed when their last step is done
    u64 next_id = 1;
    std::chrono::steady_clock::time_point t0;
    u64 steps = 0, keys = 0;
};

namespace {

double now_s(const mk_server* s) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

// A step's device work on buffer b, on the server's stream (eager, or recorded into b's
// graph).
mk_status enqueue_step(mk_server* s, u32 b) {
    mk_gpu* g = s->g;
    step_buf& B = s->buf[b];
    MK_HIPS(hipMemcpyAsync(B.d_serve, B.h_serve, sizeof(mk_serve_state), hipMemcpyHostToDevice, B.stream));
    MK_HIPS(mk_k_reset(B.stream, B.ctr, s->d_stop, MK_SERVE_MAX_SLOTS, B.d_serve));
    MK_HIPS(g->ks->search(B.stream, g->grid, g->lock_waves, g->table, g->opt.width, g->opt.batch, s->d_descs, nullptr, 0, s->step_tiles,
                        s->d_masks, s->d_dfas, B.spill, B.ctr, B.cands, kCandCap, s->d_stop, B.d_serve, s->d_next));
    MK_HIPS(g->ks->confirm(B.stream, g->table, g->opt.width, s->d_descs, s->d_masks, s->d_dfas, B.ctr, B.cands, kCandCap,
                         B.hits, kHitCap));
    MK_HIPS(hipMemcpyAsync(B.h_ctr, B.ctr, sizeof(mk_counters), hipMemcpyDeviceToHost, B.stream));
    MK_HIPS(hipMemcpyAsync(B.h_next, s->d_next, sizeof(u64) * s->opt.max_slots, hipMemcpyDeviceToHost, B.stream));
    return MK_OK;
}

void finish(mk_server* s, u32 slot, int status, const mk_hit* hit, u64 tries) {
    slot_state& st = s->slots[slot];
    mk_server_result r;
    memset(&r, 0, sizeof r);
    r.request = st.request;
    r.status = status;
    r.tries = tries;
    r.submit_s = st.submit;
    r.admit_s = st.admit;
    r.done_s = now_s(s);
    if (hit) {
        r.hit.index = hit->index;
        r.hit.search = 0;
        mk_words_to_bytes(hit->seed, r.hit.seed);
        mk_words_to_bytes(hit->pk, r.hit.pk);
    }
    s->done.push_back(r);
    // a step in flight may still walk this DFA: its pool memory is released once the
    // last step that included the slot has completed
    s->deferred.push_back({st.dfa_off, st.dfa_n, st.last_step});
    const u64 last = st.last_step;
    st = slot_state();
    st.last_step = last;
}

// Move waiting requests into free slots while their DFA fits the pool.
mk_status admit(mk_server* s) {
    mk_gpu* g = s->g;
    for (size_t i = 0; i < s->deferred.size();) {
        if (s->deferred[i].step <= s->completed) {
            s->pool.give(s->deferred[i].off, s->deferred[i].n);
            s->deferred[i] = s->deferred.back();
            s->deferred.pop_back();
        } else {
            i++;
        }
    }
    while (!s->queue.empty()) {
        // a free slot that no step in flight still searches
        u32 slot = 0;
        while (slot < s->opt.max_slots && (s->slots[slot].used || s->slots[slot].last_step > s->completed)) slot++;
        if (slot == s->opt.max_slots) break;
        pending& p = s->queue.front();
        matcher_host m;
        if (build_matcher(p.spec.needle, p.spec.needle_len, p.spec.fold, p.spec.mode, m) != 0) {
            mk_server_result r;
            memset(&r, 0, sizeof r);
            r.request = p.id;
            r.status = -1;
            r.submit_s = p.submit;
            r.admit_s = r.done_s = now_s(s);
            s->done.push_back(r);
            s->queue.pop_front();
            continue;
        }
        size_t off = 0;
        if (!s->pool.take(m.dfa.size(), off)) break;  // wait for memory (FIFO order kept)
        slot_state& st = s->slots[slot];
        st.used = true;
        st.request = p.id;
        st.first = p.spec.first;
        st.count = p.spec.count;
        st.dfa_off = off;
        st.dfa_n = m.dfa.size();
        st.submit = p.submit;
        st.admit = now_s(s);
        mk_search_desc d;
        memset(&d, 0, sizeof d);
        mk_key_to_words(p.spec.key, d.key);
        d.source = (u32)p.spec.seed_source;
        d.mode = m.mode;
        d.len = m.len;
        d.stop_on_hit = 1;
        d.mask_off = slot * 64;
        d.dfa_off = (u32)off;
        // synchronous copies from pageable memory: the sources go away when the request does
        MK_HIPS(hipMemcpy(s->d_descs + slot, &d, sizeof d, hipMemcpyHostToDevice));
        MK_HIPS(hipMemcpy(s->d_masks + (size_t)slot * 64, m.mask, sizeof m.mask, hipMemcpyHostToDevice));
        if (!m.dfa.empty())
            MK_HIPS(hipMemcpy(s->d_dfas + off, m.dfa.data(), m.dfa.size() * sizeof(u16), hipMemcpyHostToDevice));
        const u64 first = p.spec.first;
        MK_HIPS(hipMemcpy(s->d_next + slot, &first, sizeof first, hipMemcpyHostToDevice));
        // the host snapshots are refreshed by every step; start them right
        for (step_buf& B : s->buf) B.h_next[slot] = first;
        s->queue.pop_front();
        (void)g;
    }
    return MK_OK;
}

// Wait for step buffer b's step and retire its requests: each with an exact hit (its
// lowest-index hit of the step) or with its candidate budget spent.
mk_status retire_step(mk_server* s, u32 b) {
    mk_gpu* g = s->g;
    step_buf& B = s->buf[b];
    MK_HIPS(hipEventSynchronize(B.done));
    B.in_flight = false;
    s->steps++;
    s->completed = std::max(s->completed, B.id);
    const u32 nh = std::min(B.h_ctr->n_hits, kHitCap);
    if (nh > B.h_hits_cap) {
        if (B.h_hits) MK_HIPS(hipHostFree(B.h_hits));
        B.h_hits = nullptr;
        B.h_hits_cap = 0;
        const u32 want = std::max<u32>(nh, 4096);
        MK_HIPS(hipHostMalloc((void**)&B.h_hits, sizeof(mk_hit) * want, hipHostMallocDefault));
        B.h_hits_cap = want;
    }
    if (nh) {
        // on the null stream: it does not wait for the step queued behind on the server's stream
        MK_HIPS(hipMemcpy(B.h_hits, B.hits, sizeof(mk_hit) * nh, hipMemcpyDeviceToHost));
    }
    (void)g;
    // lowest-index exact hit of each slot in this step. The search kernel appends, for a
    // stop-on-hit search, each lane's lowest candidate whose match holds for either x sign
    // and the few whose match depends on it, so even a pattern that matches every key
    // appends about one candidate per lane of the tiles claimed before its stop flag is
    // seen. A full candidate list would still drop candidates but not their slots' stop
    // flags: such a slot could then finish on a later hit, exact and confirmed but not the
    // lowest of the range searched.
    std::map<u32, mk_hit> best;
    for (u32 i = 0; i < nh; i++) {
        const mk_hit& h = B.h_hits[i];
        if (h.search >= s->opt.max_slots || !s->slots[h.search].used) continue;
        auto it = best.find(h.search);
        if (it == best.end() || h.index < it->second.index) best[h.search] = h;
    }
    for (const auto& sr : B.active) {
        const u32 slot = sr.first;
        slot_state& ss = s->slots[slot];
        // retired (or cancelled) since this step was launched: its hits here are stale
        if (!ss.used || ss.request != sr.second) continue;
        const u64 tries = B.h_next[slot] - ss.first;
        ss.tries = tries;
        auto it = best.find(slot);
        if (it != best.end()) {
            finish(s, slot, 1, &it->second, tries);
        } else if (ss.count && tries >= ss.count) {
            finish(s, slot, 0, nullptr, tries);
        }
    }
    return MK_OK;
}

}  // namespace

extern "C" {

void mk_server_default_options(mk_server_options* out) {
    if (!out) return;
    out->max_slots = 1024;
    out->use_graph = 1;
    out->dfa_budget = 64ull << 20;
    out->step_keys = 1ull << 26;
    out->active_per_step = 16;
    out->overlap = 1;
    out->unit_keys = MK_SERVE_UNIT_KEYS;
}

mk_status mk_server_open(mk_gpu* g, const mk_server_options* opt_in, mk_server** out) {
    if (!g || !out) return mk_fail(MK_ERR_INVALID, "null pointer");
    *out = nullptr;
    mk_server_options opt;
    mk_server_default_options(&opt);
    if (opt_in) opt = *opt_in;
    if (opt.max_slots < 1 || opt.max_slots > MK_SERVE_MAX_SLOTS)
        return mk_fail(MK_ERR_OUT_OF_RANGE, "max_slots must be 1-%d (got %u)", MK_SERVE_MAX_SLOTS, opt.max_slots);
    if (opt.unit_keys == 0) opt.unit_keys = MK_SERVE_UNIT_KEYS;
    if (opt.unit_keys % 64 != 0 || opt.unit_keys > 256)
        return mk_fail(MK_ERR_OUT_OF_RANGE, "unit_keys must be 64, 128, 192 or 256 (got %u)", opt.unit_keys);
    MK_HIPS(hipSetDevice(g->device));
    mk_server* s = new mk_server();
    s->g = g;
    s->opt = opt;
    const u64 tile_keys = 64ull * g->opt.batch;
    // The search kernel claims a serving step in units of unit_keys keys per lane (a tile
    // of 256 keys per lane is two 128-key units), so a step needs one unit per wave at least.
    const u64 units_per_tile = mk_units_per_tile(g->opt.batch, opt.unit_keys);
    const u64 min_tiles = (g->grid + units_per_tile - 1) / units_per_tile;
    s->step_tiles = std::max<u64>(min_tiles, (opt.step_keys + tile_keys - 1) / tile_keys);
    s->slots.resize(opt.max_slots);
    const size_t dfa_units = std::max<size_t>(opt.dfa_budget / sizeof(u16), 65);
    s->pool.init(dfa_units);
    s->t0 = std::chrono::steady_clock::now();
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_descs, sizeof(mk_search_desc) * opt.max_slots);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_masks, sizeof(u64) * 64 * opt.max_slots);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_dfas, sizeof(u16) * dfa_units);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_stop, sizeof(u32) * MK_SERVE_MAX_SLOTS);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_next, sizeof(u64) * opt.max_slots);
    s->buf[0].stream = g->stream;
    s->buf[0].spill = g->spill;
    if (opt.overlap) {
        step_buf& B = s->buf[1];
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&B.stream, hipStreamNonBlocking);
        if (e == hipSuccess) B.own_stream = true;
        if (e == hipSuccess) e = hipMalloc((void**)&B.spill, g->spill_bytes);
        if (e == hipSuccess) B.own_spill = true;
    } else {
        s->buf[1].stream = g->stream;
        s->buf[1].spill = g->spill;
    }
    for (int b = 0; b < 2; b++) {
        step_buf& B = s->buf[b];
        if (e == hipSuccess) e = hipMalloc((void**)&B.d_serve, sizeof(mk_serve_state));
        if (e == hipSuccess) e = hipMalloc((void**)&B.ctr, sizeof(mk_counters));
        if (e == hipSuccess) e = hipMalloc((void**)&B.cands, sizeof(mk_cand) * kCandCap);
        if (e == hipSuccess) e = hipMalloc((void**)&B.hits, sizeof(mk_hit) * kHitCap);
        if (e == hipSuccess) e = hipHostMalloc((void**)&B.h_serve, sizeof(mk_serve_state), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&B.h_ctr, sizeof(mk_counters), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&B.h_next, sizeof(u64) * opt.max_slots, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&B.done, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        mk_server_close(s);
        return mk_fail(MK_ERR_DRIVER, "server allocation: %s", hipGetErrorString(e));
    }
    for (step_buf& B : s->buf) {
        memset(B.h_serve, 0, sizeof(mk_serve_state));
        B.h_serve->unit_keys = opt.unit_keys;
        memset(B.h_ctr, 0, sizeof(mk_counters));
        memset(B.h_next, 0, sizeof(u64) * opt.max_slots);
    }
    std::vector<u32> zeros(MK_SERVE_MAX_SLOTS, 0);
    e = hipMemcpy(s->d_stop, zeros.data(), sizeof(u32) * MK_SERVE_MAX_SLOTS, hipMemcpyHostToDevice);
    for (u32 b = 0; b < 2 && e == hipSuccess && opt.use_graph; b++) {
        step_buf& B = s->buf[b];
        e = hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            mk_status st = enqueue_step(s, b);
            hipError_t e2 = hipStreamEndCapture(B.stream, &B.graph);
            if (st != MK_OK) {
                mk_server_close(s);
                return st;
            }
            e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0);
    }
    if (e != hipSuccess) {
        mk_server_close(s);
        return mk_fail(MK_ERR_DRIVER, "server graph: %s", hipGetErrorString(e));
    }
    *out = s;
    return MK_OK;
}

double mk_server_now(const mk_server* s) { return s ? now_s(s) : 0.0; }

mk_status mk_server_submit(mk_server* s, const mk_search_spec* spec, uint64_t* request) {
    if (!s || !spec) return mk_fail(MK_ERR_INVALID, "null pointer");
    if (spec->seed_source != MK_SEED_CHACHA20 && spec->seed_source != MK_SEED_COUNTER)
        return mk_fail(MK_ERR_INVALID, "unknown seed source %d", spec->seed_source);
    pending p;
    p.id = s->next_id++;
    p.spec = *spec;
    size_t bytes = spec->mode == MK_MODE_REGEX ? (size_t)spec->needle_len * 130 : spec->needle_len;
    if (bytes && !spec->needle) return mk_fail(MK_ERR_INVALID, "null needle");
    p.needle.assign(spec->needle, spec->needle + bytes);
    p.submit = now_s(s);
    s->queue.push_back(std::move(p));
    s->queue.back().spec.needle = s->queue.back().needle.data();
    if (request) *request = s->queue.back().id;
    return MK_OK;
}

mk_status mk_server_step(mk_server* s, uint32_t* active_out, uint32_t* queued_out) {
    if (!s) return mk_fail(MK_ERR_INVALID, "null pointer");
    mk_gpu* g = s->g;
    MK_HIPS(hipSetDevice(g->device));
    mk_status st = admit(s);
    if (st != MK_OK) return st;
    // least attained service: the resident requests with the fewest candidates so far
    // (with overlap, not those the step in flight serves: the two steps take disjoint requests)
    std::vector<std::pair<u64, u32>> order;
    for (u32 i = 0; i < s->opt.max_slots; i++)
        if (s->slots[i].used && s->slots[i].last_step <= s->completed) order.emplace_back(s->slots[i].tries, i);
    const size_t cap = s->opt.active_per_step ? s->opt.active_per_step : order.size();
    if (order.size() > cap) {
        std::nth_element(order.begin(), order.begin() + cap, order.end());
        order.resize(cap);
    }
    std::sort(order.begin(), order.end(), [](const std::pair<u64, u32>& a, const std::pair<u64, u32>& b) {
        return a.second < b.second;
    });
    const u32 n = (u32)order.size();
    if (active_out) *active_out = n;
    if (queued_out) *queued_out = (u32)s->queue.size();
    const u32 prev = s->cur ^ 1u;
    if (n > 0) {
        // buffer `cur` was last used two steps ago; retire_step has waited for it
        step_buf& B = s->buf[s->cur];
        if (B.in_flight) {
            st = retire_step(s, s->cur);
            if (st != MK_OK) return st;
        }
        B.active.clear();
        for (u32 i = 0; i < n; i++) {
            const u32 slot = order[i].second;
            B.h_serve->slots[i] = slot;
            B.active.emplace_back(slot, s->slots[slot].request);
        }
        B.h_serve->n_active = n;
        B.id = ++s->launched;
        for (u32 i = 0; i < n; i++) s->slots[order[i].second].last_step = B.id;
        if (B.exec) {
            MK_HIPS(hipGraphLaunch(B.exec, B.stream));
        } else {
            st = enqueue_step(s, s->cur);
            if (st != MK_OK) return st;
        }
        MK_HIPS(hipEventRecord(B.done, B.stream));
        B.in_flight = true;
        if (!s->opt.overlap) return retire_step(s, s->cur);
        s->cur ^= 1u;
    }
    // the previous step ran while this one was queued: retire its requests now (with no
    // new step, the one just before)
    if (s->buf[prev].in_flight) return retire_step(s, prev);
    if (n == 0 && s->buf[s->cur].in_flight) return retire_step(s, s->cur);
    return MK_OK;
}

mk_status mk_server_poll(mk_server* s, mk_server_result* out, uint32_t cap, uint32_t* n) {
    if (!s || !n || (cap && !out)) return mk_fail(MK_ERR_INVALID, "null pointer");
    const u32 k = (u32)std::min<size_t>(cap, s->done.size());
    for (u32 i = 0; i < k; i++) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return MK_OK;
}

mk_status mk_server_cancel(mk_server* s, uint64_t request) {
    if (!s) return mk_fail(MK_ERR_INVALID, "null pointer");
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it)
        if (it->id == request) {
            mk_server_result r;
            memset(&r, 0, sizeof r);
            r.request = request;
            r.status = -1;
            r.submit_s = it->submit;
            r.admit_s = r.done_s = now_s(s);
            s->done.push_back(r);
            s->queue.erase(it);
            return MK_OK;
        }
    for (u32 i = 0; i < s->opt.max_slots; i++)
        if (s->slots[i].used && s->slots[i].request == request) {
            finish(s, i, -1, nullptr, s->slots[i].tries);  // as of the last retired step
            return MK_OK;
        }
    return mk_fail(MK_ERR_INVALID, "no request %llu", (unsigned long long)request);
}

mk_status mk_server_stats(const mk_server* s, uint64_t* steps, uint64_t* step_keys) {
    if (!s) return mk_fail(MK_ERR_INVALID, "null pointer");
    if (steps) *steps = s->steps;
    if (step_keys) *step_keys = s->step_tiles * 64ull * s->g->opt.batch;
    return MK_OK;
}

mk_status mk_server_close(mk_server* s) {
    if (!s) return MK_OK;
    (void)hipSetDevice(s->g->device);
    (void)hipStreamSynchronize(s->g->stream);
    for (step_buf& B : s->buf)
        if (B.own_stream) (void)hipStreamSynchronize(B.stream);
    for (step_buf& B : s->buf) {
        if (B.own_stream) (void)hipStreamDestroy(B.stream);
        if (B.own_spill) (void)hipFree(B.spill);
        if (B.exec) (void)hipGraphExecDestroy(B.exec);
        if (B.graph) (void)hipGraphDestroy(B.graph);
        if (B.done) (void)hipEventDestroy(B.done);
        if (B.d_serve) (void)hipFree(B.d_serve);
        if (B.ctr) (void)hipFree(B.ctr);
        if (B.cands) (void)hipFree(B.cands);
        if (B.hits) (void)hipFree(B.hits);
        if (B.h_serve) (void)hipHostFree(B.h_serve);
        if (B.h_ctr) (void)hipHostFree(B.h_ctr);
        if (B.h_next) (void)hipHostFree(B.h_next);
        if (B.h_hits) (void)hipHostFree(B.h_hits);
    }
    if (s->d_descs) (void)hipFree(s->d_descs);
    if (s->d_masks) (void)hipFree(s->d_masks);
    if (s->d_dfas) (void)hipFree(s->d_dfas);
    if (s->d_stop) (void)hipFree(s->d_stop);
    if (s->d_next) (void)hipFree(s->d_next);
    delete s;
    return MK_OK;
}

}  // extern "C"
