// The search driver and key output (host side of the C ABI):
//  * mk_search: the reference's Searcher::run for one GPU (gpu.rs:999-1012, 917-968) --
//    a fresh random 32-byte key, prepare, then mk_search_with over the device engine;
//  * OpenSSH text: the authorized_keys line (key.rs:132-139) and the unencrypted
//    openssh-key-v1 private key (key_file.rs:276-320).
#include <hip/hip_runtime.h>
#include <sys/random.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mk_internal.h"
#include "mk_search.h"

// mk_confirm_host.cpp (host compiler): 1 when the key at `index` re-derives from the search
// key to `seed` and `pk` and matches the pattern with its x sign, 0 when not, -1 bad pattern
extern "C" __attribute__((visibility("hidden"))) int mk_confirm_emitted(
    const uint8_t key[32], int source, uint64_t index, const uint8_t seed[32], const uint8_t pk[32],
    const uint8_t* needle, uint32_t len, uint32_t fold, uint32_t mode);

namespace {

const char kB64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

std::string b64_encode(const std::vector<uint8_t>& in) {
    std::string out;
    size_t i = 0;
    for (; i + 3 <= in.size(); i += 3) {
        const uint32_t n = (uint32_t)in[i] << 16 | (uint32_t)in[i + 1] << 8 | in[i + 2];
        for (int k = 0; k < 4; k++) out += kB64[(n >> (18 - 6 * k)) & 63];
    }
    const size_t rem = in.size() - i;
    if (rem) {
        uint32_t n = (uint32_t)in[i] << 16 | (rem == 2 ? (uint32_t)in[i + 1] << 8 : 0u);
        out += kB64[(n >> 18) & 63];
        out += kB64[(n >> 12) & 63];
        out += rem == 2 ? kB64[(n >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

// RFC 4251 wire format
void put_u32(std::vector<uint8_t>& b, uint32_t v) {
    for (int k = 3; k >= 0; k--) b.push_back((uint8_t)(v >> (8 * k)));
}
void put_string(std::vector<uint8_t>& b, const uint8_t* p, size_t n) {
    put_u32(b, (uint32_t)n);
    b.insert(b.end(), p, p + n);
}
void put_string(std::vector<uint8_t>& b, const char* s) { put_string(b, (const uint8_t*)s, strlen(s)); }

std::vector<uint8_t> public_blob(const uint8_t pk[32]) {
    std::vector<uint8_t> b;
    put_string(b, "ssh-ed25519");
    put_string(b, pk, 32);
    return b;
}

mk_status copy_out(const std::string& s, char* out, size_t cap, size_t* len) {
    if (len) *len = s.size();
    if (!out) return MK_OK;
    if (cap < s.size() + 1) return mk_fail(MK_ERR_OVERFLOW, "output needs %zu bytes", s.size() + 1);
    memcpy(out, s.c_str(), s.size() + 1);
    return MK_OK;
}

// check_comment (key.rs:64-75): no line breaks and no NUL
bool comment_ok(const char* c) { return !strchr(c, '\n') && !strchr(c, '\r'); }

struct steady_clock_s {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

struct device_engine {
    mk_gpu* g;
    mk_launch* l;
    std::vector<mk_hit_rec> hits;
    mk_hit_rec first_hit{};
    uint64_t keys_per_iter() const { return mk_keys_per_iter(g); }
    int launch(uint64_t base, uint32_t iters, int& found, uint64_t& index) {
        uint32_t n = 0;
        mk_status st = mk_launch_run(l, base, iters, hits.data(), (uint32_t)hits.size(), &n);
        if (st != MK_OK && st != MK_ERR_OVERFLOW) return st;
        found = n > 0;
        if (found) {
            first_hit = hits[0];
            index = hits[0].index;
        }
        return MK_OK;
    }
};

struct callback_progress {
    mk_progress_fn fn;
    void* user;
    void report(uint64_t tries, double elapsed) {
        if (!fn) return;
        mk_stats s;
        s.tries = tries;
        s.elapsed_s = elapsed;
        fn(user, &s);
    }
};

struct flag_cancel {
    const volatile int32_t* flag;
    bool cancelled() const { return flag && *flag != 0; }
};

}  // namespace

extern "C" {

mk_status mk_search(mk_gpu* g, const mk_search_config* cfg, mk_progress_fn progress, void* user, mk_outcome* out) {
    if (!g || !cfg || !out) return mk_fail(MK_ERR_INVALID, "null pointer");
    memset(out, 0, sizeof *out);
    uint8_t key[32];
    if (cfg->key) {
        memcpy(key, cfg->key, 32);
    } else if (getrandom(key, sizeof key, 0) != (ssize_t)sizeof key) {
        return mk_fail(MK_ERR_DRIVER, "getrandom failed");
    }
    const uint8_t* needle = nullptr;
    uint32_t len = 0, fold = 0, mode = MK_MODE_CONTAINS;
    if (cfg->pattern) {
        mk_status st = mk_pattern_needle(cfg->pattern, &needle, &len, &fold, &mode);
        if (st != MK_OK) return mk_fail(st, "pattern");
    }
    mk_launch* l = nullptr;
    mk_status st = mk_prepare(g, key, cfg->seed_source, needle, len, fold, mode, cfg->bench ? 0u : 1u, &l);
    if (st != MK_OK) return st;
    device_engine eng{g, l, std::vector<mk_hit_rec>(1u << 12), {}};
    steady_clock_s clock;
    callback_progress prog{progress, user};
    flag_cancel cancel{cfg->cancel};
    const double target = cfg->target_launch_s > 0 ? cfg->target_launch_s : 0.2;
    const double interval = cfg->progress_interval_s > 0 ? cfg->progress_interval_s : 5.0;
    mk_search_result r = mk_search_with(eng, clock, prog, cancel, cfg->bench != 0, target, cfg->time_limit_s, interval);
    mk_free_launch(l);
    out->stats.tries = r.tries;
    out->stats.elapsed_s = r.elapsed;
    out->launches = r.launches;
    memcpy(out->key, key, 32);
    if (r.status != 0) return (mk_status)r.status;
    if (r.found) {
        out->found = 1;
        out->index = r.index;
        memcpy(out->seed, eng.first_hit.seed, 32);
        memcpy(out->pk, eng.first_hit.pk, 32);
        // the reference's confirm: the key is re-derived on the host before it is handed out
        if (mk_confirm_emitted(key, cfg->seed_source, r.index, out->seed, out->pk, needle, len, fold, mode) != 1) {
            out->found = 0;
            return mk_fail(MK_ERR_SELF_CHECK, "the device's key at index %llu does not re-derive on the host",
                           (unsigned long long)r.index);
        }
    }
    return MK_OK;
}

mk_status mk_openssh_public(const uint8_t pk[32], const char* comment, char* out, size_t cap, size_t* len) {
    if (!pk) return mk_fail(MK_ERR_INVALID, "null pointer");
    if (comment && !comment_ok(comment)) return mk_fail(MK_ERR_INVALID, "invalid character in the comment");
    std::string s = "ssh-ed25519 " + b64_encode(public_blob(pk));
    if (comment && *comment) s += std::string(" ") + comment;
    return copy_out(s, out, cap, len);
}

mk_status mk_openssh_private(const uint8_t seed[32], const uint8_t pk[32], const char* comment, uint32_t checkint,
                             char* out, size_t cap, size_t* len) {
    if (!seed || !pk) return mk_fail(MK_ERR_INVALID, "null pointer");
    if (!comment) comment = "";
    if (!comment_ok(comment)) return mk_fail(MK_ERR_INVALID, "invalid character in the comment");
    std::vector<uint8_t> priv;
    put_u32(priv, checkint);
    put_u32(priv, checkint);
    put_string(priv, "ssh-ed25519");
    put_string(priv, pk, 32);
    uint8_t pair[64];
    memcpy(pair, seed, 32);
    memcpy(pair + 32, pk, 32);
    put_string(priv, pair, 64);
    put_string(priv, comment);
    for (uint8_t pad = 1; priv.size() % 8 != 0; pad++) priv.push_back(pad);
    std::vector<uint8_t> body((const uint8_t*)"openssh-key-v1", (const uint8_t*)"openssh-key-v1" + 15);
    put_string(body, "none");
    put_string(body, "none");
    put_string(body, (const uint8_t*)"", 0);
    put_u32(body, 1);
    const std::vector<uint8_t> blob = public_blob(pk);
    put_string(body, blob.data(), blob.size());
    put_string(body, priv.data(), priv.size());
    const std::string b64 = b64_encode(body);
    std::string pem = "-----BEGIN OPENSSH PRIVATE KEY-----\n";
    for (size_t i = 0; i < b64.size(); i += 70) pem += b64.substr(i, 70) + "\n";
    pem += "-----END OPENSSH PRIVATE KEY-----\n";
    memset(pair, 0, sizeof pair);
    std::fill(priv.begin(), priv.end(), 0);
    std::fill(body.begin(), body.end(), 0);
    mk_status st = copy_out(pem, out, cap, len);
    std::fill(pem.begin(), pem.end(), '\0');
    return st;
}

}  // extern "C"

// measure (gpu.rs:889-911): counter seeds, needle "meikey" case-insensitive, no stop on
// hits; one short launch sizes the iterations so each of three timed launches takes about
// 100 ms; the rate is keys / time of the three.
static mk_status measure(mk_gpu* g, double* rate) {
    uint8_t key[32];
    memset(key, 0x5a, sizeof key);
    mk_launch* l = nullptr;
    mk_status st = mk_prepare(g, key, MK_SEED_COUNTER, (const uint8_t*)"meikey", 6, 1, MK_MODE_CONTAINS, 0, &l);
    if (st != MK_OK) return st;
    std::vector<mk_hit_rec> hits(1u << 12);
    uint32_t n = 0;
    auto t0 = std::chrono::steady_clock::now();
    st = mk_launch_run(l, 0, 1, hits.data(), (uint32_t)hits.size(), &n);
    const double once = std::max(1e-6, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    const uint32_t iters = (uint32_t)std::min(65536.0, std::max(1.0, 0.1 / once));
    const uint64_t kpi = mk_keys_per_iter(g);
    t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < 3 && (st == MK_OK || st == MK_ERR_OVERFLOW); i++)
        st = mk_launch_run(l, (uint64_t)(i + 1) * kpi * iters, iters, hits.data(), (uint32_t)hits.size(), &n);
    const double dt = std::max(1e-9, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
    mk_free_launch(l);
    if (st != MK_OK && st != MK_ERR_OVERFLOW) return st;
    *rate = (double)kpi * iters * 3 / dt;
    return MK_OK;
}

extern "C" {

mk_status mk_fastest(int ordinal, const mk_options* candidates, uint32_t n, mk_tune_result* results, mk_options* best) {
    if ((n && !candidates) || !best) return mk_fail(MK_ERR_INVALID, "null pointer");
    double best_rate = -1;
    for (uint32_t i = 0; i < n; i++) {
        mk_tune_result r;
        memset(&r, 0, sizeof r);
        r.options = candidates[i];
        mk_gpu* g = nullptr;
        r.status = mk_open(ordinal, &candidates[i], &g);
        if (r.status == MK_OK) {
            // warm-up, then the measurement (gpu.rs:1041-1056 measures each setting after a warm-up)
            double warm = 0;
            r.status = measure(g, &warm);
            if (r.status == MK_OK) r.status = measure(g, &r.keys_per_s);
            mk_close(g);
        }
        if (results) results[i] = r;
        if (r.status == MK_OK && r.keys_per_s > best_rate) {
            best_rate = r.keys_per_s;
            *best = candidates[i];
        }
    }
    if (best_rate < 0) return mk_fail(MK_ERR_SELF_CHECK, "no setting could be measured");
    return MK_OK;
}

mk_status mk_fastest_staged(int ordinal, mk_options* best, mk_tune_result* results, uint32_t cap, uint32_t* n_out) {
    if (!best) return mk_fail(MK_ERR_INVALID, "null pointer");
    std::vector<mk_tune_result> all;
    auto stage = [&](const std::vector<mk_options>& cands, mk_options& won) -> mk_status {
        std::vector<mk_tune_result> rs(cands.size());
        mk_status st = mk_fastest(ordinal, cands.data(), (uint32_t)cands.size(), rs.data(), &won);
        all.insert(all.end(), rs.begin(), rs.end());
        return st;
    };
    mk_options base;
    mk_default_options(&base);
    // stage 1: the field / table layouts (gpu.rs:364-370 measures every Limbs first) at
    // W = 16 and the default keys per lane and waves per SIMD
    std::vector<mk_options> c0;
    for (uint32_t limbs : {(uint32_t)MK_LIMBS_R25U, (uint32_t)MK_LIMBS_R25, (uint32_t)MK_LIMBS_R32,
                           (uint32_t)MK_LIMBS_K26U, (uint32_t)MK_LIMBS_K26}) {
        mk_options o = base;
        o.width = 16;
        o.limbs = limbs;
        c0.push_back(o);
    }
    mk_options w0 = base;
    mk_status st = stage(c0, w0);
    // stage 2: comb widths with that layout, gathered from HBM / L2, and the W = 5 table
    // held in LDS
    std::vector<mk_options> c1;
    for (uint32_t w : {16u, 20u, 22u, 24u}) {
        mk_options o = base;
        o.limbs = w0.limbs;
        o.width = w;
        c1.push_back(o);
    }
    {
        mk_options o = base;
        o.limbs = w0.limbs;
        o.width = 5;
        o.blocks = 2;
        o.lds_table = 1;
        c1.push_back(o);
    }
    mk_options w1 = w0;
    if (st == MK_OK) st = stage(c1, w1);
    // stage 3: the inversion at the best width: keys per lane, each lane inverting its own
    // batch, or one inversion shared by the 4 waves of a workgroup (8 keys per lane, the
    // reference's pooled shape, and 32; measured on MI355X, 4-wave pools beat 8-wave ones)
    std::vector<mk_options> c2;
    for (uint32_t b : {32u, 64u, 128u, 256u}) {
        if (b == w1.batch) continue;
        mk_options o = w1;
        o.batch = b;
        c2.push_back(o);
    }
    for (uint32_t b : {8u, 32u}) {
        if (w1.lds_table) break;  // the LDS-table kernel inverts per lane
        mk_options o = w1;
        o.batch = b;
        o.blocks = 2;
        o.lockstep = 0;
        o.pool = 4;
        c2.push_back(o);
    }
    mk_options w2 = w1;
    if (st == MK_OK) {
        c2.push_back(w1);
        st = stage(c2, w2);
    }
    // stage 4: resident waves per SIMD, and workgroups that add each comb row in lockstep
    // (a pooled winner tries 8-wave pools and one wave per SIMD instead)
    mk_options w3 = w2;
    if (st == MK_OK && !w2.lds_table) {
        mk_options o = w2, l = w2;
        if (w2.pool != 0) {
            o.pool = 8;
            l.blocks = 1;
        } else {
            o.blocks = w2.blocks == 2 ? 1 : 2;
            l.lockstep = 2;
        }
        st = stage({w2, o, l}, w3);
    }
    if (n_out) *n_out = (uint32_t)all.size();
    if (results)
        for (size_t i = 0; i < all.size() && i < cap; i++) results[i] = all[i];
    if (st == MK_OK) *best = w3;
    return st;
}

}  // extern "C"
