// CPU build of the kernel arithmetic, for tests without a GPU.
//
// The headers mk_field.h, mk_hash.h, mk_point.h, mk_match.h and mk_tile.h are the
// exact code the gfx950 kernels run; compiled here with g++ they expose test hooks so
// the CPU test suite can compare every stage (seeds, scalar, field operations at their
// limb bounds, digits, table entries, the search lane's batched inversion, matching)
// with the oracle, and the host-side hit order (mk_hitsort.h). This library is test infrastructure: nothing in the product loads it.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "mk_claims.h"
#include "mk_field.h"
#include "mk_hash.h"
#include "mk_match.h"
#include "mk_point.h"
#include "mk_confirm.h"
#include "mk_tile.h"
#include "mk_search.h"
#include "mk_hitsort.h"

#define MK_MATCHER_FAIL(st, ...) (fprintf(stderr, __VA_ARGS__), fputc('\n', stderr), (st))
#include "mk_matcher_host.h"

extern "C" {

void hc_seed(const u32 key[8], u32 source, u64 index, u32 seed[8]) {
    u32 upper[8];
    mk_seed_at(key, source, index, upper, true, seed);
}

void hc_secret_scalar(const u32 seed[8], u32 s[8]) { mk_secret_scalar(seed, s); }

// The counter-seed scalar through the per-search SHA-512 mid-state.
void hc_secret_scalar_counter(const u32 key[8], u64 index, u32 s[8]) {
    mk_sha_mid mid;
    mk_sha_mid_init(key, mid);
    const u64 t = ((u64)key[7] << 32 | key[6]) + index;
    mk_secret_scalar_mid(mid, mk_be_word((u32)t, (u32)(t >> 32)), s);
}

// The build's field and table layout: 10 * MK_FIELD + MK_PACKED (100 "10u", 101 "10", 81 "8",
// 261 "10k", 260 "10ku").
int hc_layout(void) { return 10 * MK_FIELD + MK_PACKED; }

#define FE_BYTES (4 * MK_FE_LIMBS)

void hc_fe_mul_limbs(const u32* f, const u32* g, u32* out) {
    fe a, b;
    memcpy(a.v, f, FE_BYTES);
    memcpy(b.v, g, FE_BYTES);
    const fe r = fe_mul(a, b);
    memcpy(out, r.v, FE_BYTES);
}

void hc_fe_sq_limbs(const u32* f, u32* out) {
    fe a;
    memcpy(a.v, f, FE_BYTES);
    const fe r = fe_sq(a);
    memcpy(out, r.v, FE_BYTES);
}

// op on raw limbs (no reduction of the inputs): 0 add, 1 sub, 2 mul, 5 neg; out raw limbs.
void hc_fe_op_limbs(u32 op, const u32* f, const u32* g, u32* out) {
    fe a, b;
    memcpy(a.v, f, FE_BYTES);
    memcpy(b.v, g, FE_BYTES);
    fe r = op == 0 ? fe_add(a, b) : op == 1 ? fe_sub(a, b) : op == 2 ? fe_mul(a, b) : fe_neg(a);
    memcpy(out, r.v, FE_BYTES);
}

void hc_fe_to_words(const u32* f, u32 out[8]) {
    fe a;
    memcpy(a.v, f, FE_BYTES);
    fe_to_words(a, out);
}

void hc_fe_carried_to_words(const u32* f, u32 out[8]) {
    fe a;
    memcpy(a.v, f, FE_BYTES);
    fe_carried_to_words(a, out);
}

void hc_fe_from_words(const u32 w[8], u32* out) {
    const fe a = fe_from_words(w);
    memcpy(out, a.v, FE_BYTES);
}

// op: 0 add, 1 sub, 2 mul, 3 sq, 4 invert, 5 neg; a, b 256-bit words; out canonical.
void hc_fe_op(u32 op, const u32 a[8], const u32 b[8], u32 out[8]) {
    const fe x = fe_from_words(a), y = fe_from_words(b);
    fe r;
    switch (op) {
        case 0: r = fe_add(x, y); break;
        case 1: r = fe_sub(x, y); break;
        case 2: r = fe_mul(x, y); break;
        case 3: r = fe_sq(x); break;
        case 4: r = fe_invert(x); break;
        default: r = fe_neg(x); break;
    }
    fe_to_words(r, out);
}

// The claims of one launch in the order the counters hand them out: for each, the first key
// (of the launch's 64 * K * total_tiles) and the keys per lane; a claim covers 64 * Keff
// keys. serve_unit: keys per lane of a serving step's unit, 0 for a search launch. Returns
// the number of claims, or -1 when they exceed cap.
long hc_claims(u32 K, u32 serve_unit, u64 total_tiles, u64 waves, u64* starts, u32* keffs, u64 cap) {
    const mk_claim_plan p = mk_plan_claims(K, serve_unit, total_tiles, waves);
    u64 n = 0, q = 0;
    for (u64 t = 0;; t++) {
        const u64 claim = mk_claim_number(p, t, t < p.big_tiles ? 0 : q++);
        u64 tile, sub;
        u32 keff;
        if (!mk_claim_at(p, claim, tile, sub, keff)) break;
        if (n == cap) return -1;
        starts[n] = tile * 64 * K + sub * 64 * keff;
        keffs[n] = keff;
        n++;
    }
    return (long)n;
}

int hc_digits(const u32 s[8], u32 w, i32* out) {
    mk_digits g = mk_digits_of(s, w);
    for (u32 i = 0; i < g.rows; i++) out[i] = mk_next_digit(g);
    return (int)g.rows;
}

u64 hc_table_words(u32 w) { return (u64)mk_rows(w) * mk_cols(w) * MK_ENTRY_WORDS; }

// The whole comb table of width w, built on the CPU by the kernel's entry code.
void hc_table_build(u32 w, u32* table) {
    std::vector<mk_ext> bases(mk_rows(w) + 1);
    mk_row_bases(w, bases.data());
    const u32 rows = mk_rows(w), cols = mk_cols(w);
    for (u32 r = 0; r < rows; r++)
        for (u32 k = 0; k < cols; k++) mk_table_entry(w, bases.data(), r, k, table + ((u64)r * cols + k) * MK_ENTRY_WORDS);
}

void hc_table_entry(u32 w, u32 row, u32 k, u32 out[32]) {
    std::vector<mk_ext> bases(mk_rows(w) + 1);
    mk_row_bases(w, bases.data());
    mk_table_entry(w, bases.data(), row, k, out);
}

void hc_full_pk(const u32* table, u32 w, const u32 s[8], u32 pk[8]) { mk_full_pk(table, w, s, pk); }

// One search lane: keys first .. first+K-1, y words (x sign bit 0) and candidate flags,
// as the search kernel computes them.
int hc_lane(const u32* table, u32 w, const u32 key[8], u32 source, u64 first, u32 K, const u8* needle, u32 len,
            u32 fold, u32 mode, u32* ys, u8* cand) {
    matcher_host m;
    if (build_matcher(needle, len, fold, mode, m) != 0) return -1;
    mk_matcher mt;
    mt.mode = m.mode;
    mt.len = m.len;
    mt.mask = m.mask;
    mt.dfa = m.dfa.empty() ? nullptr : m.dfa.data();
    std::vector<u32> spill((size_t)K * 2 * MK_FE_LIMBS);
    mk_spill sp;
    sp.base = spill.data();
    sp.stride = 1;
    auto mul = [&](const u32 s[8], fe& yn, fe& yd) {
        fe xn, xd;
        mk_mul_base_frac(table, w, s, yn, yd, xn, xd);
    };
    mk_sha_mid mid;
    mk_sha_mid_init(key, mid);
    mk_lane_tile(
        K, key, source, first, sp, mul, [] {},
        [&](u32 j, const u32 pk[8]) {
            memcpy(ys + 8 * j, pk, 32);
            cand[j] = mk_pk_match(pk, mt, true) ? 1 : 0;
        },
        [&]() -> const mk_sha_mid& { return mid; });
    return 0;
}

int hc_build_mask(const u8* needle, u32 len, u32 fold, u32 mode, u64 mask[64]) {
    matcher_host m;
    if (build_matcher(needle, len, fold, mode, m) != 0) return -1;
    memcpy(mask, m.mask, sizeof m.mask);
    return 0;
}

// Match a full key (sign known unless sign_unknown) against a needle or DFA table.
int hc_match(const u32 pk[8], const u8* needle, u32 len, u32 fold, u32 mode, u32 sign_unknown) {
    matcher_host m;
    if (build_matcher(needle, len, fold, mode, m) != 0) return -1;
    mk_matcher mt;
    mt.mode = m.mode;
    mt.len = m.len;
    mt.mask = m.mask;
    mt.dfa = m.dfa.empty() ? nullptr : m.dfa.data();
    return mk_pk_match(pk, mt, sign_unknown != 0) ? 1 : 0;
}

// The host confirmation of an emitted key (1 = confirmed, 0 = refused, -1 = bad pattern).
int hc_confirm(const u32 key[8], u32 source, u64 index, const u8 seed[32], const u8 pk[32], const u8* needle,
               u32 len, u32 fold, u32 mode) {
    matcher_host m;
    if (build_matcher(needle, len, fold, mode, m) != 0) return -1;
    mk_matcher mt;
    mt.mode = m.mode;
    mt.len = m.len;
    mt.mask = m.mask;
    mt.dfa = m.dfa.empty() ? nullptr : m.dfa.data();
    return mk_confirm_key(key, source, index, seed, pk, mt) ? 1 : 0;
}

void hc_pk_by_ladder(const u32 s[8], u32 pk[8]) { mk_pk_by_ladder(s, pk); }

// The library's hit order over n (search, index) pairs: order[i] = position of the i-th.
void hc_sort_hits(const u32* search, const u64* index, u32 n, u32* order) {
    struct hit {
        u64 index;
        u32 search;
    };
    std::vector<hit> hs(n);
    for (u32 i = 0; i < n; i++) hs[i] = {index[i], search[i]};
    static mk_sort_pool pool;  // reused across calls, as the library reuses its own
    const u32* o = mk_sort_hits(hs.data(), n, pool);
    for (u32 i = 0; i < n; i++) order[i] = o[i];
}

// search_with over a scripted engine and clock (Python callbacks), as the reference tests
// its control loop with a mock Engine.
typedef int (*hc_launch_fn)(void* user, u64 base, u32 iters, int* found, u64* index);
typedef double (*hc_clock_fn)(void* user);
typedef void (*hc_report_fn)(void* user, u64 tries, double elapsed);
typedef int (*hc_cancel_fn)(void* user);

struct hc_engine {
    hc_launch_fn fn;
    void* user;
    u64 kpi;
    u64 keys_per_iter() const { return kpi; }
    int launch(u64 base, u32 iters, int& found, u64& index) { return fn(user, base, iters, &found, &index); }
};
struct hc_clock {
    hc_clock_fn fn;
    void* user;
    double now() { return fn(user); }
};
struct hc_progress {
    hc_report_fn fn;
    void* user;
    void report(u64 t, double e) { fn(user, t, e); }
};
struct hc_cancel {
    hc_cancel_fn fn;
    void* user;
    bool cancelled() { return fn(user) != 0; }
};

int hc_search_with(hc_launch_fn launch, hc_clock_fn clock, hc_report_fn report, hc_cancel_fn cancel, void* user,
                   u64 keys_per_iter, int bench, double target, double time_limit, double interval, u64* out) {
    hc_engine e{launch, user, keys_per_iter};
    hc_clock c{clock, user};
    hc_progress p{report, user};
    hc_cancel k{cancel, user};
    mk_search_result r = mk_search_with(e, c, p, k, bench != 0, target, time_limit, interval);
    out[0] = (u64)r.status;
    out[1] = (u64)r.found;
    out[2] = r.index;
    out[3] = r.tries;
    out[4] = r.launches;
    return r.status;
}

}  // extern "C"
