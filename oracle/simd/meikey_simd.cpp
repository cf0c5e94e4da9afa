// The SIMD CPU baseline: a C++ restatement of the reference's CPU search path, to time the
// CPU against the GPU on the same workload. Test and measurement infrastructure only:
// bench.py's cpu_baseline leg and tests/ load it (libmeikey_simd.so, oracle/Makefile), and
// nothing in the product links it.
//
// What it restates (one engine per instruction set, mks_engine.h):
// - SHA-512 of 8 seeds at a time (src/service/derive/sha512x8.rs:1-11);
// - the fixed-base comb of width 12 (src/service/derive/comb.rs:17) over a table of
//   affine Niels points (y+x, y-x, 2dxy), 8 keys in the 8 lanes of a vector, in the AVX-512
//   IFMA field of five 51-bit limbs (src/service/derive/avx512/ifma.rs:1-10), or, on a CPU
//   without IFMA, ten 25.5-bit limbs with AVX2 32x32-bit products (chosen at run time, as
//   Lanes::detect does, src/service/derive.rs:196-230);
// - one Montgomery inversion per 256 keys (BATCH, src/service/derive.rs:51, 486-543);
// - unsigned-first matching (src/service/search.rs:273-288, src/service/matcher.rs:99-129).
// Seeds are the index-addressed ChaCha20 seeds of the GPU search (src/service/gpu.rs:
// 1344-1351) or the counter seeds (src/model/search.rs:45-50). The reference's CPU worker
// reads its ChaCha20 seeds from a per-thread ChaCha20Rng stream instead
// (src/service/search.rs:249-266): the same block function, not addressable by index.
//
// Threads: one per CPU of the process's affinity mask, each pinned to its CPU.
#include <immintrin.h>
#include <pthread.h>
#include <sched.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef uint8_t u8;
typedef uint32_t u32;
typedef uint64_t u64;

#define MKS_W 12
#define MKS_ROWS 21     // ceil(251 / 12)
#define MKS_COLS 2049   // 2^11 + 1
#define MKS_BATCH 256   // keys per inversion
#define MKS_GROUPS (MKS_BATCH / 8)
#define MKS_B64_LEN 68      // base64 of the 51-byte key blob
#define MKS_SEARCH_START 25 // the hay: characters 25..67 (src/model/pattern.rs:192-208)
#define MKS_SIGN_CHAR 66    // the character that holds the x sign (src/service/matcher.rs:137)

enum { MODE_CONTAINS = 0, MODE_SUFFIX = 1, MODE_REGEX = 2 };

static const char B64_ALPHABET[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

// ------------------------------------------------------------------ matching

struct matcher {
    const u8* needle;  // the pattern text, or the DFA table for MODE_REGEX
    u32 len;
    int cs, mode;
    u8 folded[64];  // the needle in lower case (case-insensitive modes)
};

static inline u8 lower(u8 c) { return (c >= 'A' && c <= 'Z') ? (u8)(c + 32) : c; }

static inline int b64_value(u8 c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    return c == '+' ? 62 : c == '/' ? 63 : 64;
}

// DFA step (src/model/pattern/dfa.rs:117-137): state 0 accepts and absorbs, 1 starts,
// column 64 is the end of the text; 65 little-endian u16 entries per state.
static inline u32 dfa_next(const u8* t, u32 s, u32 c) {
    const u32 at = (s * 65u + c) * 2u;
    return (u32)t[at] | (u32)t[at + 1] << 8;
}

// Pattern::find semantics on the hay text[25..68) (src/model/pattern.rs:192-220).
static bool hay_match(const char* text, const matcher& m) {
    if (m.mode == MODE_REGEX) {
        u32 s = 1;
        for (int i = MKS_SEARCH_START; i < MKS_B64_LEN; i++) {
            s = dfa_next(m.needle, s, (u32)b64_value((u8)text[i]) & 63);
            if (s == 0) return true;
        }
        return dfa_next(m.needle, s, 64) != 0;
    }
    const u8* nd = m.cs ? m.needle : m.folded;
    auto same = [&](const char* h) {
        for (u32 i = 0; i < m.len; i++)
            if ((m.cs ? (u8)h[i] : lower((u8)h[i])) != nd[i]) return false;
        return true;
    };
    if (m.mode == MODE_SUFFIX) return same(text + MKS_B64_LEN - m.len);
    for (u32 st = MKS_SEARCH_START; st + m.len <= MKS_B64_LEN; st++)
        if (same(text + st)) return true;
    return false;
}

// Characters 24..67 of the base64 of the blob string("ssh-ed25519") || string(pk)
// (src/model/key.rs:9-52): blob bytes 18..50, i.e. 32 and the key. Returns the 6-bit value
// of the sign character.
static inline int encode_tail(const u8 pk[32], char text[MKS_B64_LEN]) {
    u8 b[33];
    b[0] = 32;
    memcpy(b + 1, pk, 32);
    u32 n = 0;
    for (int g = 0; g < 11; g++) {
        n = (u32)b[3 * g] << 16 | (u32)b[3 * g + 1] << 8 | b[3 * g + 2];
        char* o = text + 24 + 4 * g;
        o[0] = B64_ALPHABET[n >> 18];
        o[1] = B64_ALPHABET[(n >> 12) & 63];
        o[2] = B64_ALPHABET[(n >> 6) & 63];
        o[3] = B64_ALPHABET[n & 63];
    }
    return (int)((n >> 6) & 63);
}

// ------------------------------------------------------------------ engines

#pragma GCC push_options
#pragma GCC target("avx512f,avx512ifma,avx512vl,avx512bw,avx512dq")
#define MKS_ISA 512
#include "mks_engine.h"
#undef MKS_ISA
#pragma GCC pop_options

#pragma GCC push_options
#pragma GCC target("avx2")
#define MKS_ISA 2
#include "mks_engine.h"
#undef MKS_ISA
#pragma GCC pop_options

// ------------------------------------------------------------------ dispatch

static int g_isa = -1;

// 512: AVX-512F with IFMA; 2: AVX2; 0: neither (nothing runs). MEIKEY_SIMD_ISA=2 selects
// the AVX2 engine on a CPU that has both.
extern "C" int mks_isa(void) {
    if (g_isa < 0) {
        __builtin_cpu_init();
        const bool ifma = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512ifma") &&
                          __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512bw") &&
                          __builtin_cpu_supports("avx512dq");
        const bool avx2 = __builtin_cpu_supports("avx2");
        g_isa = ifma ? 512 : avx2 ? 2 : 0;
        const char* force = getenv("MEIKEY_SIMD_ISA");
        if (force && atoi(force) == 2 && avx2) g_isa = 2;
    }
    return g_isa;
}

static bool g_loaded = false;

// Use engine `isa` (512 or 2) from now on, if the CPU has it; the table must be loaded
// again. 0, or -1 (not supported here).
extern "C" int mks_select(int isa) {
    mks_isa();
    __builtin_cpu_init();
    const bool ok = isa == 2 ? __builtin_cpu_supports("avx2")
                  : isa == 512 ? __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512ifma") &&
                                     __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("avx512bw") &&
                                     __builtin_cpu_supports("avx512dq")
                               : false;
    if (!ok) return -1;
    g_isa = isa;
    g_loaded = false;
    return 0;
}

// The W = 12 comb table: 21 rows x 2049 columns x 96 bytes of canonical little-endian
// (y + x, y - x, 2dxy), entry (i, k) = k 2^(12 i) 8B on rows 0..19 and (k + 2^11) 2^240 8B
// on row 20 (src/service/derive/comb.rs:289-339).
extern "C" int mks_load_table(const u8* niels) {
    const int isa = mks_isa();
    const int rc = isa == 512 ? mks512::load_table(niels) : isa == 2 ? mks2::load_table(niels) : -1;
    g_loaded = rc == 0;
    return rc;
}

static void* new_state() {
    return g_isa == 512 ? (void*)mks512::new_state() : (void*)mks2::new_state();
}

static void derive_batch(void* st, const u8 key[32], int source, u64 first, u32 n, u8* pks) {
    if (g_isa == 512)
        mks512::derive((mks512::state*)st, key, source, first, n, pks);
    else
        mks2::derive((mks2::state*)st, key, source, first, n, pks);
}

static u32 search_batch(void* st, const u8 key[32], int source, u64 first, u32 n, const matcher& m, u64* hits,
                        u32 cap) {
    return g_isa == 512 ? mks512::search((mks512::state*)st, key, source, first, n, m, hits, cap)
                        : mks2::search((mks2::state*)st, key, source, first, n, m, hits, cap);
}

static bool make_matcher(matcher& m, const u8* needle, u32 len, int cs, int mode) {
    m.needle = needle;
    m.len = len;
    m.cs = cs;
    m.mode = mode;
    if (mode == MODE_REGEX) return needle != nullptr;
    if (len == 0 || len > MKS_B64_LEN - MKS_SEARCH_START || mode < 0 || mode > 1) return false;
    for (u32 i = 0; i < len; i++) m.folded[i] = lower(needle[i]);
    return true;
}

// Public keys (x sign included) of seeds first .. first + n - 1 of `source` (0 ChaCha20,
// 1 counter), 32 bytes each, on the calling thread. 0, or -1 (no table or no SIMD).
extern "C" int mks_derive(const u8 key[32], int source, u64 first, u64 n, u8* pks) {
    if (!g_loaded) return -1;
    void* st = new_state();
    if (!st) return -1;
    for (u64 off = 0; off < n; off += MKS_BATCH) {
        const u32 m = (u32)(n - off < MKS_BATCH ? n - off : MKS_BATCH);
        derive_batch(st, key, source, first + off, m, pks + 32 * off);
    }
    free(st);
    return 0;
}

// ------------------------------------------------------------------ threads

struct worker {
    const u8* key;
    int source;
    const matcher* m;
    u64 first, n;    // range (search) or start index (timed)
    double seconds;  // > 0: timed run
    int cpu;
    u64 done, count;
    u64* hits;
    u64 cap;
    int failed;
};

static double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void* run_worker(void* arg) {
    worker* w = (worker*)arg;
    if (w->cpu >= 0) {
        cpu_set_t set;
        CPU_ZERO(&set);
        CPU_SET(w->cpu, &set);
        sched_setaffinity(0, sizeof set, &set);
    }
    void* st = new_state();
    if (!st) {
        w->failed = 1;
        return nullptr;
    }
    u64 local[MKS_BATCH];
    if (w->seconds > 0) {
        const double end = now_s() + w->seconds;
        u64 idx = w->first;
        while (now_s() < end) {
            for (int i = 0; i < 8; i++, idx += MKS_BATCH)
                w->count += search_batch(st, w->key, w->source, idx, MKS_BATCH, *w->m, local, MKS_BATCH);
            w->done += 8 * MKS_BATCH;
        }
    } else {
        for (u64 off = 0; off < w->n; off += MKS_BATCH) {
            const u32 n = (u32)(w->n - off < MKS_BATCH ? w->n - off : MKS_BATCH);
            const u32 c = search_batch(st, w->key, w->source, w->first + off, n, *w->m, local, MKS_BATCH);
            for (u32 i = 0; i < c; i++, w->count++)
                if (w->count < w->cap) w->hits[w->count] = local[i];
            w->done += n;
        }
    }
    free(st);
    return nullptr;
}

// The CPUs of this process's affinity mask, in order.
static int cpu_list(int* out, int max) {
    cpu_set_t set;
    int n = 0;
    if (sched_getaffinity(0, sizeof set, &set) == 0)
        for (int c = 0; c < CPU_SETSIZE && n < max; c++)
            if (CPU_ISSET(c, &set)) out[n++] = c;
    return n;
}

#define MKS_MAX_THREADS 1024

extern "C" int mks_cpu_count(void) {
    int cpus[MKS_MAX_THREADS];
    return cpu_list(cpus, MKS_MAX_THREADS);
}

static int run_workers(worker* ws, int threads) {
    int cpus[MKS_MAX_THREADS];
    const int nc = cpu_list(cpus, MKS_MAX_THREADS);
    pthread_t* th = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    if (!th) return -1;
    int started = 0, rc = 0;
    for (int i = 0; i < threads; i++) {
        ws[i].cpu = nc > 0 ? cpus[i % nc] : -1;
        if (pthread_create(&th[i], nullptr, run_worker, &ws[i]) != 0) {
            rc = -1;
            break;
        }
        started++;
    }
    for (int i = 0; i < started; i++) pthread_join(th[i], nullptr);
    for (int i = 0; i < started; i++)
        if (ws[i].failed) rc = -1;
    free(th);
    return rc;
}

// Exact hits among seeds first .. first + n - 1, in index order, on `threads` pinned
// threads (each a contiguous share of the range). Returns the number of hits (the first
// `cap` stored), or -1.
extern "C" long long mks_search(const u8 key[32], int source, u64 first, u64 n, const u8* needle, u32 len, int cs,
                                int mode, u64* hits, u64 cap, int threads) {
    matcher m;
    if (!g_loaded || threads < 1 || threads > MKS_MAX_THREADS || !make_matcher(m, needle, len, cs, mode)) return -1;
    const u64 batches = (n + MKS_BATCH - 1) / MKS_BATCH, per = (batches + threads - 1) / threads;
    worker* ws = (worker*)calloc((size_t)threads, sizeof(worker));
    if (!ws) return -1;
    for (int i = 0; i < threads; i++) {
        const u64 a = per * MKS_BATCH * (u64)i, b = per * MKS_BATCH * (u64)(i + 1);
        const u64 lo = a < n ? a : n, hi = b < n ? b : n;
        ws[i] = worker{key, source, &m, first + lo, hi - lo, 0.0, -1, 0, 0, nullptr, cap, 0};
        ws[i].hits = (u64*)malloc(sizeof(u64) * (cap ? cap : 1));
    }
    long long total = run_workers(ws, threads);
    if (total == 0) {
        for (int i = 0; i < threads; i++) {
            for (u64 k = 0; k < ws[i].count && k < ws[i].cap; k++)
                if ((u64)total + k < cap) hits[total + k] = ws[i].hits[k];
            total += (long long)ws[i].count;
        }
    }
    for (int i = 0; i < threads; i++) free(ws[i].hits);
    free(ws);
    return total;
}

// Timed search: `threads` pinned threads (thread i from seed index base + i 2^40, batch
// after batch) for `seconds`. *keys = candidates searched, *elapsed = wall seconds,
// *n_hits = exact hits. 0, or -1.
extern "C" int mks_bench(const u8 key[32], int source, u64 base, const u8* needle, u32 len, int cs, int mode,
                         double seconds, int threads, u64* keys, double* elapsed, u64* n_hits) {
    matcher m;
    if (!g_loaded || threads < 1 || threads > MKS_MAX_THREADS || !make_matcher(m, needle, len, cs, mode)) return -1;
    worker* ws = (worker*)calloc((size_t)threads, sizeof(worker));
    if (!ws) return -1;
    for (int i = 0; i < threads; i++)
        ws[i] = worker{key, source, &m, base + ((u64)i << 40), 0, seconds, -1, 0, 0, nullptr, 0, 0};
    const double t0 = now_s();
    const int rc = run_workers(ws, threads);
    *elapsed = now_s() - t0;
    u64 k = 0, h = 0;
    for (int i = 0; i < threads; i++) {
        k += ws[i].done;
        h += ws[i].count;
    }
    free(ws);
    *keys = k;
    *n_hits = h;
    return rc;
}
