/*
 * meikey_hip — C ABI of the MI355X (gfx950) Ed25519 vanity-key search backend.
 *
 * Drop-in boundary for meikey's GPU backend. Each entry point names the reference
 * interface it replaces (paths relative to the meikey sources):
 *   - the device lifecycle and status rows of `Gpu` (src/service/gpu.rs:1127-1238,
 *     1251-1326) and `Kernel::load` (gpu.rs:678-745);
 *   - the private `trait Engine { keys_per_iter, prepare, launch }` that `search_with`
 *     drives (gpu.rs:807-887, 917-968);
 *   - the self-check / derivation kernels `meikey_derive` and `meikey_pooled`
 *     (src/service/gpu/kernel.cu:2037-2069, gpu.rs:747-801).
 * Batched multi-search and serving have no counterpart in the reference; they reuse the
 * same kernels and are declared at the end.
 *
 * Conventions: plain C types only; the caller owns every host array and passes
 * pointer + length; the library owns all device memory. Every function returns an
 * mk_status (0 = success); mk_last_error() gives the text of the last failure on the
 * calling thread. Calls on one mk_gpu must come from one host thread at a time
 * (the reference drives each GPU from one worker thread, src/service/search.rs:148-153).
 */
#ifndef MEIKEY_HIP_H
#define MEIKEY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mapped from the reference's GpuError (gpu.rs:445-499). */
typedef enum {
    MK_OK = 0,
    MK_ERR_DRIVER_NOT_FOUND = 1, /* GpuError::DriverNotFound: no HIP runtime / driver */
    MK_ERR_NO_DEVICE = 2,        /* GpuError::NoDevice */
    MK_ERR_OUT_OF_RANGE = 3,     /* GpuError::OutOfRange: width / batch / blocks outside their ranges */
    MK_ERR_SELF_CHECK = 4,       /* GpuError::SelfCheck: device results differ from the known answers */
    MK_ERR_DRIVER = 5,           /* GpuError::Driver: a HIP call failed */
    MK_ERR_INVALID = 6,          /* bad argument (null pointer, bad pattern table, ...) */
    MK_ERR_OVERFLOW = 7          /* more hits than the caller's buffer or the device lists hold */
} mk_status;

/* Seed sources (src/model/search.rs:33-42). */
#define MK_SEED_CHACHA20 0 /* default ("secure"): ChaCha20 block index/2, half index&1 */
#define MK_SEED_COUNTER 1  /* --low-security: last 8 bytes of the key + index */

/* Match modes (kernel.cu:1774-1776, src/model/pattern.rs:57-66). */
#define MK_MODE_CONTAINS 0
#define MK_MODE_SUFFIX 1
#define MK_MODE_REGEX 2

/* Field layouts (--gpu-limbs; the reference's Limbs and its codes, gpu.rs:70-106):
 *   MK_LIMBS_R25  "10"  ten 25.5-bit limbs, table entries packed as 3 x 32 bytes (96 B)
 *   MK_LIMBS_R25U "10u" ten 25.5-bit limbs, table entries stored as limbs (128 B)
 *   MK_LIMBS_R32  "8"   eight 32-bit words (full 2^256 range, carries folded by 38), packed entries
 *   MK_LIMBS_K26  "10k"  ten 26-bit limbs, Karatsuba products (75 limb products), packed entries
 *   MK_LIMBS_K26U "10ku" ten 26-bit limbs, Karatsuba products, table entries stored as limbs (128 B)
 * The floating-point "5f" is not built: the path is integer only (MK_ERR_OUT_OF_RANGE). */
#define MK_LIMBS_R25 10
#define MK_LIMBS_R25U 11
#define MK_LIMBS_R32 8
#define MK_LIMBS_K26 26
#define MK_LIMBS_K26U 27

/* Kernel settings: --gpu-width / --gpu-batch / --gpu-limbs (src/controller/cli.rs:85-111,
 * gpu.rs:185-240 GpuOptions). */
typedef struct {
    uint32_t width;  /* comb table width W, 2..28 (reference 2..20); table = rows(W) x (2^(W-1)+1) x 128 B */
    uint32_t batch;  /* keys per thread per tile, 1..256 (one inversion per `batch` keys per lane) */
    uint32_t limbs;  /* MK_LIMBS_R25 / MK_LIMBS_R25U / MK_LIMBS_R32 / MK_LIMBS_K26 / MK_LIMBS_K26U */
    uint32_t blocks; /* resident waves per SIMD the grid is sized for, 1..8 (reference: blocks per SM) */
    uint32_t lockstep; /* 0: one wave per workgroup; 1: the waves of a CU form one workgroup and add each
                          comb row together, which keeps the CU's address translation on one or two rows
                          of the table (blocks 1..2); 2: the same with workgroups of one wave per SIMD */
    uint32_t pool;     /* 0: each lane inverts its own batch; 4 or 8: workgroups of that many waves share
                          one inversion (lane l of wave 0 inverts the product of lane l's batches of all
                          the waves; the reference's block-pooled inversion, kernel.cu:1529-1541).
                          8 needs blocks 2, 4 needs blocks 1..2; not with lockstep */
    uint32_t lds_table; /* 1: each CU copies the whole comb table into LDS and its 8 waves read it there
                           (tables up to 144 KB: W <= 5, or W = 6 with limbs 10, 8 or 26); blocks 2, no
                           lockstep or pool. 0: entries gathered from HBM / L2 */
} mk_options;

/* Device status (gpu.rs:1251-1326 status rows). */
typedef struct {
    char name[64];          /* gcnArchName */
    uint32_t compute_units;
    uint32_t grid_waves;    /* workgroups (one wave64 each) per launch */
    uint64_t hbm_bytes;
    uint64_t table_bytes;
    uint64_t scratch_bytes; /* per-lane buffer of the batched inversion (80 B per key) */
    mk_options options;
    uint32_t rows, cols;    /* comb geometry of the table */
} mk_info;

/* One exact hit: the key at `index` matches with its real x sign. */
typedef struct {
    uint64_t index;
    uint32_t search; /* search number within a batch (0 for single searches) */
    uint32_t pad;
    uint8_t seed[32];
    uint8_t pk[32];
} mk_hit_rec;

typedef struct mk_gpu mk_gpu;
typedef struct mk_launch mk_launch;

/* Text of the last failure on this thread (copied into buf, NUL-terminated). */
void mk_last_error(char* buf, size_t len);

/* Number of HIP devices (device_count, gpu.rs:1127-1140). */
mk_status mk_device_count(int* n);

/* Default options (W = 24, 256 keys per thread, R25 limbs, 2 waves per SIMD): the fastest
 * setting measured on MI355X; the W = 24 table takes 11.8 GB of HBM. */
void mk_default_options(mk_options* out);

/* Open device `ordinal`: build the comb table of `opt->width` on the device and run the
 * known-answer self-check (Gpu::new + Kernel::load + self_check, gpu.rs:1162, 678-745). */
mk_status mk_open(int ordinal, const mk_options* opt, mk_gpu** out);
mk_status mk_info_get(const mk_gpu* g, mk_info* out);
mk_status mk_close(mk_gpu* g);

/* Engine::keys_per_iter (gpu.rs:816, 832-834): keys one launch covers per `iters` step. */
uint64_t mk_keys_per_iter(const mk_gpu* g);

/* Engine::prepare (gpu.rs:818-825, 836-855). `key` is the search's 32-byte ChaCha20 key
 * (or counter base). For MK_MODE_CONTAINS / MK_MODE_SUFFIX, `needle` is the text
 * (lower-cased by the caller when fold = 1, as Needle::new does, gpu.rs:627-659) and
 * needle_len its length (1..43; 0 = match nothing, for benchmarks). For MK_MODE_REGEX,
 * `needle` is the DFA table (states x 65 x u16 LE, Dfa::gpu_table, dfa.rs:187-200)
 * and needle_len the number of states. */
mk_status mk_prepare(mk_gpu* g, const uint8_t key[32], int seed_source, const uint8_t* needle, uint32_t needle_len,
                     uint32_t fold, uint32_t mode, uint32_t stop_on_hit, mk_launch** out);

/* Engine::launch (gpu.rs:826, 857-887): search indices [base, base + keys_per_iter * iters)
 * and return every exact hit, ascending by index (the reference returns one maybe-hit;
 * this returns the confirmed set, x sign included). With stop_on_hit, the launch may
 * stop early after the first candidate, and it returns the range's lowest exact hit
 * first but not necessarily every other one (a lane reports only its lowest hit whose
 * match holds for either x sign). *n_hits is the total number of exact hits returned;
 * MK_ERR_OVERFLOW when it exceeds `cap` or the device lists. */
mk_status mk_launch_run(mk_launch* l, uint64_t base, uint32_t iters, mk_hit_rec* hits, uint32_t cap,
                        uint32_t* n_hits);
mk_status mk_free_launch(mk_launch* l);

/* meikey_derive (kernel.cu:2062-2069): compressed public keys (with the x sign) of indices
 * first .. first+n-1; pks = n x 32 bytes, seeds (optional, may be NULL) = n x 32 bytes. */
mk_status mk_derive(mk_gpu* g, const uint8_t key[32], int seed_source, uint64_t first, uint64_t n, uint8_t* pks,
                    uint8_t* seeds);

/* Allocation-free mk_derive. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The keys are staged in a fixed device array of the library (2 x 2 MiB
 * per device, held until the process ends), so the call makes no hipMalloc / hipFree; it
 * takes at most MK_DERIVE_NOALLOC_MAX keys (MK_ERR_OUT_OF_RANGE beyond). All calls on one
 * GPU share that GPU's array: calls that overlap on one GPU (several threads or handles)
 * race, and the caller must serialize them. pks / seeds (optional) as in mk_derive. */
#define MK_DERIVE_NOALLOC_MAX 65536u
mk_status mk_derive_noalloc(mk_gpu* g, const uint8_t key[32], int seed_source, uint64_t first, uint32_t n,
                            uint8_t* pks, uint8_t* seeds);

/* Public keys of explicit seeds (n x 32 bytes in, n x 32 bytes out). */
mk_status mk_derive_seeds(mk_gpu* g, const uint8_t* seeds, uint32_t n, uint8_t* pks);

/* Raw comb-table entries (row, k) as stored on the device, for tests of the table build.
 * Each holds the reference's (y+x, y-x, 2dxy) halved: (y+x)/2, (y-x)/2, dxy (row 0: xy), every
 * field canonical. MK_LIMBS_R25U and MK_LIMBS_K26U: 32 words, three 10-limb fields (25.5-bit
 * or 26-bit limbs), limbs 0-7 of each in words 0-7 / 8-15 / 16-23, limbs 8-9 in words 24-25 /
 * 26-27 / 28-29, words 30-31 zero. Packed layouts (MK_LIMBS_R25, MK_LIMBS_R32, MK_LIMBS_K26):
 * 24 words, field f as a little-endian 256-bit value in words 8f .. 8f+7. */
mk_status mk_table_read(mk_gpu* g, uint32_t row, uint32_t k, uint32_t count, uint32_t* words);

/* ------------------------------------------------------------------------------------
 * Batched multi-search: many searches, one persistent launch (no reference counterpart).
 * Each search has its own key, seed source, pattern and index range; all share the
 * comb table and the kernel. Hits carry the search number (position in `specs`). */
typedef struct {
    uint8_t key[32];
    int32_t seed_source;
    uint32_t mode;
    uint32_t fold;
    uint32_t needle_len;
    const uint8_t* needle; /* text or DFA table, as in mk_prepare */
    uint32_t stop_on_hit; /* 1: the search's lowest exact hit is all it needs (as mk_launch_run) */
    uint32_t pad;
    uint64_t first; /* first index */
    uint64_t count; /* number of indices; rounded up to whole tiles of 64 * batch keys */
} mk_search_spec;

/* Run all `n` searches to completion in one launch. The exact hits come back ordered by
 * (search, index); *n_hits and MK_ERR_OVERFLOW as in mk_launch_run (no error when every
 * search stops on its first hit). tries (optional, n entries) receives the number of
 * indices each search covered. */
mk_status mk_batch_run(mk_gpu* g, const mk_search_spec* specs, uint32_t n, mk_hit_rec* hits, uint32_t cap,
                       uint32_t* n_hits, uint64_t* tries);

/* Keys one tile covers (64 lanes x batch). */
uint64_t mk_tile_keys(const mk_gpu* g);

/* Device time of the last launch's search kernel and confirm kernel, in ms, from HIP
 * events recorded on the library's stream around each kernel. */
mk_status mk_last_timing(const mk_gpu* g, float* search_ms, float* confirm_ms);

/* ------------------------------------------------------------------------------------
 * Serving: continuous batching of search requests (no reference counterpart). Requests
 * are admitted into device slots as slots and DFA-pool memory free up; each step is one
 * launch over a fixed number of tiles dealt round-robin across the requests served in that
 * step -- the resident ones with the least attained service -- whose next key index and stop
 * flag live on the device; the step is replayed from a captured hipGraph. A request ends at its first exact hit (its lowest index found in that step),
 * when `count` candidates are spent (count 0 = unbounded), or when cancelled. With `overlap`,
 * two steps are in flight on two streams, serving disjoint sets of requests: the next step
 * starts on the SIMDs the current one frees at its end, and the host retires one step while
 * the device runs the other (a request waits up to one more step before it starts). The server
 * uses the mk_gpu's stream and buffers: make no other call on that mk_gpu while it is open. */
typedef struct {
    uint32_t max_slots;       /* requests resident on the device at once, 1..4096 */
    uint32_t use_graph;       /* 1: replay the captured step graph; 0: enqueue each step */
    uint64_t dfa_budget;      /* HBM bytes for the admitted requests' DFA tables */
    uint64_t step_keys;       /* candidates per step (rounded up to whole tiles; at least one claim unit,
                                 unit_keys keys per lane or a whole smaller tile, per wave) */
    uint32_t active_per_step; /* resident requests served per step, the ones with the fewest
                                 candidates searched so far (least attained service); 0 = all;
                                 default 16 */
    uint32_t overlap;         /* 1: two steps in flight (disjoint requests); 0: one step at a time */
    uint32_t unit_keys;       /* keys per lane of a step's claim unit: 64, 128 (the default; 0 means
                                 128), 192 or 256; one that does not divide the batch leaves tiles whole.
                                 64 with half the step_keys halves a step and the median time to a
                                 hit, at about 1% of the served keys/s */
    uint32_t pad;
} mk_server_options;

typedef struct {
    uint64_t request;
    int32_t status; /* 1 found, 0 candidate budget spent, -1 invalid pattern or cancelled */
    uint32_t pad;
    uint64_t tries; /* candidates searched for this request */
    double submit_s, admit_s, done_s; /* seconds since mk_server_open */
    mk_hit_rec hit; /* when status == 1 */
} mk_server_result;

typedef struct mk_server mk_server;

void mk_server_default_options(mk_server_options* out);
mk_status mk_server_open(mk_gpu* g, const mk_server_options* opt, mk_server** out);
/* Queue a request (stop_on_hit is implied); *request receives its id. */
mk_status mk_server_submit(mk_server* s, const mk_search_spec* spec, uint64_t* request);
/* Admit what fits, run one step, retire finished requests. */
mk_status mk_server_step(mk_server* s, uint32_t* active, uint32_t* queued);
/* Take up to cap finished results. */
mk_status mk_server_poll(mk_server* s, mk_server_result* out, uint32_t cap, uint32_t* n);
mk_status mk_server_cancel(mk_server* s, uint64_t request);
mk_status mk_server_stats(const mk_server* s, uint64_t* steps, uint64_t* step_keys);
double mk_server_now(const mk_server* s);
mk_status mk_server_close(mk_server* s);

/* ------------------------------------------------------------------------------------
 * Patterns: Pattern::with_mode / probability / is_match and Dfa::new / gpu_table
 * (src/model/pattern.rs:115-262, src/model/pattern/dfa.rs:52-200), host-side. The
 * needle a pattern yields is what mk_prepare / mk_search_spec take (Needle::new,
 * gpu.rs:627-659): the lower-cased text when case-insensitive, or the DFA table. */

/* PatternError kinds (src/model/pattern.rs:23-55). */
typedef enum {
    MK_PATTERN_OK = 0,
    MK_PATTERN_EMPTY = 1,
    MK_PATTERN_INVALID_CHAR = 2,
    MK_PATTERN_TOO_LONG = 3,
    MK_PATTERN_REGEX = 4,
    MK_PATTERN_TOO_COMPLEX = 5,
    MK_PATTERN_MATCHES_EMPTY = 6,
    MK_PATTERN_NEVER_MATCHES = 7
} mk_pattern_error;

typedef struct mk_pattern mk_pattern;

/* Validate and compile `text` (text_len bytes) in mode MK_MODE_CONTAINS / SUFFIX / REGEX.
 * On failure returns MK_ERR_INVALID, sets *error_kind and the text of
 * mk_pattern_last_error(). */
mk_status mk_pattern_new(const char* text, uint32_t text_len, int case_sensitive, uint32_t mode, mk_pattern** out,
                         int* error_kind);
void mk_pattern_last_error(char* buf, size_t len);
mk_status mk_pattern_free(mk_pattern* p);
/* The needle for mk_prepare: bytes (owned by the pattern), length (characters, or DFA
 * states for a regex), fold and mode. mk_pattern_needle_bytes gives the byte size. */
mk_status mk_pattern_needle(const mk_pattern* p, const uint8_t** bytes, uint32_t* len, uint32_t* fold, uint32_t* mode);
size_t mk_pattern_needle_bytes(const mk_pattern* p);
/* Chance that one key matches (Pattern::probability); 1 / this is the expected tries. */
double mk_pattern_probability(const mk_pattern* p);
uint32_t mk_pattern_dfa_states(const mk_pattern* p);
/* Pattern::is_match on the 68-character public-key base64 (1 / 0, -1 on bad input). */
int mk_pattern_is_match(const mk_pattern* p, const char* b64);

/* ------------------------------------------------------------------------------------
 * Searcher::run for one GPU (src/service/search.rs:34-40, gpu.rs:999-1012) driving the
 * Engine through search_with (gpu.rs:917-968): a fresh random key, then launches over
 * consecutive index ranges whose iteration count adapts toward `target_launch_s` per
 * launch (x0.5..x4 per step, 1..2^16), until an exact hit (not in bench mode), the cancel
 * flag, or the time limit. */
typedef struct {
    uint64_t tries;
    double elapsed_s;
} mk_stats; /* Stats (src/model/search.rs:84-94) */

typedef void (*mk_progress_fn)(void* user, const mk_stats* stats);

typedef struct {
    const mk_pattern* pattern;      /* NULL: no pattern (matches nothing; benchmark) */
    int32_t seed_source;            /* MK_SEED_CHACHA20 / MK_SEED_COUNTER */
    uint32_t bench;                 /* 1: keep searching through hits (--bench) */
    double time_limit_s;            /* 0: none */
    double progress_interval_s;     /* progress callback period (0: 5 s, model/search.rs:76) */
    double target_launch_s;         /* per-launch target (0: 0.2 s, gpu.rs:427) */
    const volatile int32_t* cancel; /* optional; non-zero stops the search between launches */
    const uint8_t* key;             /* optional fixed 32-byte key (tests); NULL = fresh OS randomness */
} mk_search_config;

typedef struct {
    int32_t found; /* 1: index / seed / pk hold the hit */
    uint32_t pad;
    mk_stats stats;
    uint64_t launches;
    uint64_t index;
    uint8_t key[32]; /* the search key that was used */
    uint8_t seed[32];
    uint8_t pk[32];
} mk_outcome;

mk_status mk_search(mk_gpu* g, const mk_search_config* cfg, mk_progress_fn progress, void* user, mk_outcome* out);

/* Autotuner (gpu.rs:1014-1113 fastest / fastest_staged): each candidate is opened (table
 * build + self-check; a failing setting is reported with its status and skipped, like the
 * reference's "Rejected" list) and measured as gpu.rs:889-911 does (counter seeds, needle
 * "meikey" case-insensitive, three ~100 ms launches after a sizing launch). */
typedef struct {
    mk_options options;
    double keys_per_s;
    int32_t status; /* MK_OK, or the error that rejected the setting */
    uint32_t pad;
} mk_tune_result;

mk_status mk_fastest(int ordinal, const mk_options* candidates, uint32_t n, mk_tune_result* results, mk_options* best);
/* Stage 1: the layouts "10u", "10", "8", "10ku" and "10k" at W = 16 (gpu.rs:364-370); stage 2: widths
 * 16/20/22/24 gathered from HBM / L2, and W = 5 held in LDS; stage 3: the inversion,
 * 32/64/128/256 keys per lane each inverting its own batch or 8/32 keys per lane with one
 * inversion per 4-wave workgroup (pool 4); stage 4: 1 or 2 waves per SIMD, and lockstep
 * workgroups (for a per-lane inversion; 8-wave pools for a pooled one). results (cap
 * entries) receive every measured setting; *n the count. */
mk_status mk_fastest_staged(int ordinal, mk_options* best, mk_tune_result* results, uint32_t cap, uint32_t* n);

/* OpenSSH text of a key: the authorized_keys line "ssh-ed25519 <b64>[ <comment>]"
 * (key.rs:132-139) and the unencrypted openssh-key-v1 private key (key_file.rs:276-320;
 * checkint is the random check word). Comments may not contain line breaks (key.rs:64-75).
 * *len receives the text length; out (cap bytes, may be NULL to query) gets the text. */
mk_status mk_openssh_public(const uint8_t pk[32], const char* comment, char* out, size_t cap, size_t* len);
mk_status mk_openssh_private(const uint8_t seed[32], const uint8_t pk[32], const char* comment, uint32_t checkint,
                             char* out, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif
#endif
