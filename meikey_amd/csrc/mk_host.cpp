// C ABI of the gfx950 search backend (include/meikey_hip.h): device lifecycle, the
// comb-table build, the single-search Engine (prepare / launch), derivation, and the
// batched multi-search launch. Device work is launched through mk_launch.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "mk_hitsort.h"
#include "mk_internal.h"

namespace {

thread_local std::string g_last_error;

}  // namespace

mk_status mk_fail(mk_status st, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

namespace {

#define fail mk_fail

#define MK_HIP(call)                                                                                 \
    do {                                                                                             \
        hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess) return fail(MK_ERR_DRIVER, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                       \
    } while (0)

}  // namespace

#define MK_MATCHER_FAIL(st, ...) (fail((mk_status)(st), __VA_ARGS__), (st))
#include "mk_matcher_host.h"

namespace {

constexpr u32 kMaxBatch = 256;

u32 rows_of(u32 w) { return (MK_M_BITS + w - 1) / w; }
u32 cols_of(u32 w) { return (1u << (w - 1)) + 1; }

}  // namespace

void mk_key_to_words(const uint8_t key[32], u32 w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (u32)key[4 * i] | (u32)key[4 * i + 1] << 8 | (u32)key[4 * i + 2] << 16 | (u32)key[4 * i + 3] << 24;
}

void mk_words_to_bytes(const u32 w[8], uint8_t out[32]) {
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 4; k++) out[4 * i + k] = (u8)(w[i] >> (8 * k));
}

struct mk_launch {
    mk_gpu* g = nullptr;
    u32 stop_on_hit = 0;
    mk_search_desc* descs = nullptr;
    u64* masks = nullptr;
    u16* dfas = nullptr;
    mk_seg* segs = nullptr;
    u32* stop = nullptr;
};

// One launch: reset, search, confirm. On return `hits` points at the min(n_hits, kHitCap)
// exact hits in the device's pinned host buffer (valid until the next launch on g).
mk_status mk_run_step(mk_gpu* g, const mk_search_desc* descs, const u64* masks, const u16* dfas, const mk_seg* segs,
                      u32 nsegs, u64 total_tiles, u32* stop, u32 nstop, const mk_hit*& hits, u32& n_cands,
                      u32& n_hits, u32& overflow) {
    MK_HIP(mk_k_reset(g->stream, g->ctr, stop, nstop, nullptr));
    MK_HIP(hipEventRecord(g->ev[0], g->stream));
    if (total_tiles > 0)
        MK_HIP(g->ks->search(g->stream, g->grid, g->lock_waves, g->table, g->opt.width, g->opt.batch, descs, segs, nsegs, total_tiles,
                           masks, dfas, g->spill, g->ctr, g->cands, kCandCap, stop, nullptr, nullptr));
    MK_HIP(hipEventRecord(g->ev[1], g->stream));
    if (total_tiles > 0)
        MK_HIP(g->ks->confirm(g->stream, g->table, g->opt.width, descs, masks, dfas, g->ctr, g->cands, kCandCap, g->hits,
                            kHitCap));
    MK_HIP(hipEventRecord(g->ev[2], g->stream));
    MK_HIP(hipMemcpyAsync(g->h_ctr, g->ctr, sizeof(mk_counters), hipMemcpyDeviceToHost, g->stream));
    MK_HIP(hipStreamSynchronize(g->stream));
    MK_HIP(hipEventElapsedTime(&g->last_search_ms, g->ev[0], g->ev[1]));
    MK_HIP(hipEventElapsedTime(&g->last_confirm_ms, g->ev[1], g->ev[2]));
    n_cands = g->h_ctr->n_cands;
    n_hits = g->h_ctr->n_hits;
    overflow = g->h_ctr->overflow;
    const u32 n = std::min(n_hits, kHitCap);
    if (n) {
        MK_HIP(hipMemcpyAsync(g->h_hits, g->hits, sizeof(mk_hit) * n, hipMemcpyDeviceToHost, g->stream));
        MK_HIP(hipStreamSynchronize(g->stream));
    }
    hits = g->h_hits;
    return MK_OK;
}

namespace {

// The n exact hits as caller records, ordered by (search, index) (mk_hitsort.h). mk_hit
// has the record's layout, so the gather is one copy per hit.
static_assert(sizeof(mk_hit) == sizeof(mk_hit_rec) && offsetof(mk_hit, index) == offsetof(mk_hit_rec, index) &&
                  offsetof(mk_hit, search) == offsetof(mk_hit_rec, search) &&
                  offsetof(mk_hit, seed) == offsetof(mk_hit_rec, seed) && offsetof(mk_hit, pk) == offsetof(mk_hit_rec, pk),
              "mk_hit must have the layout of mk_hit_rec");
static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "hit words are copied as little-endian bytes");

void to_records(mk_gpu* g, const mk_hit* hs, u32 n, mk_hit_rec* recs, u32 cap) {
    const u32* order = mk_sort_hits(hs, n, g->sort);
    const u32 m = std::min(n, cap);
    const u32 T = mk_host_threads(m);
    mk_host_parallel(T, [&](u32 t) {
        const u32 b = (u32)((u64)m * t / T), e = (u32)((u64)m * (t + 1) / T);
        for (u32 i = b; i < e; i++) {
            if (i + 8 < e) __builtin_prefetch(&hs[order[i + 8]]);
            memcpy(&recs[i], &hs[order[i]], sizeof(mk_hit_rec));
            recs[i].pad = 0;
        }
    });
}

// RFC 8032 s7.1 known answers (secret seed -> public key): TEST 1, 2, 3, 1024, SHA(abc).
const char* const kKat[5][2] = {
    {"9d61b19deffd5a60ba844af492ec2cc44449c5697b326919703bac031cae7f60",
     "d75a980182b10ab7d54bfed3c964073a0ee172f3daa62325af021a68f707511a"},
    {"4ccd089b28ff96da9db6c346ec114e0f5b8a319f35aba624da8cf6ed4fb8a6fb",
     "3d4017c3e843895a92b70aa74d1b7ebc9c982ccf2ec4968cc0cd55f12af4660c"},
    {"c5aa8df43f9f837bedb7442f31dcb7b166d38535076f094b85ce3a2e0b4458f7",
     "fc51cd8e6218a1a38da47ed00230f0580816ed13ba3303ac5deb911548908025"},
    {"f5e5767cf153319517630f226876b86c8160cc583bc013744c6bf255f5cc0ee5",
     "278117fc144c72340f67d0f2316e8386ceffbf2b2428c9c51fef7c597f1d426e"},
    {"833fe62409237b9d62ec77587520911e9a759cec1d19755b7da901b96dca3d42",
     "ec172b93ad5e563bf4932c70e1245034c35467ef2efd4d64ebf819683467e2bf"},
};

void unhex(const char* h, u8* out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        unsigned v;
        sscanf(h + 2 * i, "%2x", &v);
        out[i] = (u8)v;
    }
}

// 68-char base64 of the ssh-ed25519 blob of pk (src/model/key.rs:43-52).
void public_b64(const u8 pk[32], char out[68]) {
    u8 blob[51] = {0, 0, 0, 11, 's', 's', 'h', '-', 'e', 'd', '2', '5', '5', '1', '9', 0, 0, 0, 32};
    memcpy(blob + 19, pk, 32);
    for (int g = 0; g < 17; g++) {
        const u32 n = (u32)blob[3 * g] << 16 | (u32)blob[3 * g + 1] << 8 | blob[3 * g + 2];
        for (int k = 0; k < 4; k++) out[4 * g + k] = mk_b64_alphabet[(n >> (18 - 6 * k)) & 63];
    }
}

// Self-check (Kernel::self_check, gpu.rs:713-745): the derivation kernel must reproduce the
// RFC 8032 public keys, and the search path must find a key whose text it was given.
mk_status self_check(mk_gpu* g) {
    u8 seeds[5][32], want[5][32], got[5][32];
    for (int i = 0; i < 5; i++) {
        unhex(kKat[i][0], seeds[i], 32);
        unhex(kKat[i][1], want[i], 32);
    }
    mk_status st = mk_derive_seeds(g, &seeds[0][0], 5, &got[0][0]);
    if (st != MK_OK) return st;
    for (int i = 0; i < 5; i++)
        if (memcmp(got[i], want[i], 32) != 0)
            return fail(MK_ERR_SELF_CHECK, "self-check: RFC 8032 vector %d differs (W=%u, batch %u)", i,
                        g->opt.width, g->opt.batch);
    // search path: take 8 characters of the key at index `target` as a case-sensitive
    // substring and search a range around it
    u8 key[32];
    for (int i = 0; i < 32; i++) key[i] = (u8)(0xa5 ^ (i * 29));
    const u64 tile = 64ull * g->opt.batch;
    const u64 target = 3 * tile + 37;
    u8 pk[32];
    st = mk_derive(g, key, MK_SEED_CHACHA20, target, 1, pk, nullptr);
    if (st != MK_OK) return st;
    char b64[68];
    public_b64(pk, b64);
    mk_search_spec spec{};
    memcpy(spec.key, key, 32);
    spec.seed_source = MK_SEED_CHACHA20;
    spec.mode = MK_MODE_CONTAINS;
    spec.needle_len = 8;
    spec.needle = (const uint8_t*)(b64 + 40);
    spec.first = 0;
    spec.count = 5 * tile;
    std::vector<mk_hit_rec> hits(64);
    u32 n = 0;
    st = mk_batch_run(g, &spec, 1, hits.data(), (u32)hits.size(), &n, nullptr);
    if (st != MK_OK) return st;
    bool found = false;
    for (u32 i = 0; i < std::min<u32>(n, (u32)hits.size()); i++)
        if (hits[i].index == target && memcmp(hits[i].pk, pk, 32) == 0) found = true;
    if (!found)
        return fail(MK_ERR_SELF_CHECK, "self-check: the search did not report index %llu (W=%u, batch %u)",
                    (unsigned long long)target, g->opt.width, g->opt.batch);
    return MK_OK;
}

}  // namespace

extern "C" {

void mk_last_error(char* buf, size_t len) {
    if (!buf || !len) return;
    snprintf(buf, len, "%s", g_last_error.c_str());
}

mk_status mk_device_count(int* n) {
    if (!n) return fail(MK_ERR_INVALID, "null pointer");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e == hipErrorNoDevice) {
        *n = 0;
        return MK_OK;
    }
    if (e != hipSuccess) return fail(MK_ERR_DRIVER_NOT_FOUND, "hipGetDeviceCount: %s", hipGetErrorString(e));
    return MK_OK;
}

void mk_default_options(mk_options* out) {
    if (!out) return;
    out->width = 24;
    out->batch = 256;
    out->limbs = MK_LIMBS_R25U;
    out->blocks = 2;
    out->lockstep = 0;
    out->pool = 0;
    out->lds_table = 0;
}

mk_status mk_open(int ordinal, const mk_options* opt_in, mk_gpu** out) {
    if (!out) return fail(MK_ERR_INVALID, "null output pointer");
    *out = nullptr;
    mk_options opt;
    mk_default_options(&opt);
    if (opt_in) opt = *opt_in;
    if (opt.width < MK_MIN_W || opt.width > MK_MAX_W || opt.batch < 1 || opt.batch > kMaxBatch || opt.blocks < 1 ||
        opt.blocks > 8)
        return fail(MK_ERR_OUT_OF_RANGE,
                    "the table width must be %d-%d, the keys per thread 1-%u and the waves per SIMD 1-8 (W=%u, "
                    "batch %u, blocks %u)",
                    MK_MIN_W, MK_MAX_W, kMaxBatch, opt.width, opt.batch, opt.blocks);
    const mk_kernel_set* ks = opt.limbs == MK_LIMBS_R25U   ? &mk_kernels_u10
                              : opt.limbs == MK_LIMBS_R25  ? &mk_kernels_p10
                              : opt.limbs == MK_LIMBS_R32  ? &mk_kernels_r32
                              : opt.limbs == MK_LIMBS_K26  ? &mk_kernels_k26
                              : opt.limbs == MK_LIMBS_K26U ? &mk_kernels_k26u
                                                           : nullptr;
    if (!ks)
        return fail(MK_ERR_OUT_OF_RANGE,
                    "limb layout %u is not available (10, 11 = \"10u\", 8, 26 = \"10k\" or 27 = \"10ku\"; the "
                    "floating-point \"5f\" is not built)",
                    opt.limbs);
    if (opt.lockstep > 2 || (opt.lockstep == 1 && opt.blocks > 2))
        return fail(MK_ERR_OUT_OF_RANGE,
                    "lockstep is 0, 1 (one workgroup per CU, 1 or 2 waves per SIMD) or 2 (workgroups of one wave per "
                    "SIMD) (lockstep %u, blocks %u)", opt.lockstep, opt.blocks);
    if ((opt.pool != 0 && opt.pool != 4 && opt.pool != 8) || (opt.pool == 8 && opt.blocks != 2) ||
        (opt.pool == 4 && opt.blocks > 2) || (opt.pool != 0 && opt.lockstep != 0))
        return fail(MK_ERR_OUT_OF_RANGE,
                    "pool is 0, 4 (workgroups of one wave per SIMD, 1 or 2 waves per SIMD) or 8 (one workgroup per "
                    "CU, 2 waves per SIMD), and not combined with lockstep (pool %u, blocks %u, lockstep %u)",
                    opt.pool, opt.blocks, opt.lockstep);
    if (opt.lds_table > 1 ||
        (opt.lds_table == 1 &&
         ((u64)rows_of(opt.width) * cols_of(opt.width) * ks->entry_words * 4 > MK_LDS_TABLE_BYTES || opt.blocks != 2 ||
          opt.lockstep != 0 || opt.pool != 0)))
        return fail(MK_ERR_OUT_OF_RANGE,
                    "lds_table 1 needs a table of at most %u bytes (W <= 5, or W = 6 with packed entries: limbs 10, 8 or 26), 2 waves per "
                    "SIMD and no lockstep or pool (lds_table %u, W=%u, limbs %u, blocks %u)",
                    MK_LDS_TABLE_BYTES, opt.lds_table, opt.width, opt.limbs, opt.blocks);
    int count = 0;
    mk_status st = mk_device_count(&count);
    if (st != MK_OK) return st;
    if (count == 0) return fail(MK_ERR_NO_DEVICE, "no HIP device found");
    if (ordinal < 0 || ordinal >= count) return fail(MK_ERR_NO_DEVICE, "device %d of %d", ordinal, count);

    mk_gpu* g = new mk_gpu();
    g->device = ordinal;
    g->opt = opt;
    g->ks = ks;
    void* bases = nullptr;  // row bases, needed only while the table is built
    auto bail = [&](mk_status s) {
        if (bases) (void)hipFree(bases);
        mk_close(g);
        return s;
    };
#define MK_HIP_G(call)                                                                                    \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return bail(fail(MK_ERR_DRIVER, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__)); \
    } while (0)
    MK_HIP_G(hipSetDevice(ordinal));
    hipDeviceProp_t prop;
    MK_HIP_G(hipGetDeviceProperties(&prop, ordinal));
    snprintf(g->name, sizeof g->name, "%s", prop.gcnArchName);
    g->cus = (u32)prop.multiProcessorCount;
    g->hbm = prop.totalGlobalMem;
    g->grid = g->cus * 4 * opt.blocks;
    // lockstep 1: one workgroup per CU; 2: workgroups of one wave per SIMD; pool: workgroups
    // of `pool` waves sharing an inversion
    g->lock_waves = opt.lds_table ? (MK_WG_LDST | 8)
                    : opt.pool != 0 ? (MK_WG_POOL | opt.pool)
                    : opt.lockstep == 1 ? 4 * opt.blocks
                    : opt.lockstep == 2 ? 4
                                        : 0;
    g->lanes = (u64)g->grid * 64;
    g->rows = rows_of(opt.width);
    g->cols = cols_of(opt.width);
    MK_HIP_G(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    MK_HIP_G(mk_k_noalloc_scratch(ordinal, &g->noalloc));
    for (auto& e : g->ev) MK_HIP_G(hipEventCreate(&e));

    // comb table, built on the device
    g->table_bytes = (u64)g->rows * g->cols * g->ks->entry_words * 4;
    MK_HIP_G(hipMalloc((void**)&g->table, g->table_bytes));
    MK_HIP_G(hipMalloc(&bases, (size_t)(g->rows + 1) * MK_EXT_BYTES));
    MK_HIP_G(g->ks->table_bases(g->stream, opt.width, bases));
    const u64 entries = (u64)g->rows * g->cols;
    const u64 chunk = 1ull << 22;
    for (u64 e0 = 0; e0 < entries; e0 += chunk)
        MK_HIP_G(g->ks->table_fill(g->stream, opt.width, bases, e0, std::min(chunk, entries - e0), g->table));
    MK_HIP_G(hipStreamSynchronize(g->stream));
    MK_HIP_G(hipFree(bases));
    bases = nullptr;

    g->spill_bytes = g->lanes * opt.batch * 20 * 4;  // two field elements per key (mk_lane_tile)
    MK_HIP_G(hipMalloc((void**)&g->spill, g->spill_bytes));
    MK_HIP_G(hipMalloc((void**)&g->ctr, sizeof(mk_counters)));
    MK_HIP_G(hipMalloc((void**)&g->cands, sizeof(mk_cand) * kCandCap));
    MK_HIP_G(hipMalloc((void**)&g->hits, sizeof(mk_hit) * kHitCap));
    MK_HIP_G(hipHostMalloc((void**)&g->h_ctr, sizeof(mk_counters), hipHostMallocDefault));
    MK_HIP_G(hipHostMalloc((void**)&g->h_hits, sizeof(mk_hit) * kHitCap, hipHostMallocDefault));
#undef MK_HIP_G
    st = self_check(g);
    if (st != MK_OK) return bail(st);
    *out = g;
    return MK_OK;
}

mk_status mk_info_get(const mk_gpu* g, mk_info* out) {
    if (!g || !out) return fail(MK_ERR_INVALID, "null pointer");
    memset(out, 0, sizeof *out);
    snprintf(out->name, sizeof out->name, "%s", g->name);
    out->compute_units = g->cus;
    out->grid_waves = g->grid;
    out->hbm_bytes = g->hbm;
    out->table_bytes = g->table_bytes;
    out->scratch_bytes = g->spill_bytes;
    out->options = g->opt;
    out->rows = g->rows;
    out->cols = g->cols;
    return MK_OK;
}

mk_status mk_close(mk_gpu* g) {
    if (!g) return MK_OK;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->table) (void)hipFree(g->table);
    if (g->spill) (void)hipFree(g->spill);
    if (g->ctr) (void)hipFree(g->ctr);
    if (g->cands) (void)hipFree(g->cands);
    if (g->hits) (void)hipFree(g->hits);
    if (g->h_ctr) (void)hipHostFree(g->h_ctr);
    if (g->h_hits) (void)hipHostFree(g->h_hits);
    g->descs.release();
    g->masks.release();
    g->dfas.release();
    g->segs.release();
    g->stop.release();
    for (auto& e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
    return MK_OK;
}

mk_status mk_last_timing(const mk_gpu* g, float* search_ms, float* confirm_ms) {
    if (!g) return fail(MK_ERR_INVALID, "null pointer");
    if (search_ms) *search_ms = g->last_search_ms;
    if (confirm_ms) *confirm_ms = g->last_confirm_ms;
    return MK_OK;
}

uint64_t mk_keys_per_iter(const mk_gpu* g) { return g ? g->lanes * g->opt.batch : 0; }

uint64_t mk_tile_keys(const mk_gpu* g) { return g ? 64ull * g->opt.batch : 0; }

mk_status mk_prepare(mk_gpu* g, const uint8_t key[32], int seed_source, const uint8_t* needle, uint32_t needle_len,
                     uint32_t fold, uint32_t mode, uint32_t stop_on_hit, mk_launch** out) {
    if (!g || !key || !out) return fail(MK_ERR_INVALID, "null pointer");
    *out = nullptr;
    if (seed_source != MK_SEED_CHACHA20 && seed_source != MK_SEED_COUNTER)
        return fail(MK_ERR_INVALID, "unknown seed source %d", seed_source);
    matcher_host m;
    mk_status st = (mk_status)build_matcher(needle, needle_len, fold, mode, m);
    if (st != MK_OK) return st;
    MK_HIP(hipSetDevice(g->device));
    mk_launch* l = new mk_launch();
    l->g = g;
    l->stop_on_hit = stop_on_hit ? 1u : 0u;
    mk_search_desc d{};
    mk_key_to_words(key, d.key);
    d.source = (u32)seed_source;
    d.mode = m.mode;
    d.len = m.len;
    d.stop_on_hit = l->stop_on_hit;
    d.mask_off = 0;
    d.dfa_off = 0;
    const size_t dfa_n = std::max<size_t>(m.dfa.size(), 1);
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = hipMalloc((void**)&l->descs, sizeof d);
    if (e == hipSuccess) e = hipMalloc((void**)&l->masks, sizeof m.mask);
    if (e == hipSuccess) e = hipMalloc((void**)&l->dfas, dfa_n * sizeof(u16));
    if (e == hipSuccess) e = hipMalloc((void**)&l->segs, sizeof(mk_seg));
    if (e == hipSuccess) e = hipMalloc((void**)&l->stop, sizeof(u32));
    if (e == hipSuccess) e = hipMemcpy(l->descs, &d, sizeof d, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(l->masks, m.mask, sizeof m.mask, hipMemcpyHostToDevice);
    if (e == hipSuccess && !m.dfa.empty()) e = hipMemcpy(l->dfas, m.dfa.data(), m.dfa.size() * sizeof(u16), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        mk_free_launch(l);
        return fail(MK_ERR_DRIVER, "prepare: %s", hipGetErrorString(e));
    }
    *out = l;
    return MK_OK;
}

mk_status mk_launch_run(mk_launch* l, uint64_t base, uint32_t iters, mk_hit_rec* hits, uint32_t cap,
                        uint32_t* n_hits) {
    if (!l || !n_hits || (cap && !hits)) return fail(MK_ERR_INVALID, "null pointer");
    mk_gpu* g = l->g;
    MK_HIP(hipSetDevice(g->device));
    mk_seg seg{};
    seg.first = base;
    seg.tile_begin = 0;
    seg.tile_end = (u64)g->grid * iters;
    seg.search = 0;
    MK_HIP(hipMemcpyAsync(l->segs, &seg, sizeof seg, hipMemcpyHostToDevice, g->stream));
    const mk_hit* hs = nullptr;
    u32 n_cands = 0, nh = 0, overflow = 0;
    mk_status st = mk_run_step(g, l->descs, l->masks, l->dfas, l->segs, 1, seg.tile_end, l->stop, 1, hs, n_cands, nh,
                               overflow);
    if (st != MK_OK) return st;
    *n_hits = nh;
    to_records(g, hs, std::min(nh, kHitCap), hits, cap);
    // a stop-on-hit search needs one hit, not the full set: lists that overflowed while
    // the tiles already in flight finished are not an error
    if ((overflow || nh > cap) && !l->stop_on_hit)
        return fail(MK_ERR_OVERFLOW, "%u candidates / %u exact hits exceed the lists (%u / %u, caller %u)", n_cands,
                    nh, kCandCap, kHitCap, cap);
    if (nh > cap) *n_hits = cap;
    return MK_OK;
}

mk_status mk_free_launch(mk_launch* l) {
    if (!l) return MK_OK;
    (void)hipSetDevice(l->g->device);
    if (l->descs) (void)hipFree(l->descs);
    if (l->masks) (void)hipFree(l->masks);
    if (l->dfas) (void)hipFree(l->dfas);
    if (l->segs) (void)hipFree(l->segs);
    if (l->stop) (void)hipFree(l->stop);
    delete l;
    return MK_OK;
}

mk_status mk_derive(mk_gpu* g, const uint8_t key[32], int seed_source, uint64_t first, uint64_t n, uint8_t* pks,
                    uint8_t* seeds) {
    if (!g || !key || (n && !pks)) return fail(MK_ERR_INVALID, "null pointer");
    if (seed_source != MK_SEED_CHACHA20 && seed_source != MK_SEED_COUNTER)
        return fail(MK_ERR_INVALID, "unknown seed source %d", seed_source);
    if (n == 0) return MK_OK;
    MK_HIP(hipSetDevice(g->device));
    u32 kw[8];
    mk_key_to_words(key, kw);
    const u64 chunk = 1ull << 22;
    u32 *dkey = nullptr, *dpk = nullptr, *dseed = nullptr;
    MK_HIP(hipMalloc((void**)&dkey, 32));
    hipError_t e = hipMalloc((void**)&dpk, std::min(n, chunk) * 32);
    if (e == hipSuccess && seeds) e = hipMalloc((void**)&dseed, std::min(n, chunk) * 32);
    if (e == hipSuccess) e = hipMemcpy(dkey, kw, 32, hipMemcpyHostToDevice);
    for (u64 off = 0; e == hipSuccess && off < n; off += chunk) {
        const u64 m = std::min(chunk, n - off);
        e = g->ks->derive(g->stream, g->table, g->opt.width, dkey, (u32)seed_source, first + off, m, dpk, dseed);
        if (e == hipSuccess) e = hipMemcpyAsync(pks + 32 * off, dpk, m * 32, hipMemcpyDeviceToHost, g->stream);
        if (e == hipSuccess && seeds)
            e = hipMemcpyAsync(seeds + 32 * off, dseed, m * 32, hipMemcpyDeviceToHost, g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    }
    (void)hipFree(dkey);
    if (dpk) (void)hipFree(dpk);
    if (dseed) (void)hipFree(dseed);
    if (e != hipSuccess) return fail(MK_ERR_DRIVER, "derive: %s", hipGetErrorString(e));
    return MK_OK;
}

static_assert(MK_DERIVE_NOALLOC_MAX == MK_NOALLOC_KEYS, "the header's limit is the staging array's size");

mk_status mk_derive_noalloc(mk_gpu* g, const uint8_t key[32], int seed_source, uint64_t first, uint32_t n,
                            uint8_t* pks, uint8_t* seeds) {
    if (!g || !key || (n && !pks)) return fail(MK_ERR_INVALID, "null pointer");
    if (seed_source != MK_SEED_CHACHA20 && seed_source != MK_SEED_COUNTER)
        return fail(MK_ERR_INVALID, "unknown seed source %d", seed_source);
    if (n > MK_DERIVE_NOALLOC_MAX)
        return fail(MK_ERR_OUT_OF_RANGE,
                    "mk_derive_noalloc derives at most %u keys per call (%u asked): its device staging is a fixed "
                    "array", MK_DERIVE_NOALLOC_MAX, n);
    if (n == 0) return MK_OK;
    if (!g->noalloc) return fail(MK_ERR_DRIVER, "the staging array of mk_derive_noalloc has no device address");
    MK_HIP(hipSetDevice(g->device));
    mk_key_words kw;
    mk_key_to_words(key, kw.w);
    // public keys in the first half of the array, seeds in the second; every word read
    // back is written by this call's kernel, so nothing is cleared
    u32* dpk = g->noalloc;
    u32* dseed = seeds ? g->noalloc + 8 * MK_NOALLOC_KEYS : nullptr;
    MK_HIP(g->ks->derive_kv(g->stream, g->table, g->opt.width, kw, (u32)seed_source, first, n, dpk, dseed));
    MK_HIP(hipMemcpyAsync(pks, dpk, (size_t)n * 32, hipMemcpyDeviceToHost, g->stream));
    if (seeds) MK_HIP(hipMemcpyAsync(seeds, dseed, (size_t)n * 32, hipMemcpyDeviceToHost, g->stream));
    MK_HIP(hipStreamSynchronize(g->stream));
    return MK_OK;
}

mk_status mk_derive_seeds(mk_gpu* g, const uint8_t* seeds, uint32_t n, uint8_t* pks) {
    if (!g || (n && (!seeds || !pks))) return fail(MK_ERR_INVALID, "null pointer");
    if (n == 0) return MK_OK;
    MK_HIP(hipSetDevice(g->device));
    u32 *ds = nullptr, *dp = nullptr;
    MK_HIP(hipMalloc((void**)&ds, (size_t)n * 32));
    hipError_t e = hipMalloc((void**)&dp, (size_t)n * 32);
    if (e == hipSuccess) e = hipMemcpy(ds, seeds, (size_t)n * 32, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = g->ks->derive_seeds(g->stream, g->table, g->opt.width, ds, n, dp);
    if (e == hipSuccess) e = hipMemcpyAsync(pks, dp, (size_t)n * 32, hipMemcpyDeviceToHost, g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    (void)hipFree(ds);
    if (dp) (void)hipFree(dp);
    if (e != hipSuccess) return fail(MK_ERR_DRIVER, "derive_seeds: %s", hipGetErrorString(e));
    return MK_OK;
}

mk_status mk_table_read(mk_gpu* g, uint32_t row, uint32_t k, uint32_t count, uint32_t* words) {
    if (!g || (count && !words)) return fail(MK_ERR_INVALID, "null pointer");
    if (row >= g->rows || (u64)k + count > g->cols) return fail(MK_ERR_OUT_OF_RANGE, "entry (%u, %u+%u)", row, k, count);
    MK_HIP(hipSetDevice(g->device));
    const u64 off = ((u64)row * g->cols + k) * g->ks->entry_words;
    MK_HIP(hipMemcpy(words, g->table + off, (size_t)count * g->ks->entry_words * 4, hipMemcpyDeviceToHost));
    return MK_OK;
}

mk_status mk_batch_run(mk_gpu* g, const mk_search_spec* specs, uint32_t n, mk_hit_rec* hits, uint32_t cap,
                       uint32_t* n_hits, uint64_t* tries) {
    if (!g || !n_hits || (n && !specs) || (cap && !hits)) return fail(MK_ERR_INVALID, "null pointer");
    *n_hits = 0;
    if (n == 0) return MK_OK;
    MK_HIP(hipSetDevice(g->device));
    const u64 tile_keys = 64ull * g->opt.batch;
    std::vector<mk_search_desc> descs(n);
    std::vector<u64> masks((size_t)n * 64);
    std::vector<u16> dfas;
    std::vector<mk_seg> segs(n);
    u64 tiles = 0;
    for (u32 i = 0; i < n; i++) {
        const mk_search_spec& s = specs[i];
        if (s.seed_source != MK_SEED_CHACHA20 && s.seed_source != MK_SEED_COUNTER)
            return fail(MK_ERR_INVALID, "search %u: unknown seed source %d", i, s.seed_source);
        matcher_host m;
        mk_status st = (mk_status)build_matcher(s.needle, s.needle_len, s.fold, s.mode, m);
        if (st != MK_OK) return fail(st, "search %u: %s", i, g_last_error.c_str());
        mk_search_desc& d = descs[i];
        memset(&d, 0, sizeof d);
        mk_key_to_words(s.key, d.key);
        d.source = (u32)s.seed_source;
        d.mode = m.mode;
        d.len = m.len;
        d.stop_on_hit = s.stop_on_hit ? 1u : 0u;
        d.mask_off = i * 64;
        d.dfa_off = (u32)dfas.size();
        memcpy(&masks[(size_t)i * 64], m.mask, sizeof m.mask);
        dfas.insert(dfas.end(), m.dfa.begin(), m.dfa.end());
        const u64 t = (s.count + tile_keys - 1) / tile_keys;
        segs[i].first = s.first;
        segs[i].tile_begin = tiles;
        segs[i].tile_end = tiles + t;
        segs[i].search = i;
        tiles += t;
        if (tries) tries[i] = t * tile_keys;
    }
    if (dfas.empty()) dfas.push_back(0);
    // segments with no tiles would break the tile -> segment search; keep only non-empty ones
    std::vector<mk_seg> live;
    for (auto& s : segs)
        if (s.tile_end > s.tile_begin) live.push_back(s);
    mk_status st;
    if ((st = g->descs.ensure(sizeof(mk_search_desc) * n)) != MK_OK) return st;
    if ((st = g->masks.ensure(sizeof(u64) * masks.size())) != MK_OK) return st;
    if ((st = g->dfas.ensure(sizeof(u16) * dfas.size())) != MK_OK) return st;
    if ((st = g->segs.ensure(sizeof(mk_seg) * std::max<size_t>(live.size(), 1))) != MK_OK) return st;
    if ((st = g->stop.ensure(sizeof(u32) * n)) != MK_OK) return st;
    MK_HIP(hipMemcpyAsync(g->descs.p, descs.data(), sizeof(mk_search_desc) * n, hipMemcpyHostToDevice, g->stream));
    MK_HIP(hipMemcpyAsync(g->masks.p, masks.data(), sizeof(u64) * masks.size(), hipMemcpyHostToDevice, g->stream));
    MK_HIP(hipMemcpyAsync(g->dfas.p, dfas.data(), sizeof(u16) * dfas.size(), hipMemcpyHostToDevice, g->stream));
    if (!live.empty())
        MK_HIP(hipMemcpyAsync(g->segs.p, live.data(), sizeof(mk_seg) * live.size(), hipMemcpyHostToDevice, g->stream));
    const mk_hit* hs = nullptr;
    u32 n_cands = 0, nh = 0, overflow = 0;
    st = mk_run_step(g, (const mk_search_desc*)g->descs.p, (const u64*)g->masks.p, (const u16*)g->dfas.p,
                     (const mk_seg*)g->segs.p, (u32)live.size(), live.empty() ? 0 : tiles, (u32*)g->stop.p, n, hs,
                     n_cands, nh, overflow);
    if (st != MK_OK) return st;
    *n_hits = nh;
    to_records(g, hs, std::min(nh, kHitCap), hits, cap);
    bool all_stop = true;
    for (u32 i = 0; i < n; i++) all_stop = all_stop && specs[i].stop_on_hit;
    if ((overflow || nh > cap) && !all_stop)
        return fail(MK_ERR_OVERFLOW, "%u candidates / %u exact hits exceed the lists (%u / %u, caller %u)", n_cands,
                    nh, kCandCap, kHitCap, cap);
    if (nh > cap) *n_hits = cap;
    return MK_OK;
}

}  // extern "C"
