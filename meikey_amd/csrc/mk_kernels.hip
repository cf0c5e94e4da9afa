// gfx950 kernels of the vanity-key search: comb-table build, key derivation, the
// search kernel and the exact-hit confirmation.
//
// Search kernel (the hot path). A launch is a list of segments (index ranges of one
// or more searches) cut into tiles of 64 lanes x K keys. Each workgroup is one
// wave64; a wave takes the next tile from a launch-wide atomic counter, so searches
// of uneven length share the chip without a static split, and exits when the tiles
// run out. Lane l of a tile owns K consecutive key indices, so the ChaCha20 block of
// an even index also gives the seed of the odd index after it. Per key:
//   ChaCha20 -> SHA-512 + clamp -> comb multiply (rows - 2 extended+Niels additions and
//   a 3-multiplication fraction for the last row) -> y numerator/denominator.
// The K denominators of a lane are inverted together (Montgomery's trick with the
// numerators folded in, mk_lane_tile: one inversion, 4 multiplications per key). Two
// field elements per key wait in a global buffer laid out [key][word][lane] so every
// access of a wave is one contiguous 256-byte line (HBM traffic 160 B per key, against
// about 16 k instructions of integer work per key). y is compressed with the x sign
// still unknown and matched with sign_unknown (superset of the true hits); lanes
// with a candidate append it through one wave ballot and one atomic.
// The confirm kernel re-derives each candidate with x, matches exactly and writes
// exact hits (index, seed, public key).
//
// This file is compiled once per field representation and table layout (mk_common.h):
// as it stands for ten 25.5-bit limbs and 128-byte entries of limbs ("10u"), and through
// mk_kernels_p10.hip (packed 96-byte entries, "10"), mk_kernels_r32.hip (eight 32-bit
// words, "8"), mk_kernels_k26.hip and mk_kernels_k26u.hip (ten 26-bit limbs with Karatsuba
// products, packed "10k" and limbs "10ku"), which set MK_FIELD / MK_PACKED / MK_VARIANT first. Each build lives in its
// own namespace with kernel names suffixed by MK_VARIANT, and exports one mk_kernel_set.
#include <hip/hip_runtime.h>

#include <mutex>

#include "mk_launch.h"

#ifndef MK_VARIANT
#define MK_VARIANT _u10
#define MK_BASE_TU 1
#endif
#define MK_CAT2(a, b) a##b
#define MK_CAT(a, b) MK_CAT2(a, b)
#define MK_K(name) MK_CAT(name, MK_VARIANT)
#define MK_NS MK_K(mk_ns)

namespace MK_NS {

#include "mk_claims.h"
#include "mk_field.h"
#include "mk_hash.h"
#include "mk_match.h"
#include "mk_point.h"
#include "mk_tile.h"

// -------------------------------------------------------------- table build

// Row bases and the top row's offset (mk_row_bases). One thread.
extern "C" __global__ void MK_K(mk_table_bases_kernel)(u32 w, mk_ext* out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    mk_row_bases(w, out);
}

// One thread per entry, entries [e0, e0 + n) of the row-major table.
extern "C" __global__ void __launch_bounds__(64) MK_K(mk_table_fill_kernel)(u32 w, const mk_ext* __restrict__ bases,
                                                                      u64 e0, u64 n, u32* __restrict__ table) {
    const u64 e = e0 + (u64)blockIdx.x * 64 + threadIdx.x;
    if (e >= e0 + n) return;
    const u32 rows = mk_rows(w), cols = mk_cols(w);
    const u32 row = (u32)(e / cols), k = (u32)(e % cols);
    if (row >= rows) return;
    mk_table_entry(w, bases, row, k, table + e * MK_ENTRY_WORDS);
}

// --------------------------------------------------------------- derivation

// Public key (and optionally seed) number i of indices first, first + 1, ...
MK_DEV void mk_derive_one(const u32* __restrict__ table, u32 w, const u32 k[8], u32 source, u64 first, u64 i,
                          u32* __restrict__ out_pk, u32* __restrict__ out_seed) {
    u32 upper[8], seed[8], s[8], pk[8];
    mk_seed_at(k, source, first + i, upper, true, seed);
    mk_secret_scalar(seed, s);
    mk_full_pk(table, w, s, pk);
    for (int j = 0; j < 8; j++) out_pk[i * 8 + j] = pk[j];
    if (out_seed)
        for (int j = 0; j < 8; j++) out_seed[i * 8 + j] = seed[j];
}

// Public keys (and optionally seeds) of indices first .. first + n - 1 of one search
// key. One thread per key.
extern "C" __global__ void __launch_bounds__(64) MK_K(mk_derive_kernel)(const u32* __restrict__ table, u32 w,
                                                                  const u32* __restrict__ key, u32 source,
                                                                  u64 first, u64 n, u32* __restrict__ out_pk,
                                                                  u32* __restrict__ out_seed) {
    const u64 i = (u64)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    u32 k[8];
    for (int j = 0; j < 8; j++) k[j] = key[j];
    mk_derive_one(table, w, k, source, first, i, out_pk, out_seed);
}

// The same with the key passed by value, so the launch needs no device copy of it.
extern "C" __global__ void __launch_bounds__(64) MK_K(mk_derive_kv_kernel)(const u32* __restrict__ table, u32 w,
                                                                     mk_key_words key, u32 source, u64 first,
                                                                     u64 n, u32* __restrict__ out_pk,
                                                                     u32* __restrict__ out_seed) {
    const u64 i = (u64)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    mk_derive_one(table, w, key.w, source, first, i, out_pk, out_seed);
}

// Public keys of explicit 32-byte seeds (known-answer self-check).
extern "C" __global__ void __launch_bounds__(64) MK_K(mk_derive_seeds_kernel)(const u32* __restrict__ table, u32 w,
                                                                        const u32* __restrict__ seeds, u32 n,
                                                                        u32* __restrict__ out_pk) {
    const u32 i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    u32 seed[8], s[8], pk[8];
    for (int j = 0; j < 8; j++) seed[j] = seeds[i * 8 + j];
    mk_secret_scalar(seed, s);
    mk_full_pk(table, w, s, pk);
    for (int j = 0; j < 8; j++) out_pk[i * 8 + j] = pk[j];
}

// ------------------------------------------------------------------- search

__device__ static u64 mk_uniform64(u64 v) {
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
    return (u64)hi << 32 | lo;
}

// Resident waves per SIMD the search kernel's register budget is sized for
// (512 VGPRs per SIMD lane are shared by the waves: 2 waves -> 256 each).
#ifndef MK_SEARCH_WAVES
#define MK_SEARCH_WAVES 2
#endif

// Largest DFA that fits a wave's staging buffers (2 x MK_STAGE_CHUNKS x 1 KB; states x 65
// x 2 bytes): 126 states with 128-byte entries, 94 with 96-byte ones.
#define MK_LDS_DFA_STATES ((MK_STAGE_BUFFERS * MK_STAGE_CHUNKS * 64 * 16) / 130)

// Orders this wave's LDS writes before its later LDS reads by other lanes (the per-wave
// buffers are shared by no other wave, so no workgroup barrier is needed).
__device__ __forceinline__ void mk_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The kernel arguments of both search kernels.
struct mk_search_args {
    const u32* __restrict__ table;
    u32 w, K;
    const mk_search_desc* __restrict__ descs;
    const mk_seg* __restrict__ segs;
    u32 nsegs;
    u64 total_tiles;
    const u64* __restrict__ masks;
    const u16* __restrict__ dfas;
    u32* __restrict__ spill;
    mk_counters* __restrict__ ctr;
    mk_cand* __restrict__ cands;
    u32 cand_cap;
    u32* __restrict__ stop;
    const mk_serve_state* __restrict__ serve;
    u64* __restrict__ next_index;
};

// One wave's LDS and identity.
struct mk_wave_ctx {
    u32 lane;
    u64 gl;     // global lane number (spill column)
    u64 lanes;  // lanes of the grid (spill row stride)
    u32 cached; // search whose bitap table is in lmask
    u64* lmask;
    uint4* stage;
    mk_sha_mid* smid;
    const u32* ltable;  // the comb table in LDS (the LDS-table kernel), else null
};

// Tail balancing (the claim layout and its measurements): mk_claims.h. Workgroup claims
// (the lockstep and pooled kernels below) cut one tile per wave of the tail into units.

// One claim of one wave: tile `tile` (sub-unit `sub` of a tail tile) with Keff keys per
// lane; nth is the claim's number in claim order (serving deals claims round-robin over
// the active requests). Returns false, having searched nothing, when the claim's search
// has stopped; otherwise the comb multiply calls row_sync() Keff x (rows - 2) times and
// the batched inversion calls invert(P) once.
// With LDST the comb reads its entries from the LDS copy of the table (cx.ltable) instead
// of staging them from HBM.
template <bool LDST, class R, class I>
__device__ __forceinline__ bool mk_search_claim(const mk_search_args& a, mk_wave_ctx& cx, u64 tile, u64 sub,
                                                u32 Keff, u64 nth, R row_sync, I invert) {
    const u32 lane = cx.lane;
    u32 search;
    u64 tile_first;
    if (a.serve != nullptr) {
        // serving: round-robin over the active slots, indices from the slot's counter
        search = __builtin_amdgcn_readfirstlane(a.serve->slots[nth % a.serve->n_active]);
        // the decision is taken once for the whole wave (readfirstlane), so the lockstep
        // kernel's barrier count never differs between its lanes
        if (__builtin_amdgcn_readfirstlane(__atomic_load_n(&a.stop[search], __ATOMIC_RELAXED)) != 0) return false;
        u64 f = 0;
        if (lane == 0) f = atomicAdd((unsigned long long*)&a.next_index[search], 64ull * Keff);
        tile_first = mk_uniform64(f);
    } else {
        // segment holding this tile (binary search; segments are few and in order)
        u32 lo = 0, hi = a.nsegs;
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (a.segs[mid].tile_begin <= tile) lo = mid;
            else hi = mid;
        }
        const mk_seg sg = a.segs[lo];
        search = sg.search;
        tile_first = sg.first + (tile - sg.tile_begin) * 64ull * a.K + sub * 64ull * Keff;
    }
    const mk_search_desc& dsc = a.descs[search];
    // A serving unit has already advanced its request's next index, so it is searched
    // even if the request stopped meanwhile: skipping it would leave a hole of
    // unsearched indices that still count as tries (the stop flag was checked before
    // the claim, and the request may go on after a candidate that the exact check
    // rejects).
    if (a.serve == nullptr && dsc.stop_on_hit &&
        __builtin_amdgcn_readfirstlane(__atomic_load_n(&a.stop[search], __ATOMIC_RELAXED)) != 0)
        return false;
    if (search != cx.cached) {
        mk_wave_sync();
        cx.lmask[lane] = a.masks[(u64)dsc.mask_off + lane];
        mk_wave_sync();
        cx.cached = search;
    }
    mk_matcher mt;
    mt.mode = dsc.mode;
    mt.len = dsc.len;
    mt.mask = cx.lmask;
    mt.dfa = a.dfas + dsc.dfa_off;
    u32 key[8];
    MK_UNROLL
    for (int i = 0; i < 8; i++) key[i] = dsc.key[i];
    const u64 first = tile_first + (u64)lane * Keff;
    const u32 stop_on_hit = dsc.stop_on_hit;
    if (dsc.source == 1) {
        // the counter seeds' SHA-512 mid-state, the same for the whole wave
        mk_sha_mid m;
        mk_sha_mid_init(key, m);
        mk_wave_sync();
        if (lane == 0) *cx.smid = m;
        mk_wave_sync();
    }
    mk_spill sp;
    sp.base = a.spill + cx.gl;
    sp.stride = cx.lanes;
    uint4* stage = cx.stage;
    auto mul = [&](const u32 s[8], fe& yn, fe& yd) {
        fe xn, xd;
        if constexpr (LDST)
            mk_mul_base_frac_near(cx.ltable, a.w, s, yn, yd, xn, xd);
        else
            mk_mul_base_frac_staged(a.table, a.w, s, stage, lane, yn, yd, xn, xd, row_sync);
    };
    // During the back-sweep the staging buffers are idle: a DFA of up to
    // MK_LDS_DFA_STATES states is copied into them so its 43 dependent lookups per
    // key hit LDS instead of L2.
    auto pre_sweep = [&] {
        if (!LDST && mt.mode == MK_MODE_REGEX && mt.len <= MK_LDS_DFA_STATES) {
            u16* ldfa = (u16*)stage;
            const u16* src = a.dfas + dsc.dfa_off;
            for (u32 i = lane; i < mt.len * 65; i += 64) ldfa[i] = src[i];
            mk_wave_sync();
            mt.dfa = ldfa;
        }
    };
    // Loaded where each key's hash needs it, not kept in registers through the comb: the
    // empty asm with a memory clobber keeps the compiler from hoisting the LDS reads out
    // of the key loop. (A volatile pointer did the same through serialized flat loads.)
    const mk_sha_mid* smid = cx.smid;
    auto mid = [&]() {
        asm volatile("" ::: "memory");
        mk_sha_mid m;
        MK_UNROLL
        for (int i = 0; i < 8; i++) m.v[i] = smid->v[i];
        m.w2 = smid->w2;
        m.w16 = smid->w16;
        m.w17 = smid->w17;
        return m;
    };
    mk_counters* ctr = a.ctr;
    mk_cand* cands = a.cands;
    const u32 cand_cap = a.cand_cap;
    u32* stop = a.stop;
    // Append the candidates of the lanes with `cand` set (index `index` each).
    auto append = [&](bool cand, u64 index) {
        const u64 bal = __ballot(cand);
        if (bal == 0) return;
        const u32 leader = (u32)__builtin_ctzll(bal);
        const u32 below = __builtin_amdgcn_mbcnt_hi((u32)(bal >> 32), __builtin_amdgcn_mbcnt_lo((u32)bal, 0u));
        u32 base = 0;
        if (lane == leader) base = atomicAdd(&ctr->n_cands, (u32)__builtin_popcountll(bal));
        base = __shfl(base, (int)leader);
        if (cand) {
            const u32 slot = base + below;
            if (slot < cand_cap) {
                cands[slot].search = search;
                cands[slot].pad = 0;
                cands[slot].index = index;
            } else {
                atomicOr(&ctr->overflow, 1u);
            }
        }
    };
    // A stop-on-hit search needs only its lowest exact hit. Per lane that is its lowest
    // candidate whose match holds for either x sign (the back-sweep meets a lane's keys from
    // the highest index down, so the last such candidate seen is the lowest), appended once
    // at the end of the claim, and the candidates whose match depends on the sign, appended
    // as found for the exact check to decide. A pattern that matches most keys then appends
    // about one candidate per lane instead of one per key, and the list does not overflow
    // and drop a search's lowest candidates.
    u64 lane_best = ~0ull;
    mk_lane_tile(Keff, key, dsc.source, first, sp, mul, pre_sweep, [&](u32 jj, const u32 pk[8]) {
        bool cand = mk_pk_match(pk, mt, true);
        const u64 bal = __ballot(cand);
        if (bal == 0) return;
        if (stop_on_hit) {
            if (lane == (u32)__builtin_ctzll(bal)) atomicExch(&stop[search], 1u);
            if (cand) {
                u32 pk1[8];
                MK_UNROLL
                for (int i = 0; i < 8; i++) pk1[i] = pk[i];
                pk1[7] |= 0x80000000u;  // the x sign, bit 1 of the value at the sign position
                if (mk_pk_match(pk, mt, false) && mk_pk_match(pk1, mt, false)) {
                    lane_best = first + jj;
                    cand = false;
                }
            }
        }
        append(cand, first + jj);
    }, mid, invert);
    if (stop_on_hit) append(lane_best != ~0ull, lane_best);
    return true;
}

// The lane's own inversion (one per lane per claim).
struct mk_lane_invert {
    __device__ __forceinline__ fe operator()(const fe& p) const { return fe_invert(p); }
};

// The search kernel: one wave per workgroup, each wave claims its own tiles.
// The claim loop of a wave that claims on its own (`waves` = waves of the grid).
template <bool LDST>
__device__ __forceinline__ void mk_wave_claims(const mk_search_args& a, mk_wave_ctx& cx, u64 waves) {
    const u32 serve_unit = a.serve != nullptr ? __builtin_amdgcn_readfirstlane(a.serve->unit_keys) : 0u;
    const mk_claim_plan plan = mk_plan_claims(a.K, serve_unit, a.total_tiles, waves);
    for (;;) {
        u64 claim = 0;
        if (cx.lane == 0) {
            const u64 t = atomicAdd((unsigned long long*)&a.ctr->queue, 1ull);
            claim = mk_claim_number(plan, t, t < plan.big_tiles ? 0 : atomicAdd((unsigned long long*)&a.ctr->queue2, 1ull));
        }
        claim = mk_uniform64(claim);
        u64 tile, sub;
        u32 Keff;
        if (!mk_claim_at(plan, claim, tile, sub, Keff)) break;
        if (a.serve != nullptr && a.serve->n_active == 0) break;
        // claim number: whole tiles first, then the units of the tail
        const u64 nth = tile < plan.big_tiles ? tile : plan.big_tiles + (tile - plan.big_tiles) * plan.U + sub;
        mk_search_claim<LDST>(a, cx, tile, sub, Keff, nth, [] {}, mk_lane_invert());
    }
}

extern "C" __global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_kernel)(mk_search_args a) {
    __shared__ u64 lmask[64];
    __shared__ uint4 stage[MK_STAGE_BUFFERS * MK_STAGE_CHUNKS * 64];
    __shared__ mk_sha_mid smid;
    mk_wave_ctx cx;
    cx.lane = threadIdx.x;
    cx.lanes = (u64)gridDim.x * 64;
    cx.gl = (u64)blockIdx.x * 64 + cx.lane;
    cx.cached = 0xffffffffu;
    cx.lmask = lmask;
    cx.stage = stage;
    cx.smid = &smid;
    cx.ltable = nullptr;
    mk_wave_claims<false>(a, cx, gridDim.x);
}

// The LDS-table search kernel: one workgroup of 8 waves per CU copies the whole comb table
// (at most MK_LDS_TABLE_BYTES: W <= 5, or W = 6 with 96-byte entries) into LDS once, then
// every wave claims its own tiles as in mk_search_kernel, its comb reading entries from
// LDS instead of gathering them from HBM / L2. The price is the field work of the small
// width's rows.
extern "C" __global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_ldst_kernel)(mk_search_args a) {
    __shared__ uint4 ltab[MK_LDS_TABLE_BYTES / 16];
    __shared__ u64 lmask[8][64];
    __shared__ mk_sha_mid smid[8];
    const u32 n16 = mk_rows(a.w) * mk_cols(a.w) * (MK_ENTRY_WORDS / 4);  // checked on the host
    const uint4* src = (const uint4*)a.table;
    for (u32 i = threadIdx.x; i < n16; i += 512) ltab[i] = src[i];
    __syncthreads();
    const u32 wave = threadIdx.x / 64;
    mk_wave_ctx cx;
    cx.lane = threadIdx.x % 64;
    cx.lanes = (u64)gridDim.x * 512;
    cx.gl = (u64)blockIdx.x * 512 + threadIdx.x;
    cx.cached = 0xffffffffu;
    cx.lmask = lmask[wave];
    cx.stage = nullptr;
    cx.smid = &smid[wave];
    cx.ltable = (const u32*)ltab;
    mk_wave_claims<true>(a, cx, (u64)gridDim.x * 8);
}

// The inversion shared by the NW waves of a workgroup (the reference's block-pooled
// inversion, kernel.cu:1529-1541, 2024-2028): lane l of every wave puts its product P in
// LDS; lane l of wave 0 inverts the product of the NW values of lane l (Montgomery:
// NW - 1 prefix products, one inversion, 2 (NW - 1) multiplications back) and hands each
// wave its own 1/P. One inversion's instruction stream per workgroup instead of one per
// wave; the other waves wait at the barrier, and their SIMDs run the other resident
// waves meanwhile. Values live in the waves' staging buffers, which are idle between a
// key's last comb row and the next key's first (every copy into them has been waited
// for). Every wave of the workgroup calls this once per claim round, a wave with nothing
// to search with P = 1; two barriers.
template <u32 NW>
__device__ __forceinline__ fe mk_pool_invert(const fe& p, uint4 (*stage)[2 * MK_STAGE_CHUNKS * 64], u32 wave,
                                             u32 lane) {
    auto put = [&](u32 w, u32 slot, const fe& v) {
        u32* r = (u32*)stage[w] + slot * MK_FE_LIMBS * 64 + lane;
        MK_UNROLL
        for (int i = 0; i < MK_FE_LIMBS; i++) r[i * 64] = v.v[i];
    };
    auto get = [&](u32 w, u32 slot) {
        const u32* r = (const u32*)stage[w] + slot * MK_FE_LIMBS * 64 + lane;
        fe v;
        MK_UNROLL
        for (int i = 0; i < MK_FE_LIMBS; i++) v.v[i] = r[i * 64];
        return v;
    };
    put(wave, 0, p);
    __syncthreads();
    if (wave == 0) {
        // slot 1 of wave w: the prefix product P_0 ... P_w
        fe acc = get(0, 0);
        put(0, 1, acc);
        for (u32 w = 1; w < NW; w++) {
            acc = fe_mul(get(w, 0), acc);
            put(w, 1, acc);
        }
        fe inv = fe_invert(acc);
        for (u32 w = NW - 1; w > 0; w--) {
            const fe q = get(w, 0);
            put(w, 0, fe_mul(get(w - 1, 1), inv));  // 1 / P_w
            inv = fe_mul(q, inv);                     // 1 / (P_0 ... P_(w-1))
        }
        put(0, 0, inv);
    }
    __syncthreads();
    return get(wave, 0);
}

// The lockstep search kernel: one workgroup of 4 x (waves per SIMD) waves per CU. The
// workgroup claims a tile (or tail unit) for each of its waves at once, all with the same
// number of keys per lane, and its waves pass a barrier at every comb row, so the waves of
// a CU gather from one or two rows of the table at a time instead of all of them: the
// CU's address-translation cache then covers what it reads even for tables far larger
// than its reach. A wave whose claim has nothing to search passes the same barriers
// (Keff x (rows - 2) of them) without work, so every wave of the workgroup executes the
// same barrier sequence. Every branch that decides how many barriers a wave passes
// (the claim, the end of the queue, an empty serving step) reads values that are equal
// for the whole workgroup.
//
// With POOL the waves share one inversion per claim round (mk_pool_invert); without ROWS
// there is no barrier per comb row.
template <u32 NW, bool ROWS, bool POOL>
__device__ __forceinline__ void mk_search_lock(const mk_search_args& a) {
    __shared__ u64 lmask[NW][64];
    __shared__ uint4 stage[NW][2 * MK_STAGE_CHUNKS * 64];
    __shared__ mk_sha_mid smid[NW];
    __shared__ u64 wg_claim[2];
    const u32 nw = NW;  // waves of this workgroup
    const u32 wave = threadIdx.x / 64;
    mk_wave_ctx cx;
    cx.lane = threadIdx.x % 64;
    cx.lanes = (u64)gridDim.x * blockDim.x;
    cx.gl = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    cx.cached = 0xffffffffu;
    cx.lmask = lmask[wave];
    cx.stage = stage[wave];
    cx.smid = &smid[wave];
    cx.ltable = nullptr;
    const u32 K = a.K;
    const u32 U = mk_units_per_tile(K, a.serve != nullptr ? __builtin_amdgcn_readfirstlane(a.serve->unit_keys) : 0u);
    const u64 total_tiles = a.total_tiles;
    const u64 waves = (u64)gridDim.x * nw;
    // whole tiles come in groups of nw, so a group never straddles the tail
    u64 big_tiles = total_tiles - (U > 1 ? (total_tiles < waves ? total_tiles : waves) : 0);
    if (U > 1) big_tiles -= big_tiles % nw;
    const u64 total_claims = big_tiles + (total_tiles - big_tiles) * U;  // whole tiles, then units
    const u32 rows = mk_rows(a.w);
    for (u32 round = 0;; round++) {
        if (threadIdx.x == 0) {
            const u64 t = atomicAdd((unsigned long long*)&a.ctr->queue, (unsigned long long)nw);
            wg_claim[round & 1] =
                t < big_tiles ? t : big_tiles + atomicAdd((unsigned long long*)&a.ctr->queue2, (unsigned long long)nw);
        }
        __syncthreads();
        const u64 c0 = mk_uniform64(wg_claim[round & 1]);  // first claim of the group, the same for all waves
        if (c0 >= total_claims) break;
        if (a.serve != nullptr && __builtin_amdgcn_readfirstlane(a.serve->n_active) == 0) break;
        const bool whole = c0 < big_tiles;
        const u32 Keff = whole ? K : K / U;
        const u64 c = c0 + wave;  // this wave's claim number
        const u64 tile = whole ? c : big_tiles + (c - big_tiles) / U;
        const u64 sub = whole ? 0 : (c - big_tiles) % U;
        bool searched = false;
        auto row_sync = [] {
            if constexpr (ROWS) __builtin_amdgcn_s_barrier();
        };
        auto invert = [&](const fe& p) {
            if constexpr (POOL) return mk_pool_invert<NW>(p, stage, wave, cx.lane);
            else return fe_invert(p);
        };
        if (c < total_claims && (!whole || c < big_tiles))
            searched = mk_search_claim<false>(a, cx, tile, sub, Keff, c, row_sync, invert);
        if (!searched) {
            if constexpr (ROWS)
                for (u32 j = 0; j < Keff * (rows - 2); j++) __builtin_amdgcn_s_barrier();
            if constexpr (POOL) (void)mk_pool_invert<NW>(fe_small(1), stage, wave, cx.lane);
        }
    }
}

// Workgroups of 8 waves (2 per SIMD: one per CU) and of 4 waves (1 per SIMD; with 2 waves
// per SIMD, two workgroups share a CU and step through the rows independently).
extern "C" __global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_lock8_kernel)(mk_search_args a) {
    mk_search_lock<8, true, false>(a);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_lock4_kernel)(mk_search_args a) {
    mk_search_lock<4, true, false>(a);
}
// The pooled inversion over workgroups of 8 and of 4 waves, without row barriers.
extern "C" __global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_pool8_kernel)(mk_search_args a) {
    mk_search_lock<8, false, true>(a);
}
extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MK_SEARCH_WAVES, MK_SEARCH_WAVES)))
    MK_K(mk_search_pool4_kernel)(mk_search_args a) {
    mk_search_lock<4, false, true>(a);
}
// Exact check of every candidate: full key with its x sign, matched with the sign
// known. One thread per candidate slot.
extern "C" __global__ void __launch_bounds__(64)
    MK_K(mk_confirm_kernel)(const u32* __restrict__ table, u32 w, const mk_search_desc* __restrict__ descs,
                      const u64* __restrict__ masks, const u16* __restrict__ dfas, mk_counters* __restrict__ ctr,
                      const mk_cand* __restrict__ cands, u32 cand_cap, mk_hit* __restrict__ hits, u32 hit_cap) {
    const u32 i = blockIdx.x * 64 + threadIdx.x;
    const u32 n = ctr->n_cands < cand_cap ? ctr->n_cands : cand_cap;
    if (i >= n) return;
    const mk_cand c = cands[i];
    const mk_search_desc& dsc = descs[c.search];
    u32 key[8], upper[8], seed[8], s[8], pk[8];
    for (int j = 0; j < 8; j++) key[j] = dsc.key[j];
    mk_seed_at(key, dsc.source, c.index, upper, true, seed);
    mk_secret_scalar(seed, s);
    mk_full_pk(table, w, s, pk);
    mk_matcher mt;
    mt.mode = dsc.mode;
    mt.len = dsc.len;
    mt.mask = masks + dsc.mask_off;
    mt.dfa = dfas + dsc.dfa_off;
    if (!mk_pk_match(pk, mt, false)) return;
    const u32 slot = atomicAdd(&ctr->n_hits, 1u);
    if (slot >= hit_cap) {
        atomicOr(&ctr->overflow, 2u);
        return;
    }
    mk_hit& h = hits[slot];
    h.search = c.search;
    h.pad = 0;
    h.index = c.index;
    for (int j = 0; j < 8; j++) h.seed[j] = seed[j];
    for (int j = 0; j < 8; j++) h.pk[j] = pk[j];
}

// ------------------------------------------------------------ host launchers

static hipError_t mk_k_table_bases(hipStream_t s, u32 w, void* bases) {
    hipLaunchKernelGGL(MK_K(mk_table_bases_kernel), dim3(1), dim3(64), 0, s, w, (mk_ext*)bases);
    return hipGetLastError();
}

static hipError_t mk_k_table_fill(hipStream_t s, u32 w, const void* bases, u64 e0, u64 n, u32* table) {
    const u64 blocks = (n + 63) / 64;
    hipLaunchKernelGGL(MK_K(mk_table_fill_kernel), dim3((u32)blocks), dim3(64), 0, s, w, (const mk_ext*)bases, e0, n, table);
    return hipGetLastError();
}

static hipError_t mk_k_derive(hipStream_t s, const u32* table, u32 w, const u32* key, u32 source, u64 first, u64 n,
                       u32* out_pk, u32* out_seed) {
    const u64 blocks = (n + 63) / 64;
    hipLaunchKernelGGL(MK_K(mk_derive_kernel), dim3((u32)blocks), dim3(64), 0, s, table, w, key, source, first, n, out_pk,
                       out_seed);
    return hipGetLastError();
}

static hipError_t mk_k_derive_kv(hipStream_t s, const u32* table, u32 w, const mk_key_words& key, u32 source, u64 first,
                          u64 n, u32* out_pk, u32* out_seed) {
    const u64 blocks = (n + 63) / 64;
    hipLaunchKernelGGL(MK_K(mk_derive_kv_kernel), dim3((u32)blocks), dim3(64), 0, s, table, w, key, source, first, n,
                       out_pk, out_seed);
    return hipGetLastError();
}

static hipError_t mk_k_derive_seeds(hipStream_t s, const u32* table, u32 w, const u32* seeds, u32 n, u32* out_pk) {
    hipLaunchKernelGGL(MK_K(mk_derive_seeds_kernel), dim3((n + 63) / 64), dim3(64), 0, s, table, w, seeds, n, out_pk);
    return hipGetLastError();
}

static hipError_t mk_k_search(hipStream_t s, u32 grid, u32 lock_waves, const u32* table, u32 w, u32 K,
                       const mk_search_desc* descs, const mk_seg* segs, u32 nsegs, u64 total_tiles, const u64* masks,
                       const u16* dfas, u32* spill, mk_counters* ctr, mk_cand* cands, u32 cand_cap, u32* stop,
                       const mk_serve_state* serve, u64* next_index) {
    mk_search_args a;
    a.table = table;
    a.w = w;
    a.K = K;
    a.descs = descs;
    a.segs = segs;
    a.nsegs = nsegs;
    a.total_tiles = total_tiles;
    a.masks = masks;
    a.dfas = dfas;
    a.spill = spill;
    a.ctr = ctr;
    a.cands = cands;
    a.cand_cap = cand_cap;
    a.stop = stop;
    a.serve = serve;
    a.next_index = next_index;
    if (lock_waves == 0) {
        hipLaunchKernelGGL(MK_K(mk_search_kernel), dim3(grid), dim3(64), 0, s, a);
    } else if (lock_waves == (MK_WG_LDST | 8)) {
        const u64 table_bytes = (u64)((251 + w - 1) / w) * ((1u << (w - 1)) + 1) * MK_ENTRY_WORDS * 4;
        if (grid % 8 != 0 || table_bytes > MK_LDS_TABLE_BYTES)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(MK_K(mk_search_ldst_kernel), dim3(grid / 8), dim3(512), 0, s, a);
    } else {
        // grid waves in workgroups of nw (4 or 8, dividing the grid)
        const u32 nw = lock_waves & ~MK_WG_POOL;
        const bool pool = (lock_waves & MK_WG_POOL) != 0;
        if ((nw != 4 && nw != 8) || grid % nw != 0) return hipErrorInvalidValue;
        if (pool && nw == 8)
            hipLaunchKernelGGL(MK_K(mk_search_pool8_kernel), dim3(grid / 8), dim3(512), 0, s, a);
        else if (pool)
            hipLaunchKernelGGL(MK_K(mk_search_pool4_kernel), dim3(grid / 4), dim3(256), 0, s, a);
        else if (nw == 8)
            hipLaunchKernelGGL(MK_K(mk_search_lock8_kernel), dim3(grid / 8), dim3(512), 0, s, a);
        else
            hipLaunchKernelGGL(MK_K(mk_search_lock4_kernel), dim3(grid / 4), dim3(256), 0, s, a);
    }
    return hipGetLastError();
}

static hipError_t mk_k_confirm(hipStream_t s, const u32* table, u32 w, const mk_search_desc* descs, const u64* masks,
                        const u16* dfas, mk_counters* ctr, const mk_cand* cands, u32 cand_cap, mk_hit* hits,
                        u32 hit_cap) {
    hipLaunchKernelGGL(MK_K(mk_confirm_kernel), dim3((cand_cap + 63) / 64), dim3(64), 0, s, table, w, descs, masks, dfas,
                       ctr, cands, cand_cap, hits, hit_cap);
    return hipGetLastError();
}

}  // namespace MK_NS

#if !defined(__HIP_DEVICE_COMPILE__)  // host code only: the device pass has no launchers
extern const mk_kernel_set MK_K(mk_kernels);
const mk_kernel_set MK_K(mk_kernels) = {
    MK_LIMBS_CODE,
    MK_ENTRY_WORDS,
    (u32)sizeof(MK_NS::mk_ext),
    MK_NS::mk_k_table_bases,
    MK_NS::mk_k_table_fill,
    MK_NS::mk_k_derive,
    MK_NS::mk_k_derive_kv,
    MK_NS::mk_k_derive_seeds,
    MK_NS::mk_k_search,
    MK_NS::mk_k_confirm,
};
#endif

#ifdef MK_BASE_TU
// -------------------------------------------- layout-independent kernels (one copy)

// Staging of the allocation-free derivation (mk_derive_noalloc): MK_NOALLOC_KEYS public
// keys, then as many seeds. File scope, one copy per device for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress gives per device.
static __device__ __align__(256) u32 scratch[MK_NOALLOC_WORDS];

// Clears the launch counters and the first nstop stop flags with a kernel (a plain
// fill rather than a memset node, so the step can be captured in a graph).
// With `serve`, the stop flags of the step's active slots are cleared instead.
extern "C" __global__ void mk_reset_kernel(mk_counters* ctr, u32* stop, u32 nstop, const mk_serve_state* serve) {
    const u32 i = blockIdx.x * 64 + threadIdx.x;
    if (i == 0) {
        ctr->queue = 0;
        ctr->queue2 = 0;
        ctr->n_cands = 0;
        ctr->n_hits = 0;
        ctr->overflow = 0;
    }
    if (serve != nullptr) {
        if (i < serve->n_active && i < MK_SERVE_MAX_SLOTS) stop[serve->slots[i]] = 0;
    } else if (stop != nullptr && i < nstop) {
        stop[i] = 0;
    }
}

// Device address of `scratch` on the current device, looked up once per device (each
// device holds its own copy at its own address) and kept for the life of the process.
hipError_t mk_k_noalloc_scratch(int device, u32** out) {
    static std::mutex mu;
    static u32* ptrs[MK_MAX_DEVICES] = {};
    if (device < 0 || device >= MK_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (ptrs[device] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptrs[device] = (u32*)p;
    }
    *out = ptrs[device];
    return hipSuccess;
}

hipError_t mk_k_reset(hipStream_t s, mk_counters* ctr, u32* stop, u32 nstop, const mk_serve_state* serve) {
    hipLaunchKernelGGL(mk_reset_kernel, dim3(nstop > 64 ? (nstop + 63) / 64 : 1), dim3(64), 0, s, ctr, stop, nstop,
                       serve);
    return hipGetLastError();
}
#endif  // MK_BASE_TU
