// Host-side launchers of the kernels in mk_kernels.hip (defined there, called by the
// C ABI in mk_host.cpp and mk_serve.cpp). The kernels that depend on the field
// representation and table layout come as one mk_kernel_set per layout.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/meikey_hip.h"
#include "mk_kernels.h"

// mk_kernel_set::search lock_waves flags: workgroups of 4 or 8 waves sharing one inversion
// (MK_WG_POOL | 4 or 8); the comb table copied into LDS, workgroups of 8 waves
// (MK_WG_LDST | 8).
#define MK_WG_POOL 0x100u
#define MK_WG_LDST 0x200u

struct mk_kernel_set {
    u32 limbs;        // mk_options.limbs code of this layout
    u32 entry_words;  // 32-bit words per comb-table entry
    u32 ext_bytes;    // bytes of one extended point (the table build's row bases)
    hipError_t (*table_bases)(hipStream_t s, u32 w, void* bases);
    hipError_t (*table_fill)(hipStream_t s, u32 w, const void* bases, u64 e0, u64 n, u32* table);
    hipError_t (*derive)(hipStream_t s, const u32* table, u32 w, const u32* key, u32 source, u64 first, u64 n,
                         u32* out_pk, u32* out_seed);
    hipError_t (*derive_kv)(hipStream_t s, const u32* table, u32 w, const mk_key_words& key, u32 source, u64 first,
                            u64 n, u32* out_pk, u32* out_seed);
    hipError_t (*derive_seeds)(hipStream_t s, const u32* table, u32 w, const u32* seeds, u32 n, u32* out_pk);
    // The search kernel over `grid` waves: one wave per workgroup (lock_waves = 0), the
    // lockstep variant in workgroups of lock_waves waves (4 or 8; grid a multiple of it),
    // with MK_WG_POOL set, workgroups of 4 or 8 waves that share one inversion, or with
    // MK_WG_LDST, 8-wave workgroups reading the comb table from LDS.
    hipError_t (*search)(hipStream_t s, u32 grid, u32 lock_waves, const u32* table, u32 w, u32 K,
                         const mk_search_desc* descs, const mk_seg* segs, u32 nsegs, u64 total_tiles,
                         const u64* masks, const u16* dfas, u32* spill, mk_counters* ctr, mk_cand* cands,
                         u32 cand_cap, u32* stop, const mk_serve_state* serve, u64* next_index);
    hipError_t (*confirm)(hipStream_t s, const u32* table, u32 w, const mk_search_desc* descs, const u64* masks,
                          const u16* dfas, mk_counters* ctr, const mk_cand* cands, u32 cand_cap, mk_hit* hits,
                          u32 hit_cap);
};

// "10u": ten 25.5-bit limbs, 128-byte entries of limbs; "10": the same arithmetic with
// packed 96-byte entries; "8": eight 32-bit words, packed entries; "10k" / "10ku": ten
// 26-bit limbs with Karatsuba products, packed entries / entries of limbs.
extern const mk_kernel_set mk_kernels_u10;
extern const mk_kernel_set mk_kernels_p10;
extern const mk_kernel_set mk_kernels_r32;
extern const mk_kernel_set mk_kernels_k26;
extern const mk_kernel_set mk_kernels_k26u;

// Largest extended point of any layout (4 field elements x 10 limbs).
#define MK_EXT_BYTES (4 * 10 * 4)

// Address of the file-scope staging array of mk_derive_noalloc on `device` (current).
hipError_t mk_k_noalloc_scratch(int device, u32** out);
hipError_t mk_k_reset(hipStream_t s, mk_counters* ctr, u32* stop, u32 nstop, const mk_serve_state* serve);
