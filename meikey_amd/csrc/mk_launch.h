// Host-side launchers of the kernels in mk_kernels.hip (defined there, called by the
// C ABI in mk_host.cpp). Sizes of the extended-point records the table build uses.
#pragma once

#include <hip/hip_runtime.h>

#include "mk_kernels.h"

// mk_ext = 4 field elements x 10 limbs
#define MK_EXT_BYTES (4 * 10 * 4)

hipError_t mk_k_table_bases(hipStream_t s, u32 w, void* bases);
hipError_t mk_k_table_fill(hipStream_t s, u32 w, const void* bases, u64 e0, u64 n, u32* table);
hipError_t mk_k_derive(hipStream_t s, const u32* table, u32 w, const u32* key, u32 source, u64 first, u64 n,
                       u32* out_pk, u32* out_seed);
hipError_t mk_k_derive_kv(hipStream_t s, const u32* table, u32 w, const mk_key_words& key, u32 source, u64 first,
                          u64 n, u32* out_pk, u32* out_seed);
// Address of the file-scope staging array of mk_derive_noalloc on `device` (current).
hipError_t mk_k_noalloc_scratch(int device, u32** out);
hipError_t mk_k_derive_seeds(hipStream_t s, const u32* table, u32 w, const u32* seeds, u32 n, u32* out_pk);
hipError_t mk_k_reset(hipStream_t s, mk_counters* ctr, u32* stop, u32 nstop, const mk_serve_state* serve);
// The search kernel over `grid` waves: one wave per workgroup (lock_waves = 0), or the
// lockstep variant in workgroups of lock_waves waves (4 or 8; grid a multiple of it).
hipError_t mk_k_search(hipStream_t s, u32 grid, u32 lock_waves, const u32* table, u32 w, u32 K,
                       const mk_search_desc* descs, const mk_seg* segs, u32 nsegs, u64 total_tiles, const u64* masks,
                       const u16* dfas, u32* spill, mk_counters* ctr, mk_cand* cands, u32 cand_cap, u32* stop,
                       const mk_serve_state* serve, u64* next_index);
hipError_t mk_k_confirm(hipStream_t s, const u32* table, u32 w, const mk_search_desc* descs, const u64* masks,
                        const u16* dfas, mk_counters* ctr, const mk_cand* cands, u32 cand_cap, mk_hit* hits,
                        u32 hit_cap);
