// Device-side data structures shared by the kernels and the host launch code.
#pragma once

#include "mk_common.h"

// One search: its seed key, how seeds are made and what it matches.
struct mk_search_desc {
    u32 key[8];       // ChaCha20 key, or the counter base (little-endian words of the 32 bytes)
    u32 source;       // 0 = ChaCha20 seeds, 1 = counter seeds (low-security)
    u32 mode;         // MK_MODE_CONTAINS / SUFFIX / REGEX
    u32 len;          // needle length, or DFA state count for MK_MODE_REGEX
    u32 stop_on_hit;  // skip the rest of this search's tiles once it has a candidate
    u32 mask_off;     // offset of its 64-entry bitap table in the mask pool (u64 units)
    u32 dfa_off;      // offset of its DFA table in the DFA pool (u16 units)
    u32 pad[2];
};

// A contiguous index range of one search, cut into tiles of 64 * keys_per_thread keys.
// Segments of one launch cover tiles [tile_begin, tile_end) of the launch in order.
struct mk_seg {
    u64 first;       // key index of the segment's first tile
    u64 tile_begin;  // launch-wide tile number where the segment starts
    u64 tile_end;
    u32 search;      // index into the search descriptors
    u32 pad;
};

// A search-kernel candidate: the key may match (its x sign is not yet known).
struct mk_cand {
    u32 search;
    u32 pad;
    u64 index;
};

// An exact hit: the key matches with its real x sign. The fields are in the order of
// the caller's record (mk_hit_rec: index, search, pad, seed bytes, pk bytes), and the
// words are little-endian, so on the host a hit converts to a record by one copy.
struct mk_hit {
    u64 index;
    u32 search;
    u32 pad;
    u32 seed[8];
    u32 pk[8];
};

// Serving: the active request slots of a step. Tile t of the step goes to slot
// slots[t % n_active]; the slot's next index (next_index[slot], device-resident) is
// advanced atomically by each tile it receives. The step is claimed in units of
// unit_keys keys per lane.
#define MK_SERVE_MAX_SLOTS 4096
struct mk_serve_state {
    u32 n_active;
    u32 unit_keys;
    u32 slots[MK_SERVE_MAX_SLOTS];
};

// Launch-wide counters, cleared before each step.
struct mk_counters {
    u64 queue;      // next whole tile to hand out
    u64 queue2;     // next tail unit to hand out (tiles cut into smaller units at the end)
    u32 n_cands;    // candidates appended (may exceed the buffer; the excess is dropped and flagged)
    u32 n_hits;     // exact hits appended
    u32 overflow;   // set when candidates or hits did not fit
    u32 pad;
};

// A search key's eight little-endian words, passed to a kernel by value.
struct mk_key_words {
    u32 w[8];
};

// Staging of the allocation-free derivation: public keys and seeds of at most
// MK_NOALLOC_KEYS indices per call (2 x 2 MiB).
#define MK_NOALLOC_KEYS (1u << 16)
#define MK_NOALLOC_WORDS (2u * 8u * MK_NOALLOC_KEYS)
// Devices whose staging address is kept.
#define MK_MAX_DEVICES 64

// Largest comb table the LDS-table search kernel holds in LDS (with its 8 waves' bitap
// masks and SHA mid-states the workgroup stays below the CU's 160 KB): W <= 5 in any
// layout (W = 5: 51 x 17 entries, 111 KB of 128-byte entries), W = 6 with 96-byte entries
// (42 x 33 entries, 133 KB).
#define MK_LDS_TABLE_BYTES (144u * 1024u)
