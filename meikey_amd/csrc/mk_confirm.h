// Host re-derivation of an emitted key (the reference's confirm, src/service/gpu.rs:973-996):
// the seed is recomputed from the search key and index, the public key from the seed by
// SHA-512 and a table-free double-and-add, and the pattern is matched with the real x
// sign. A key the device reports is only handed out when all three agree.
#pragma once

#include "mk_hash.h"
#include "mk_match.h"
#include "mk_point.h"

MK_DEV bool mk_confirm_key(const u32 key[8], u32 source, u64 index, const u8 seed[32], const u8 pk[32],
                           const mk_matcher& mt) {
    u32 upper[8], sw[8], s[8], pw[8];
    mk_seed_at(key, source, index, upper, true, sw);
    for (int i = 0; i < 32; i++)
        if ((u8)(sw[i >> 2] >> (8 * (i & 3))) != seed[i]) return false;
    mk_secret_scalar(sw, s);
    mk_pk_by_ladder(s, pw);
    for (int i = 0; i < 32; i++)
        if ((u8)(pw[i >> 2] >> (8 * (i & 3))) != pk[i]) return false;
    return mk_pk_match(pw, mt, false);
}
