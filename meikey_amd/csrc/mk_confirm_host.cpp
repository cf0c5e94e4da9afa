// Host re-derivation of emitted keys, compiled by the host C++ compiler (the arithmetic
// headers then take their plain C++ form) and linked into the library. mk_search calls
// it before handing a key out, as the reference confirms every GPU hit on the CPU
// (src/service/gpu.rs:973-996).
#include <stdio.h>

#include "mk_confirm.h"

#define MK_MATCHER_FAIL(st, ...) (st)
#include "mk_matcher_host.h"

extern "C" __attribute__((visibility("hidden"))) int mk_confirm_emitted(
    const uint8_t key[32], int source, uint64_t index, const uint8_t seed[32], const uint8_t pk[32],
    const uint8_t* needle, uint32_t len, uint32_t fold, uint32_t mode) {
    matcher_host m;
    if (build_matcher(needle, len, fold, mode, m) != 0) return -1;
    mk_matcher mt;
    mt.mode = m.mode;
    mt.len = m.len;
    mt.mask = m.mask;
    mt.dfa = m.dfa.empty() ? nullptr : m.dfa.data();
    u32 kw[8];
    for (int i = 0; i < 8; i++)
        kw[i] = (u32)key[4 * i] | (u32)key[4 * i + 1] << 8 | (u32)key[4 * i + 2] << 16 | (u32)key[4 * i + 3] << 24;
    return mk_confirm_key(kw, (u32)source, index, seed, pk, mt) ? 1 : 0;
}
