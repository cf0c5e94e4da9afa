// Host-side matcher construction (shared by the C ABI and the CPU test build): the
// per-search bitap table for substring / suffix patterns and the validated DFA table
// for regular expressions (the reference's Needle, src/service/gpu.rs:617-660).
#pragma once

#include <stdint.h>

#include <vector>

#include "mk_common.h"

static const char mk_b64_alphabet[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

// MK_MATCHER_FAIL(status, fmt, ...) reports an error and returns status; the includer defines it.
// Host-side form of one search's matcher: descriptor fields plus the bitap table or
// the DFA table (src/service/gpu.rs:617-660 Needle).
struct matcher_host {
    u32 mode = MK_MODE_CONTAINS, len = 0, fold = 0;
    u64 mask[64] = {};
    std::vector<u16> dfa;  // states * 65
};

static int build_matcher(const uint8_t* needle, uint32_t len, uint32_t fold, uint32_t mode, matcher_host& m) {
    m = matcher_host();
    m.mode = mode;
    m.fold = fold ? 1u : 0u;
    if (mode == MK_MODE_REGEX) {
        if (needle == nullptr || len < 2 || len > (1u << 14))
            return MK_MATCHER_FAIL(6, "regex DFA needs 2..16384 states, got %u", len);
        m.len = len;
        m.dfa.resize((size_t)len * 65);
        for (size_t i = 0; i < m.dfa.size(); i++) {
            const u16 v = (u16)(needle[2 * i] | (u16)needle[2 * i + 1] << 8);
            const size_t col = i % 65;
            if (col < 64 ? v >= len : v > 1)
                return MK_MATCHER_FAIL(6, "DFA entry %zu (state %zu, column %zu) = %u out of range", i, i / 65, col,
                            v);
            m.dfa[i] = v;
        }
        return 0;
    }
    if (mode != MK_MODE_CONTAINS && mode != MK_MODE_SUFFIX) return MK_MATCHER_FAIL(6, "unknown mode %u", mode);
    if (len > MK_TAIL_B64 - MK_SEARCH_FROM) return MK_MATCHER_FAIL(6, "needle of %u characters (at most 43)", len);
    if (len > 0 && needle == nullptr) return MK_MATCHER_FAIL(6, "null needle");
    m.len = len;
    // mask[v] bit k: the character of value v equals needle[k] (both folded to lower
    // case when fold, matching b64_char(v, fold) == needle[k] of kernel.cu:1703-1715)
    for (u32 v = 0; v < 64; v++) {
        u8 c = (u8)mk_b64_alphabet[v];
        if (m.fold && c >= 'A' && c <= 'Z') c = (u8)(c + 32);
        for (u32 k = 0; k < len; k++) {
            u8 n = needle[k];
            if (m.fold && n >= 'A' && n <= 'Z') n = (u8)(n + 32);
            if (c == n) m.mask[v] |= 1ull << k;
        }
    }
    return 0;
}

