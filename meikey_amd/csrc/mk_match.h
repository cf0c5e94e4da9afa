// Base64 of the key-dependent tail and the pattern matchers.
//
// The public-key blob is string("ssh-ed25519") || string(pk), 51 bytes, 68 base64
// characters. Blob bytes 18..50 (0x20 then the 32 key bytes) give the 44 characters
// b64[24..68]; tail position p is b64[24 + p]. Patterns match the 43-character hay
// b64[25..68], i.e. tail positions 1..43 (reference src/model/pattern.rs:15-19,
// 192-220; src/service/gpu/kernel.cu:1703-1728).
//
// Matching works on 6-bit values, never on characters:
//  * substring / suffix: a shift-and (bitap) automaton. For each value v the host
//    precomputes mask[v] = the set of needle positions k whose character equals the
//    character of v (both cases of a letter when case-insensitive). One table lookup
//    and three 64-bit operations per tail position, whatever the needle length.
//  * regex: the DFA table of the reference's format (states x 65 x u16; state 0 =
//    ACCEPT and absorbing, 1 = START, column 64 = "accepts at end of input";
//    src/model/pattern/dfa.rs:14-20, 187-200).
// The x sign lives in bit 1 of the value at tail position 42 (kernel.cu:1730-1731).
// With sign_unknown the matcher accepts when either sign would match: the bitap
// step at position 42 uses mask[v] | mask[v | 2], and the DFA forks a second state
// there (kernel.cu:1819-1822, 1866-1871).
#pragma once

#include "mk_common.h"

// Byte b of the 33-byte tail (0x20 followed by pk[0..31]); b is a compile-time
// constant in the unrolled loops below.
MK_DEV u32 mk_tail_byte(const u32 pk[8], int b) {
    if (b == 0) return 0x20u;
    const int i = b - 1;
    return (pk[i >> 2] >> (8 * (i & 3))) & 0xffu;
}

// Value (0..63) of tail position pos.
MK_DEV u32 mk_tail_value(const u32 pk[8], int pos) {
    const int g = pos >> 2, k = pos & 3;
    const u32 n = mk_tail_byte(pk, 3 * g) << 16 | mk_tail_byte(pk, 3 * g + 1) << 8 | mk_tail_byte(pk, 3 * g + 2);
    return (n >> (18 - 6 * k)) & 63u;
}

// Shift-and over tail positions From..43; returns the final state D and, in seen, the OR
// of all states. Bit k of D at position p is set when needle[0..k] ends at p. S is the
// state word: u32 for needles of at most 32 characters (the low half of each mask entry,
// one 32-bit LDS read and 32-bit shift-and per position), u64 for longer ones.
template <int From, class S>
MK_DEV S mk_bitap_run(const u32 pk[8], const u64* mask, bool sign_unknown, S& seen) {
    S D = 0;
    seen = 0;
    MK_UNROLL
    for (int pos = From; pos < MK_TAIL_B64; pos++) {
        const u32 v = mk_tail_value(pk, pos);
        S m = (S)mask[v];
        if (pos == MK_SIGN_CHAR && sign_unknown) m |= (S)mask[v | 2u];
        D = ((D << 1) | (S)1) & m;
        seen |= D;
    }
    return D;
}

template <class S>
MK_DEV bool mk_bitap_match_w(const u32 pk[8], const u64* mask, u32 len, bool suffix, bool sign_unknown) {
    const S last = (S)1 << (len - 1);
    S seen = 0;
    if (suffix) {
        const S D = len <= 8    ? mk_bitap_run<MK_TAIL_B64 - 8>(pk, mask, sign_unknown, seen)
                    : len <= 16 ? mk_bitap_run<MK_TAIL_B64 - 16>(pk, mask, sign_unknown, seen)
                                : mk_bitap_run<MK_SEARCH_FROM>(pk, mask, sign_unknown, seen);
        return (D & last) != 0;
    }
    mk_bitap_run<MK_SEARCH_FROM>(pk, mask, sign_unknown, seen);
    return (seen & last) != 0;
}

// Substring (suffix = false) or suffix match by shift-and over tail positions 1..43.
// mask: 64 entries; len in 1..43. A suffix of len characters can only start at
// position 44 - len, so the automaton starts at 36 for len <= 8 and at 28 for len <= 16
// (the final state is the same as from position 1).
MK_DEV bool mk_bitap_match(const u32 pk[8], const u64* mask, u32 len, bool suffix, bool sign_unknown) {
    return len <= 32 ? mk_bitap_match_w<u32>(pk, mask, len, suffix, sign_unknown)
                     : mk_bitap_match_w<u64>(pk, mask, len, suffix, sign_unknown);
}

MK_DEV u32 mk_dfa_next(const u16* dfa, u32 s, u32 col) { return dfa[s * 65u + col]; }

MK_DEV bool mk_dfa_done(const u16* dfa, u32 s) { return s == MK_DFA_ACCEPT || mk_dfa_next(dfa, s, 64) != 0; }

MK_DEV bool mk_dfa_match(const u32 pk[8], const u16* dfa, bool sign_unknown) {
    u32 s = MK_DFA_START, alt = MK_DFA_START;
    MK_UNROLL
    for (int pos = MK_SEARCH_FROM; pos < MK_TAIL_B64; pos++) {
        const u32 v = mk_tail_value(pk, pos);
        if (pos == MK_SIGN_CHAR) {
            alt = mk_dfa_next(dfa, s, sign_unknown ? (v | 2u) : v);
        } else if (pos > MK_SIGN_CHAR) {
            alt = mk_dfa_next(dfa, alt, v);
        }
        s = mk_dfa_next(dfa, s, v);
    }
    return mk_dfa_done(dfa, s) || mk_dfa_done(dfa, alt);
}

// One search's matcher. len = needle length (substring/suffix) or DFA state count.
struct mk_matcher {
    u32 mode, len;
    const u64* mask;
    const u16* dfa;
};

MK_DEV bool mk_pk_match(const u32 pk[8], const mk_matcher& m, bool sign_unknown) {
    if (m.mode == MK_MODE_REGEX) return mk_dfa_match(pk, m.dfa, sign_unknown);
    if (m.len == 0 || m.len > MK_MAX_NEEDLE + 1) return false;
    return mk_bitap_match(pk, m.mask, m.len, m.mode == MK_MODE_SUFFIX, sign_unknown);
}
