// One lane's share of a search tile: K consecutive key indices, their y coordinates
// from a single field inversion, handed one by one (last to first) to a callback.
// The search kernel and the CPU test build both run this exact code.
#pragma once

#include "mk_hash.h"
#include "mk_point.h"

// Per-lane spill area for the K numerators, denominators and prefix products.
// Element `i` of field slot `slot` lives at base[(slot * 10 + i) * stride]; with
// stride = number of lanes and base offset by the lane, the words of one slot are
// contiguous across a wave.
struct mk_spill {
    u32* base;
    u64 stride;
    MK_MEMBER void store(u32 slot, const fe& a) const {
        u32* p = base + (u64)slot * 10 * stride;
        MK_UNROLL
        for (int i = 0; i < 10; i++) p[(u64)i * stride] = a.v[i];
    }
    MK_MEMBER fe load(u32 slot) const {
        const u32* p = base + (u64)slot * 10 * stride;
        fe a;
        MK_UNROLL
        for (int i = 0; i < 10; i++) a.v[i] = p[(u64)i * stride];
        return a;
    }
};

// Keys first .. first + K - 1: seed -> scalar -> (yn, yd) by mul(s, yn, yd); one
// inversion of the product of the yd; then pre_sweep(); then, from the last key to the
// first, y = yn / yd is compressed (x sign bit left 0) and on_key(j, pk_words) is called.
// Counter seeds (source 1) hash through the search's SHA-512 mid-state, which mid()
// returns; the search kernel keeps it in LDS and reads it afresh for each key, so its
// 22 registers are not held through the comb multiply.
template <class M, class P, class F, class G>
MK_DEV void mk_lane_tile(u32 K, const u32 key[8], u32 source, u64 first, const mk_spill& sp, M mul, P pre_sweep,
                         F on_key, G mid) {
    u32 upper[8];
    fe pre = fe_small(1);  // prefix product of the yd (each yd is a first operand of fe_mul)
    const u64 tail = (u64)key[7] << 32 | key[6];
    MK_NO_UNROLL
    for (u32 j = 0; j < K; j++) {
        u32 s[8];
        if (source == 1) {
            const u64 t = tail + first + j;  // bytes 24..31 of the counter seed (LE)
            mk_secret_scalar_mid(mid(), mk_be_word((u32)t, (u32)(t >> 32)), s);
        } else {
            u32 seed[8];
            mk_seed_at(key, source, first + j, upper, j == 0, seed);
            mk_secret_scalar(seed, s);
        }
        fe yn, yd;
        mul(s, yn, yd);
        pre = fe_mul(yd, pre);
        sp.store(3 * j, yn);
        sp.store(3 * j + 1, yd);
        sp.store(3 * j + 2, pre);
    }
    fe inv = fe_invert(pre);
    pre_sweep();
    // 1/yd_j = inv * pre_(j-1) with inv = 1/pre_j; then inv <- inv * yd_j = 1/pre_(j-1).
    // Each key's three spilled values are loaded one key ahead, so their HBM latency
    // overlaps the previous key's multiplications and matching.
    fe n_yn = sp.load(3 * (K - 1)), n_yd = inv, n_pre = inv;
    if (K > 1) {
        n_yd = sp.load(3 * (K - 1) + 1);
        n_pre = sp.load(3 * (K - 2) + 2);
    }
    MK_NO_UNROLL
    for (u32 jj = K; jj-- > 0;) {
        const fe yn = n_yn, yd = n_yd, pre_prev = n_pre;
        if (jj > 0) {
            n_yn = sp.load(3 * (jj - 1));
            if (jj > 1) {
                n_yd = sp.load(3 * (jj - 1) + 1);
                n_pre = sp.load(3 * (jj - 2) + 2);
            }
        }
        fe zinv = inv;
        if (jj > 0) {
            zinv = fe_mul(inv, pre_prev);
            inv = fe_mul(yd, inv);
        }
        u32 pk[8];
        fe_carried_to_words(fe_mul(yn, zinv), pk);
        on_key(jj, pk);
    }
}
