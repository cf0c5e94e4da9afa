// One lane's share of a search tile: K consecutive key indices, their y coordinates
// from a single field inversion, handed one by one (last to first) to a callback.
// The search kernel and the CPU test build both run this exact code.
#pragma once

#include "mk_hash.h"
#include "mk_point.h"

// Per-lane spill area of the batched inversion: two field elements per key (the key's
// numerator times the prefix product before it, and its denominator).
// Element `i` of field slot `slot` lives at base[(slot * MK_FE_LIMBS + i) * stride]; with
// stride = number of lanes and base offset by the lane, the words of one slot are
// contiguous across a wave.
struct mk_spill {
    u32* base;
    u64 stride;
    MK_MEMBER void store(u32 slot, const fe& a) const {
        u32* p = base + (u64)slot * MK_FE_LIMBS * stride;
        MK_UNROLL
        for (int i = 0; i < MK_FE_LIMBS; i++) p[(u64)i * stride] = a.v[i];
    }
    MK_MEMBER fe load(u32 slot) const {
        const u32* p = base + (u64)slot * MK_FE_LIMBS * stride;
        fe a;
        MK_UNROLL
        for (int i = 0; i < MK_FE_LIMBS; i++) a.v[i] = p[(u64)i * stride];
        return a;
    }
};

// Keys first .. first + K - 1: seed -> scalar -> (yn, yd) by mul(s, yn, yd); one
// inversion of the product of the yd, by invert(P) (fe_invert, or an inversion shared
// with other lanes that returns this lane's 1/P); then pre_sweep(); then, from the last
// key to the first, y = yn / yd is compressed (x sign bit left 0) and on_key(j,
// pk_words) is called.
// Counter seeds (source 1) hash through the search's SHA-512 mid-state, which mid()
// returns; the search kernel keeps it in LDS and reads it afresh for each key, so its
// 22 registers are not held through the comb multiply.
//
// Montgomery's simultaneous inversion with the numerators folded in: with
// P_j = yd_0 ... yd_j, the forward pass spills Q_j = yn_j * P_(j-1) and yd_j, and after one
// inversion of P_(K-1) the back-sweep takes y_j = Q_j / P_j and steps 1/P_j down to
// 1/P_(j-1) by a multiplication with yd_j. Four multiplications per key, as with the
// textbook form that spills yn, yd and P_(j-1), but two spilled fields instead of three and
// one multiplication between a key's loads and its y.
template <class M, class P, class F, class G, class I>
MK_DEV void mk_lane_tile(u32 K, const u32 key[8], u32 source, u64 first, const mk_spill& sp, M mul, P pre_sweep,
                         F on_key, G mid, I invert) {
    u32 upper[8];
    fe pre = fe_small(1);  // P_(j-1); each yd and yn is a first operand of fe_mul
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
        sp.store(2 * j, fe_mul(yn, pre));
        sp.store(2 * j + 1, yd);
        pre = fe_mul(yd, pre);
    }
    fe inv = invert(pre);  // 1 / P_(K-1)
    pre_sweep();
    // Each key's two spilled values are loaded one key ahead, so their HBM latency
    // overlaps the previous key's multiplications and matching.
    fe n_q = sp.load(2 * (K - 1)), n_yd = sp.load(2 * (K - 1) + 1);
    MK_NO_UNROLL
    for (u32 jj = K; jj-- > 0;) {
        const fe q = n_q, yd = n_yd;
        if (jj > 0) {
            n_q = sp.load(2 * (jj - 1));
            n_yd = sp.load(2 * (jj - 1) + 1);
        }
        u32 pk[8];
        fe_carried_to_words(fe_mul(q, inv), pk);  // y_jj = Q_jj / P_jj
        if (jj > 0) inv = fe_mul(yd, inv);          // 1 / P_(jj-1)
        on_key(jj, pk);
    }
}

template <class M, class P, class F, class G>
MK_DEV void mk_lane_tile(u32 K, const u32 key[8], u32 source, u64 first, const mk_spill& sp, M mul, P pre_sweep,
                         F on_key, G mid) {
    mk_lane_tile(K, key, source, first, sp, mul, pre_sweep, on_key, mid, [](const fe& p) { return fe_invert(p); });
}
