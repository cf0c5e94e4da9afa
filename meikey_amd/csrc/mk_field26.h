// GF(2^255 - 19) in ten uniform 26-bit limbs with Karatsuba products: the reference's
// "10k" / "10ku" representation (src/service/gpu/kernel.cu:1102-1316), written for gfx950.
//
// An element is V = sum v_i 2^(26 i) with any limbs below 2^29; V may exceed p and 2^255
// (up to about 2^263). 2^260 = 32 * 2^255 = 608 (mod p).
//
// Multiplication. With a = a_L + 2^130 a_H and b = b_L + 2^130 b_H (five limbs each) and
// D = (a_L - a_H)(b_H - b_L) in signed limbs,
//   a * b = P0 + 2^130 (P0 + P2 + D) + 2^260 P2,   P0 = a_L b_L,  P2 = a_H b_H,
// so a product takes 75 limb products instead of 100 and the middle term needs additions
// only: D's 25 products are v_mad_i64_i32 (signed 32 x 32 + 64-bit addend, the same issue
// rate as v_mad_u64_u32), accumulated onto P0 + P2 of the column. The 19 columns of
// weight 2^(26 t) then fold: column t >= 10 adds 608 times its low 32 bits to column
// t - 10 and 608 * 2^6 times its high 32 bits to column t - 9 (2^32 = 2^6 * 2^26), two
// multiply-adds and no carry pass over the high columns. A carry pass of 12 steps over
// the ten low columns follows. About 157 VALU instructions, against about 151 for the
// ten 25.5-bit limbs of mk_field.h.
//
// Limb bounds ("carried" = what fe_mul, fe_sq and fe_carry return):
//   carried:   limb 1 < 2^26 + 2^22, limb 6 < 2^26 + 2^13, the others < 2^26
//   fe_add of two carried values: < 2^27 + 2^23
//   fe_sub(a, b) = a + 64p - b with b carried (64p's limbs are 2^27 - 1216 and 2^27 - 2):
//              < a's bound + 2^27
// fe_mul and fe_sq take any limbs below 2^29 on either side: every column sum, with the
// middle term's signed products, then stays below 2^63.
#pragma once

#include "mk_common.h"

struct fe {
    u32 v[10];
};

#define MK_K26_MASK 0x3ffffffu
#define MK_K26_MASK21 0x1fffffu

MK_DEV fe fe_small(u32 x) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = 0;
    r.v[0] = x;
    return r;
}

MK_DEV fe fe_add(const fe& a, const fe& b) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = a.v[i] + b.v[i];
    return r;
}

// 64p = 2^261 - 1216 as limbs: limb 0 = 2^27 - 1216, the others 2^27 - 2.
MK_DEV u32 fe_64p(int i) { return i == 0 ? (1u << 27) - 1216u : (1u << 27) - 2u; }

// a - b + 64p; b must be carried (its limbs at most 64p's).
MK_DEV fe fe_sub(const fe& a, const fe& b) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = a.v[i] + fe_64p(i) - b.v[i];
    return r;
}

MK_DEV fe fe_neg(const fe& a) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = fe_64p(i) - a.v[i];
    return r;
}

// Carry pass over ten 64-bit column sums (each below 2^63) back to carried limbs. Two
// interleaved chains, 0 -> 1 -> ... -> 5 and 5 -> 6 -> ... -> 9 -> 0 -> 1. The carry
// out of column 9 is below 2^38; its low 32 bits fold into column 0 (times 608) and its
// high bits into column 1 (times 608 * 2^6), so the fold needs no 64-bit product.
MK_DEV fe fe_carry64(u64 h[10]) {
    u64 c;
    c = h[0] >> 26; h[1] += c; h[0] &= MK_K26_MASK;
    c = h[5] >> 26; h[6] += c; h[5] &= MK_K26_MASK;
    c = h[1] >> 26; h[2] += c; h[1] &= MK_K26_MASK;
    c = h[6] >> 26; h[7] += c; h[6] &= MK_K26_MASK;
    c = h[2] >> 26; h[3] += c; h[2] &= MK_K26_MASK;
    c = h[7] >> 26; h[8] += c; h[7] &= MK_K26_MASK;
    c = h[3] >> 26; h[4] += c; h[3] &= MK_K26_MASK;
    c = h[8] >> 26; h[9] += c; h[8] &= MK_K26_MASK;
    c = h[4] >> 26; h[5] += c; h[4] &= MK_K26_MASK;
    c = h[9] >> 26; h[9] &= MK_K26_MASK;
    h[0] += (u64)(u32)c * 608u;
    h[1] += (u64)(u32)(c >> 32) * (608u << 6);
    c = h[5] >> 26; h[6] += c; h[5] &= MK_K26_MASK;
    c = h[0] >> 26; h[1] += c; h[0] &= MK_K26_MASK;
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = (u32)h[i];
    return r;
}

// Weak reduction of 32-bit limbs (any limb < 2^31) to carried form.
MK_DEV fe fe_carry(const fe& a) {
    u32 v[10];
    MK_UNROLL
    for (int i = 0; i < 10; i++) v[i] = a.v[i];
    MK_UNROLL
    for (int i = 0; i < 9; i++) {
        v[i + 1] += v[i] >> 26;
        v[i] &= MK_K26_MASK;
    }
    const u32 top = v[9] >> 26;  // below 2^6
    v[9] &= MK_K26_MASK;
    v[0] += 608u * top;
    v[1] += v[0] >> 26;
    v[0] &= MK_K26_MASK;
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = v[i];
    return r;
}

// Column t of a 19-column product lands in h[t] (t < 10) or folds into h[t - 10] and
// h[t - 9] (t >= 10).
MK_DEV void fe_fold_col(u64 h[10], int t, u64 col) {
    if (t < 10) {
        h[t] += col;
    } else {
        h[t - 10] += (u64)(u32)col * 608u;
        h[t - 9] += (u64)(u32)(col >> 32) * (608u << 6);
    }
}

// Columns k of P0, P2 and the middle term M = P0 + P2 + D in order of k: P0_k lands in
// column k, M_k in column k + 5 and P2_k in column k + 10. Columns 10-13 take both a P2
// and an M term, so P2_0 .. P2_3 wait for M_5 .. M_8 and each high column folds once.
MK_DEV void fe_place_cols(u64 h[10], u64 hold[4], int k, u64 p0, u64 p2, u64 m) {
    h[k] += p0;
    if (k < 4)
        hold[k] = p2;
    else
        fe_fold_col(h, k + 10, p2);
    if (k < 5)
        h[k + 5] += m;
    else
        fe_fold_col(h, k + 5, hold[k - 5] + m);
}

// f * g by one level of Karatsuba on the five-limb halves (see the top of this file).
MK_DEV fe fe_mul(const fe& f, const fe& g) {
    i32 df[5], dg[5];
    MK_UNROLL
    for (int i = 0; i < 5; i++) {
        df[i] = (i32)(f.v[i] - f.v[i + 5]);
        dg[i] = (i32)(g.v[i + 5] - g.v[i]);
    }
    u64 h[10], hold[4];
    MK_UNROLL
    for (int t = 0; t < 10; t++) h[t] = 0;
    MK_UNROLL
    for (int k = 0; k < 9; k++) {
        u64 p0 = 0, p2 = 0;
        MK_UNROLL
        for (int i = 0; i < 5; i++) {
            const int j = k - i;
            if (j < 0 || j > 4) continue;
            p0 += (u64)f.v[i] * g.v[j];
            p2 += (u64)f.v[5 + i] * g.v[5 + j];
        }
        u64 m = p0 + p2;
        MK_UNROLL
        for (int i = 0; i < 5; i++) {
            const int j = k - i;
            if (j < 0 || j > 4) continue;
            m += (u64)((int64_t)df[i] * dg[j]);
        }
        fe_place_cols(h, hold, k, p0, p2, m);
    }
    return fe_carry64(h);
}

// f^2: the same split with five-limb squares (15 products each, cross terms doubled) and
// D = (f_L - f_H)(f_H - f_L), 45 limb products in all.
MK_DEV fe fe_sq(const fe& f) {
    i32 df[5], nd[5];
    u32 f2[10];
    MK_UNROLL
    for (int i = 0; i < 5; i++) {
        df[i] = (i32)(f.v[i] - f.v[i + 5]);
        nd[i] = -df[i];
    }
    MK_UNROLL
    for (int i = 0; i < 10; i++) f2[i] = 2u * f.v[i];
    u64 h[10], hold[4];
    MK_UNROLL
    for (int t = 0; t < 10; t++) h[t] = 0;
    MK_UNROLL
    for (int k = 0; k < 9; k++) {
        u64 p0 = 0, p2 = 0;
        MK_UNROLL
        for (int i = 0; i < 5; i++) {
            const int j = k - i;
            if (j < i || j > 4) continue;
            p0 += (u64)(i == j ? f.v[i] : f2[i]) * f.v[j];
            p2 += (u64)(i == j ? f.v[5 + i] : f2[5 + i]) * f.v[5 + j];
        }
        u64 m = p0 + p2;
        MK_UNROLL
        for (int i = 0; i < 5; i++) {
            const int j = k - i;
            if (j < i || j > 4) continue;
            m += (u64)((int64_t)(i == j ? df[i] : 2 * df[i]) * nd[j]);
        }
        fe_place_cols(h, hold, k, p0, p2, m);
    }
    return fe_carry64(h);
}

MK_DEV fe fe_pow2k(fe a, int k) {
    MK_NO_UNROLL
    for (int i = 0; i < k; i++) a = fe_sq(a);
    return a;
}

// a^(p-2): 254 squarings and 11 multiplications (the addition chain of the reference's
// src/service/derive/field.rs:108-124). 0 maps to 0.
MK_DEV fe fe_invert(const fe& a) {
    fe a2 = fe_sq(a);
    fe a9 = fe_mul(fe_sq(fe_sq(a2)), a);
    fe a11 = fe_mul(a9, a2);
    fe a_5_0 = fe_mul(fe_sq(a11), a9);
    fe a_10_0 = fe_mul(fe_pow2k(a_5_0, 5), a_5_0);
    fe a_20_0 = fe_mul(fe_pow2k(a_10_0, 10), a_10_0);
    fe a_40_0 = fe_mul(fe_pow2k(a_20_0, 20), a_20_0);
    fe a_50_0 = fe_mul(fe_pow2k(a_40_0, 10), a_10_0);
    fe a_100_0 = fe_mul(fe_pow2k(a_50_0, 50), a_50_0);
    fe a_200_0 = fe_mul(fe_pow2k(a_100_0, 100), a_100_0);
    fe a_250_0 = fe_mul(fe_pow2k(a_200_0, 50), a_50_0);
    return fe_mul(fe_pow2k(a_250_0, 5), a11);
}

// Canonical value in [0, p) as eight little-endian 32-bit words, of any element whose
// limbs are below 2^30 (carried values included).
MK_DEV void fe_carried_to_words(const fe& h, u32 w[8]) {
    u32 v[10];
    MK_UNROLL
    for (int i = 0; i < 10; i++) v[i] = h.v[i];
    MK_UNROLL
    for (int i = 0; i < 9; i++) {
        v[i + 1] += v[i] >> 26;
        v[i] &= MK_K26_MASK;
    }
    // bits 255 and up (limb 9 from its bit 21, below 2^10) fold back as 19 each; the value
    // is then below 2^255 + 2^15, below 2p
    const u32 top = v[9] >> 21;
    v[9] &= MK_K26_MASK21;
    v[0] += 19u * top;
    // q = floor((V + 19) / 2^255) is 1 exactly when V >= p (the chain of floor divisions
    // is exact for any limbs)
    u32 q = (v[0] + 19u) >> 26;
    MK_UNROLL
    for (int i = 1; i < 9; i++) q = (v[i] + q) >> 26;
    q = (v[9] + q) >> 21;
    v[0] += 19u * q;
    MK_UNROLL
    for (int i = 0; i < 9; i++) {
        v[i + 1] += v[i] >> 26;
        v[i] &= MK_K26_MASK;
    }
    v[9] &= MK_K26_MASK21;  // drops 2^255 q, completing V + 19q - 2^255 q = V - pq
    // limb i starts at bit 26 i
    w[0] = v[0] | v[1] << 26;
    w[1] = v[1] >> 6 | v[2] << 20;
    w[2] = v[2] >> 12 | v[3] << 14;
    w[3] = v[3] >> 18 | v[4] << 8;
    w[4] = v[4] >> 24 | v[5] << 2 | v[6] << 28;
    w[5] = v[6] >> 4 | v[7] << 22;
    w[6] = v[7] >> 10 | v[8] << 16;
    w[7] = v[8] >> 16 | v[9] << 10;
}

// Canonical value of any element (limbs below 2^30) as eight little-endian words.
MK_DEV void fe_to_words(const fe& a, u32 w[8]) { fe_carried_to_words(a, w); }

// Any 256-bit little-endian value to carried limbs (bit 255 stays in limb 9, which holds
// bits 234-255).
MK_DEV fe fe_from_words(const u32 w[8]) {
    fe r;
    r.v[0] = w[0] & MK_K26_MASK;
    r.v[1] = (w[0] >> 26 | w[1] << 6) & MK_K26_MASK;
    r.v[2] = (w[1] >> 20 | w[2] << 12) & MK_K26_MASK;
    r.v[3] = (w[2] >> 14 | w[3] << 18) & MK_K26_MASK;
    r.v[4] = (w[3] >> 8 | w[4] << 24) & MK_K26_MASK;
    r.v[5] = (w[4] >> 2) & MK_K26_MASK;
    r.v[6] = (w[4] >> 28 | w[5] << 4) & MK_K26_MASK;
    r.v[7] = (w[5] >> 22 | w[6] << 10) & MK_K26_MASK;
    r.v[8] = (w[6] >> 16 | w[7] << 16) & MK_K26_MASK;
    r.v[9] = w[7] >> 10;
    return r;
}

MK_DEV bool fe_is_odd(const fe& a) {
    u32 w[8];
    fe_to_words(a, w);
    return (w[0] & 1u) != 0;
}
