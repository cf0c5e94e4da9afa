// GF(2^255 - 19) in ten 25.5-bit limbs (26, 25, 26, 25, ... bits at bit offsets
// 0, 26, 51, 77, 102, 128, 153, 179, 204, 230), laid out for gfx950.
//
// Why this layout on CDNA4: v_mad_u64_u32 (32x32 -> 64 plus a 64-bit addend) issues
// at full VALU rate on gfx950, so each of the 100 limb products is one instruction
// that accumulates straight into a 64-bit column sum; the column sums never exceed
// 2^64, so no carry flags are tracked, and the carry pass uses the 64-bit
// v_lshl_add_u64. The reference's R25 field (src/service/gpu/kernel.cu:399-571) and
// src/service/derive/avx2.rs use the same radix; the code here is written for this
// machine and its bounds are stated per function.
//
// Limb bounds used below ("carried" = what fe_mul / fe_carry return):
//   carried:    even limbs < 2^26 + 2^13, odd limbs < 2^25 + 2^13
//   fe_add of two carried values:  < 2^27 / 2^26
//   fe_sub(a, b) = a + 2p - b with b carried and a carried:  < 3*2^26 / 3*2^25
// fe_mul(f, g): g's limbs must stay below 2^32 / 19 (~2^27.75), so g is carried,
// an fe_add result or an fe_sub result; f may be up to 2^28 per limb (every 64-bit
// column then stays below 2^63.2). fe_sq needs a carried input.
#pragma once

#include "mk_common.h"

#if MK_FIELD == 8
#include "mk_field8.h"
#elif MK_FIELD == 26
#include "mk_field26.h"
#else

struct fe {
    u32 v[10];
};

#define MK_M25 0x1ffffffu
#define MK_M26 0x3ffffffu

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

// 2p in this radix: limb 0 = 2(2^26 - 19), other even limbs 2(2^26 - 1), odd 2(2^25 - 1).
MK_DEV u32 fe_2p(int i) { return i == 0 ? 0x7ffffdau : (i & 1) ? 0x3fffffeu : 0x7fffffeu; }

// a - b + 2p; b must be carried (or any value whose limbs are below 2p's).
MK_DEV fe fe_sub(const fe& a, const fe& b) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = a.v[i] + fe_2p(i) - b.v[i];
    return r;
}

MK_DEV fe fe_neg(const fe& a) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = fe_2p(i) - a.v[i];
    return r;
}

// Carry pass over 64-bit column sums (each < 2^63.2) back to carried 32-bit limbs.
// Two interleaved chains (0->1->2->3->4->5, 4->5->6->7->8->9->0->1) as in ref10.
MK_DEV fe fe_carry64(u64 h[10]) {
    u64 c;
    c = h[0] >> 26; h[1] += c; h[0] &= MK_M26;
    c = h[4] >> 26; h[5] += c; h[4] &= MK_M26;
    c = h[1] >> 25; h[2] += c; h[1] &= MK_M25;
    c = h[5] >> 25; h[6] += c; h[5] &= MK_M25;
    c = h[2] >> 26; h[3] += c; h[2] &= MK_M26;
    c = h[6] >> 26; h[7] += c; h[6] &= MK_M26;
    c = h[3] >> 25; h[4] += c; h[3] &= MK_M25;
    c = h[7] >> 25; h[8] += c; h[7] &= MK_M25;
    c = h[4] >> 26; h[5] += c; h[4] &= MK_M26;
    c = h[8] >> 26; h[9] += c; h[8] &= MK_M26;
    c = h[9] >> 25; h[0] += c * 19; h[9] &= MK_M25;
    c = h[0] >> 26; h[1] += c; h[0] &= MK_M26;
    fe r;
    MK_UNROLL
    for (int i = 0; i < 10; i++) r.v[i] = (u32)h[i];
    return r;
}

// Weak reduction of 32-bit limbs (any limb < 2^31) to carried form.
MK_DEV fe fe_carry(const fe& a) {
    u64 h[10];
    MK_UNROLL
    for (int i = 0; i < 10; i++) h[i] = a.v[i];
    return fe_carry64(h);
}

// f * g: h_k = sum_{i+j=k} c_ij f_i g_j + 19 sum_{i+j=k+10} c_ij f_i g_j, c_ij = 2 when
// both i and j are odd (the half-bit offsets add up to one extra bit).
MK_DEV fe fe_mul(const fe& f, const fe& g) {
    u32 g19[10], f2[10];
    MK_UNROLL
    for (int j = 0; j < 10; j++) g19[j] = 19u * g.v[j];
    MK_UNROLL
    for (int i = 0; i < 10; i++) f2[i] = (i & 1) ? 2u * f.v[i] : f.v[i];
    u64 h[10];
    MK_UNROLL
    for (int k = 0; k < 10; k++) h[k] = 0;
    MK_UNROLL
    for (int i = 0; i < 10; i++) {
        MK_UNROLL
        for (int j = 0; j < 10; j++) {
            const int k = i + j;
            const u32 a = (j & 1) ? f2[i] : f.v[i];
            const u32 b = k >= 10 ? g19[j] : g.v[j];
            h[k >= 10 ? k - 10 : k] += (u64)a * b;
        }
    }
    return fe_carry64(h);
}

// f^2 with the 55 distinct products; f must be carried.
MK_DEV fe fe_sq(const fe& f) {
    u64 h[10];
    MK_UNROLL
    for (int k = 0; k < 10; k++) h[k] = 0;
    MK_UNROLL
    for (int i = 0; i < 10; i++) {
        MK_UNROLL
        for (int j = i; j < 10; j++) {
            const int k = i + j;
            const u32 a = i == j ? f.v[i] : 2u * f.v[i];
            const u32 scale = ((i & 1) && (j & 1) ? 2u : 1u) * (k >= 10 ? 19u : 1u);
            const u32 b = scale * f.v[j];
            h[k >= 10 ? k - 10 : k] += (u64)a * b;
        }
    }
    return fe_carry64(h);
}

MK_DEV fe fe_pow2k(fe a, int k) {
    MK_NO_UNROLL
    for (int i = 0; i < k; i++) a = fe_sq(a);
    return a;
}

// a^(p-2): 254 squarings and 11 multiplications (addition chain of the reference's
// src/service/derive/field.rs:108-124). a must be carried; 0 maps to 0.
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

// Canonical value in [0, p) as eight little-endian 32-bit words, of h whose value is
// below 2p: what fe_mul, fe_sq and fe_carry return (their value is below
// 2^255 + 2^142; limb 1 below 2^25 + 2^18.3, limb 5 below 2^25 + 2^14, the others
// within their radix).
MK_DEV void fe_carried_to_words(const fe& h, u32 w[8]) {
    // q = floor((h + 19) / 2^255) is 1 exactly when h >= p; the chain of floor
    // divisions is exact for any limbs, and h < 2p keeps q at 0 or 1
    u32 q = (h.v[0] + 19u) >> 26;
    q = (h.v[1] + q) >> 25;
    q = (h.v[2] + q) >> 26;
    q = (h.v[3] + q) >> 25;
    q = (h.v[4] + q) >> 26;
    q = (h.v[5] + q) >> 25;
    q = (h.v[6] + q) >> 26;
    q = (h.v[7] + q) >> 25;
    q = (h.v[8] + q) >> 26;
    q = (h.v[9] + q) >> 25;
    u32 v[10];
    MK_UNROLL
    for (int i = 0; i < 10; i++) v[i] = h.v[i];
    v[0] += 19u * q;
    MK_UNROLL
    for (int i = 0; i < 9; i++) {
        const int bits = (i & 1) ? 25 : 26;
        v[i + 1] += v[i] >> bits;
        v[i] &= (i & 1) ? MK_M25 : MK_M26;
    }
    v[9] &= MK_M25;  // drops 2^255, completing h + 19q - 2^255 q = h - pq
    w[0] = v[0] | v[1] << 26;
    w[1] = v[1] >> 6 | v[2] << 19;
    w[2] = v[2] >> 13 | v[3] << 13;
    w[3] = v[3] >> 19 | v[4] << 6;
    w[4] = v[5] | v[6] << 25;
    w[5] = v[6] >> 7 | v[7] << 19;
    w[6] = v[7] >> 13 | v[8] << 12;
    w[7] = v[8] >> 20 | v[9] << 6;
}

// Canonical value of any element (limbs below 2^31) as eight little-endian words.
MK_DEV void fe_to_words(const fe& a, u32 w[8]) { fe_carried_to_words(fe_carry(a), w); }

// Any 256-bit little-endian value (bit 255 folded as 19) to carried limbs.
MK_DEV fe fe_from_words(const u32 w[8]) {
    fe r;
    r.v[0] = w[0] & MK_M26;
    r.v[1] = (w[0] >> 26 | w[1] << 6) & MK_M25;
    r.v[2] = (w[1] >> 19 | w[2] << 13) & MK_M26;
    r.v[3] = (w[2] >> 13 | w[3] << 19) & MK_M25;
    r.v[4] = (w[3] >> 6) & MK_M26;
    r.v[5] = w[4] & MK_M25;
    r.v[6] = (w[4] >> 25 | w[5] << 7) & MK_M26;
    r.v[7] = (w[5] >> 19 | w[6] << 13) & MK_M25;
    r.v[8] = (w[6] >> 12 | w[7] << 20) & MK_M26;
    r.v[9] = (w[7] >> 6) & MK_M25;
    r.v[0] += 19u * (w[7] >> 31);
    return r;
}

MK_DEV bool fe_is_odd(const fe& a) {
    u32 w[8];
    fe_to_words(a, w);
    return (w[0] & 1u) != 0;
}

#endif  // MK_FIELD
