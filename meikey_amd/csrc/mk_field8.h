// GF(2^255 - 19) in eight 32-bit words: the reference's "8" representation
// (src/service/gpu/kernel.cu:650-848), written here for gfx950.
//
// An element is any 256-bit value V = sum v_i 2^(32 i) and stands for V mod p. Every
// function accepts and returns values below 2^256 (no other bound), and a carry out of
// bit 256 folds back as 38, since 2^256 = 38 (mod p).
//
// Cost on gfx950. A multiplication takes 64 v_mad_u64_u32 products. Each adds into its
// 64-bit column sum with the carry-out counted by a v_addc_co_u32. The columns are then
// combined (the even columns' sums do not overlap, nor do the odd ones', so two 16-word
// carry chains add them and the counts) and folded by 38. That is about 190 VALU
// instructions, all of the one-per-quad-cycle kind. The ten-limb form (mk_field.h) needs
// about 150 and has carry-free additions.
#pragma once

#include "mk_common.h"

struct fe {
    u32 v[8];
};

MK_DEV fe fe_small(u32 x) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 8; i++) r.v[i] = 0;
    r.v[0] = x;
    return r;
}

// a + b: the carry out of bit 256 adds 38; a second carry leaves the value below 38, so
// the last 38 cannot carry.
MK_DEV fe fe_add(const fe& a, const fe& b) {
    fe r;
    u32 c = 0;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)a.v[i] + b.v[i] + c;
        r.v[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    c *= 38u;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)r.v[i] + c;
        r.v[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    r.v[0] += 38u * c;
    return r;
}

// a - b: a borrow out of bit 256 (the result is then a - b + 2^256) subtracts 38; a
// second borrow leaves the value at 2^256 - 38 or above, so the last 38 cannot borrow.
MK_DEV fe fe_sub(const fe& a, const fe& b) {
    fe r;
    u32 br = 0;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)a.v[i] - b.v[i] - br;
        r.v[i] = (u32)t;
        br = (u32)(t >> 63);
    }
    br *= 38u;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)r.v[i] - br;
        r.v[i] = (u32)t;
        br = (u32)(t >> 63);
    }
    r.v[0] -= 38u * br;
    return r;
}

MK_DEV fe fe_neg(const fe& a) { return fe_sub(fe_small(0), a); }

// Already in range: every value below 2^256 is an element.
MK_DEV fe fe_carry(const fe& a) { return a; }

// acc += a * b, the carry out of the 64-bit sum counted into hi. On gfx950 one
// v_mad_u64_u32, whose carry-out (a lane mask) feeds one v_addc_co_u32.
MK_DEV void fe8_mac(u64& acc, u32& hi, u32 a, u32 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    u64 cc;
    asm("v_mad_u64_u32 %0, %1, %3, %4, %0\n\tv_addc_co_u32 %2, %1, 0, %2, %1"
        : "+v"(acc), "=&s"(cc), "+v"(hi)
        : "v"(a), "v"(b));
#else
    const u64 p = (u64)a * b;
    acc += p;
    hi += acc < p ? 1u : 0u;
#endif
}

// The 512-bit product folded to 256 bits: lo + 38 * hi, then the few bits above 2^256 once
// more by 38.
MK_DEV fe fe_fold512(const u32 w[16]) {
    // t_i = w_i + 38 w_(8+i) < 2^38; the words of sum t_i 2^(32 i) with one carry chain
    u32 lo[8], hi[8];
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)w[8 + i] * 38u + w[i];
        lo[i] = (u32)t;
        hi[i] = (u32)(t >> 32);
    }
    fe r;
    u32 c = 0;
    r.v[0] = lo[0];
    MK_UNROLL
    for (int i = 1; i < 8; i++) {
        const u64 t = (u64)lo[i] + hi[i - 1] + c;
        r.v[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    // the part above 2^256 (below 2^6 + 1), times 38
    u32 top = (hi[7] + c) * 38u;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)r.v[i] + top;
        r.v[i] = (u32)t;
        top = (u32)(t >> 32);
    }
    r.v[0] += 38u * top;  // a last carry leaves the value below 2^12: no further carry
    return r;
}

// f * g, any values below 2^256.
MK_DEV fe fe_mul(const fe& f, const fe& g) {
    // column sums: acc[k] = sum_{i+j=k} f_i g_j mod 2^64, cnt[k] = its carries out
    u64 acc[15];
    u32 cnt[15];
    MK_UNROLL
    for (int k = 0; k < 15; k++) {
        acc[k] = 0;
        cnt[k] = 0;
    }
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        MK_UNROLL
        for (int j = 0; j < 8; j++) fe8_mac(acc[i + j], cnt[i + j], f.v[i], g.v[j]);
    }
    // sum_k acc[k] 2^(32k) + cnt[k] 2^(32k + 64): the even columns' acc do not overlap,
    // nor the odd ones', nor the counts
    u32 e[16], o[16], h[16];
    MK_UNROLL
    for (int i = 0; i < 16; i++) e[i] = o[i] = h[i] = 0;
    MK_UNROLL
    for (int k = 0; k < 15; k++) {
        u32* d = (k & 1) ? o : e;
        d[k] = (u32)acc[k];
        d[k + 1] = (u32)(acc[k] >> 32);
        if (k + 2 < 16) h[k + 2] = cnt[k];
    }
    u32 w[16];
    u32 c = 0;
    MK_UNROLL
    for (int i = 0; i < 16; i++) {
        const u64 t = (u64)e[i] + o[i] + c;
        w[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    c = 0;
    MK_UNROLL
    for (int i = 2; i < 16; i++) {
        const u64 t = (u64)w[i] + h[i] + c;
        w[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    return fe_fold512(w);
}

MK_DEV fe fe_sq(const fe& f) { return fe_mul(f, f); }

MK_DEV fe fe_pow2k(fe a, int k) {
    MK_NO_UNROLL
    for (int i = 0; i < k; i++) a = fe_sq(a);
    return a;
}

// a^(p-2): 254 squarings and 11 multiplications (the addition chain of the reference's
// src/service/derive/field.rs:108-124); 0 maps to 0.
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

// Canonical value in [0, p) as eight little-endian words, of any value below 2^256.
MK_DEV void fe_carried_to_words(const fe& h, u32 w[8]) {
    // below 2^255 + 19: bit 255 folds as 19
    u32 c = (h.v[7] >> 31) * 19u;
    u32 v[8];
    MK_UNROLL
    for (int i = 0; i < 8; i++) v[i] = h.v[i];
    v[7] &= 0x7fffffffu;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)v[i] + c;
        v[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    // v >= p exactly when v + 19 reaches 2^255; then the result is v + 19 - 2^255
    u32 s[8];
    c = 19u;
    MK_UNROLL
    for (int i = 0; i < 8; i++) {
        const u64 t = (u64)v[i] + c;
        s[i] = (u32)t;
        c = (u32)(t >> 32);
    }
    const u32 q = s[7] >> 31;
    s[7] &= 0x7fffffffu;
    MK_UNROLL
    for (int i = 0; i < 8; i++) w[i] = q ? s[i] : v[i];
}

MK_DEV void fe_to_words(const fe& a, u32 w[8]) { fe_carried_to_words(a, w); }

// Any 256-bit little-endian value is an element as it is.
MK_DEV fe fe_from_words(const u32 w[8]) {
    fe r;
    MK_UNROLL
    for (int i = 0; i < 8; i++) r.v[i] = w[i];
    return r;
}

MK_DEV bool fe_is_odd(const fe& a) {
    u32 w[8];
    fe_to_words(a, w);
    return (w[0] & 1u) != 0;
}
