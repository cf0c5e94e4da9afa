// Points of edwards25519 (-x^2 + y^2 = 1 + d x^2 y^2) and the fixed-base comb
// multiply s*B used for every candidate key.
//
// The clamped scalar is s = 8 * (2^251 + m). The comb table of width w has
// rows = ceil(251 / w) rows; entry (i, k) is k * 2^(w*i) * 8B as an affine point in
// "Niels" form, and the top row also adds 2^254 * B, so that
// s*B = sum_i d_i * entry(i, |d_i|) for the signed base-2^w digits d_i of m
// (reference src/service/derive/comb.rs:1-8, 289-339, 414-445).
//
// The reference stores (y+x, y-x, 2dxy), with 2xy in row 0 (src/service/gpu.rs:1455-1480).
// This table stores every field halved: ((y+x)/2, (y-x)/2, dxy), row 0 with xy. The mixed
// addition then takes Z where the textbook formula takes 2Z: its A, B, C and D come out
// halved together and the result is the textbook one divided by 4, the same projective
// point, without the doubling of Z. Row 0 becomes extended coordinates with no multiply
// (X = x, Y = y, Z = 1, T = xy).
#pragma once

#include "mk_field.h"

struct mk_ext {
    fe x, y, z, t;
};

// A table entry's fields: (y+x)/2, (y-x)/2 and dxy (row 0: xy).
struct mk_niels {
    fe ypx, ymx, xyd;
};

// Extended + halved affine Niels point, 7 multiplications (madd for a = -1 with the
// d factor folded into the table). pp, mm and txy are the textbook B, A and C halved,
// and Z stands for D = 2Z halved. Inputs: p carried (or row 0's point from
// mk_row0_extended), q's fields carried or negated.
MK_DEV mk_ext mk_add_niels(const mk_ext& p, const mk_niels& q) {
    const fe pp = fe_mul(fe_add(p.y, p.x), q.ypx);
    const fe mm = fe_mul(fe_sub(p.y, p.x), q.ymx);
    const fe txy = fe_mul(p.t, q.xyd);
    const fe x = fe_sub(pp, mm), y = fe_add(pp, mm);
    const fe z = fe_add(p.z, txy), t = fe_sub(p.z, txy);
    mk_ext r;
    // t (limbs < 2^28) only ever on the first operand; x and z are each the second
    // operand of two products, so their 19x multiples are computed once
    r.t = fe_mul(y, x);
    r.x = fe_mul(t, x);
    r.y = fe_mul(y, z);
    r.z = fe_mul(t, z);
    return r;
}

// The last addition as fractions: y = (pp + mm) / (Z - T*dxy) and
// x = (pp - mm) / (Z + T*dxy), 3 multiplications (the reference's (pp + mm) / (2Z - T*2dxy),
// kernel.cu:1613-1627, with numerator and denominator halved by the table).
// yd is left uncarried: it is mk_add_niels's t (limbs below 2^27 + 2^26), so callers take
// it only as the first operand of fe_mul. xd is carried; yn and xn are sums and
// differences of products (second-operand bounds).
MK_DEV void mk_add_niels_frac(const mk_ext& p, const mk_niels& q, fe& yn, fe& yd, fe& xn, fe& xd) {
    const fe pp = fe_mul(fe_add(p.y, p.x), q.ypx);
    const fe mm = fe_mul(fe_sub(p.y, p.x), q.ymx);
    const fe txy = fe_mul(p.t, q.xyd);
    yn = fe_add(pp, mm);
    yd = fe_sub(p.z, txy);
    xn = fe_sub(pp, mm);
    xd = fe_carry(fe_add(p.z, txy));
}

// Row 0's entry ((y+x)/2, (y-x)/2, xy) to extended coordinates with no multiplication and
// no carry pass: X = ypx + 2p - ymx = x, Y = ypx + ymx = y, Z = 1, T = xy. The point only
// ever enters mk_add_niels as p, where Y + X = 2 ypx + 2p (limbs < 2^28) and
// Y - X = Y + 2p - X = 2 ymx (exact limb by limb: no limb borrows, although X's limbs
// exceed 2p's) are both first operands of fe_mul, as that addition requires.
MK_DEV mk_ext mk_row0_extended(const mk_niels& n) {
    mk_ext p;
    p.x = fe_sub(n.ypx, n.ymx);
    p.y = fe_add(n.ypx, n.ymx);
    p.z = fe_small(1);
    p.t = n.xyd;
    return p;
}

// Signed digits of m, produced one per row from low to high. m is bits 3..253 of the
// clamped scalar; each digit is the next w bits plus the carry, mapped to
// [-2^(w-1), 2^(w-1)) with a carry out, except the top digit, which keeps the carry
// and lies in [0, 2^(w-1)] (comb.rs:410-445).
struct mk_digits {
    u64 m[4];
    u32 w, rows, i;
    i32 carry;
};

MK_DEV mk_digits mk_digits_of(const u32 s[8], u32 w) {
    mk_digits g;
    u64 x[4];
    MK_UNROLL
    for (int i = 0; i < 4; i++) x[i] = (u64)s[2 * i + 1] << 32 | s[2 * i];
    g.m[0] = x[0] >> 3 | x[1] << 61;
    g.m[1] = x[1] >> 3 | x[2] << 61;
    g.m[2] = x[2] >> 3 | x[3] << 61;
    g.m[3] = (x[3] >> 3) & ((1ull << (MK_M_BITS - 192)) - 1);
    g.w = w;
    g.rows = mk_rows(w);
    g.i = 0;
    g.carry = 0;
    return g;
}

MK_DEV u64 mk_digit_word(const mk_digits& g, u32 q) {
    return q == 0 ? g.m[0] : q == 1 ? g.m[1] : q == 2 ? g.m[2] : q == 3 ? g.m[3] : 0ull;
}

MK_DEV i32 mk_next_digit(mk_digits& g) {
    const u32 off = g.i * g.w, q = off >> 6, r = off & 63;
    u64 bits = mk_digit_word(g, q) >> r;
    if (r != 0) bits |= mk_digit_word(g, q + 1) << (64 - r);
    const i32 v = (i32)(bits & ((1ull << g.w) - 1)) + g.carry;
    i32 d = v;
    if (g.i + 1 < g.rows) {
        g.carry = (v + (1 << (g.w - 1))) >> g.w;
        d = v - (g.carry << g.w);
    }
    g.i++;
    return d;
}

// Raw entry words as loaded from the table (30 limbs + 2 pad words).
struct mk_entry {
    u32 w[MK_ENTRY_WORDS];
};

// 16-byte chunks of one entry, as staged in LDS.
#define MK_STAGE_CHUNKS (MK_ENTRY_WORDS / 4)

#if !MK_PACKED
// Word of limb i of field f (0 = y+x, 1 = y-x, 2 = third field) in a table entry. Limbs
// 0-7 of the three fields fill 16-byte chunks 0-1, 2-3 and 4-5; limbs 8-9 of y+x and y-x
// share chunk 6 and those of the third field start chunk 7 (2 pad words). y+x and y-x
// thus sit at symmetric chunk positions, and a negative digit (which swaps them) only
// changes which chunks a lane reads.
MK_DEV int mk_entry_word(int f, int i) { return i < 8 ? 8 * f + i : 24 + 2 * f + (i - 8); }
#endif
// Packed entries: field f is the canonical value in words 8f .. 8f + 7 (chunks 2f, 2f + 1),
// y+x and y-x again at mirrored chunk positions.

#if defined(__HIPCC__)
// One 128-byte entry: eight 16-byte vector loads of one cache line.
MK_DEV void mk_load_entry(const u32* __restrict__ table, u64 row, i32 d, u32 cols, mk_entry& e) {
    const u32 k = (u32)(d < 0 ? -d : d);
    const uint4* p = (const uint4*)(table + (row * cols + k) * MK_ENTRY_WORDS);
    MK_UNROLL
    for (int q = 0; q < MK_ENTRY_WORDS / 4; q++) {
        const uint4 v = p[q];
        e.w[4 * q] = v.x;
        e.w[4 * q + 1] = v.y;
        e.w[4 * q + 2] = v.z;
        e.w[4 * q + 3] = v.w;
    }
}
#else
MK_DEV void mk_load_entry(const u32* table, u64 row, i32 d, u32 cols, mk_entry& e) {
    const u32 k = (u32)(d < 0 ? -d : d);
    const u32* p = table + (row * cols + k) * MK_ENTRY_WORDS;
    for (int i = 0; i < MK_ENTRY_WORDS; i++) e.w[i] = p[i];
}
#endif

// Entry words to a Niels point, negated for a negative digit (-P = (-x, y): swap y+x
// and y-x, negate the third field).
MK_DEV mk_niels mk_entry_niels(const mk_entry& e, i32 d) {
    mk_niels n;
#if MK_PACKED
    n.ypx = fe_from_words(&e.w[0]);
    n.ymx = fe_from_words(&e.w[8]);
    n.xyd = fe_from_words(&e.w[16]);
#else
    MK_UNROLL
    for (int i = 0; i < 10; i++) {
        n.ypx.v[i] = e.w[mk_entry_word(0, i)];
        n.ymx.v[i] = e.w[mk_entry_word(1, i)];
        n.xyd.v[i] = e.w[mk_entry_word(2, i)];
    }
#endif
    if (d < 0) {
        const fe t = n.ypx;
        n.ypx = n.ymx;
        n.ymx = t;
        n.xyd = fe_neg(n.xyd);
    }
    return n;
}

// s*B as fractions yn/yd (and xn/xd) from the comb table of width w. The entry of the
// next row is loaded before the current row is added, so its latency overlaps the
// seven multiplications of the addition.
MK_DEV void mk_mul_base_frac(const u32* __restrict__ table, u32 w, const u32 s[8], fe& yn, fe& yd, fe& xn,
                             fe& xd) {
    mk_digits g = mk_digits_of(s, w);
    const u32 rows = g.rows, cols = mk_cols(w);
    i32 d = mk_next_digit(g);
    mk_entry cur, next;
    mk_load_entry(table, 0, d, cols, cur);
    i32 dn = mk_next_digit(g);
    mk_load_entry(table, 1, dn, cols, next);
    mk_ext acc = mk_row0_extended(mk_entry_niels(cur, d));
    MK_NO_UNROLL
    for (u32 row = 1; row + 1 < rows; row++) {
        cur = next;
        d = dn;
        dn = mk_next_digit(g);
        mk_load_entry(table, row + 1, dn, cols, next);
        acc = mk_add_niels(acc, mk_entry_niels(cur, d));
    }
    mk_add_niels_frac(acc, mk_entry_niels(next, dn), yn, yd, xn, xd);
}

// mk_mul_base_frac over a table in LDS: each row's entry is read just before its addition
// (LDS latency is short), so no entry is held in registers across an addition.
MK_DEV void mk_mul_base_frac_near(const u32* __restrict__ table, u32 w, const u32 s[8], fe& yn, fe& yd, fe& xn,
                                  fe& xd) {
    mk_digits g = mk_digits_of(s, w);
    const u32 rows = g.rows, cols = mk_cols(w);
    mk_entry e;
    i32 d = mk_next_digit(g);
    mk_load_entry(table, 0, d, cols, e);
    mk_ext acc = mk_row0_extended(mk_entry_niels(e, d));
    MK_NO_UNROLL
    for (u32 row = 1; row + 1 < rows; row++) {
        d = mk_next_digit(g);
        mk_load_entry(table, row, d, cols, e);
        acc = mk_add_niels(acc, mk_entry_niels(e, d));
    }
    d = mk_next_digit(g);
    mk_load_entry(table, rows - 1, d, cols, e);
    mk_add_niels_frac(acc, mk_entry_niels(e, d), yn, yd, xn, xd);
}

#if defined(__HIPCC__)
// Wait for every outstanding memory operation of this wave (the LDS copies included).
MK_DEV void mk_wait_all() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0);
#endif
}

// Staging buffers per wave: 2 (the next row's gather lands in the other buffer) or 1 (it
// lands in the same buffer once the current entry is in registers).
#ifndef MK_STAGE_BUFFERS
#define MK_STAGE_BUFFERS 2
#endif

// Every LDS read issued so far has completed, and no memory access moves across.
MK_DEV void mk_wait_lds() {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
#endif
}

// LDS staging of the table gathers (gfx950 global_load_lds_dwordx4): the next row's
// entry is copied straight from HBM into LDS without passing through VGPRs, so the
// prefetch costs no registers. `stage` is this wave's 2 x 8 KB double buffer laid out
// [buffer][16-byte chunk][lane]; chunk q of lane l sits at stage[(b * 8 + q) * 64 + l],
// so both the copy (one instruction per chunk for the whole wave) and the read
// (ds_read_b128, consecutive lanes on consecutive 16-byte words) are conflict-free.
MK_DEV void mk_stage_entry(const u32* __restrict__ table, u64 row, i32 d, u32 cols, uint4* stage, u32 b) {
    const u32 k = (u32)(d < 0 ? -d : d);
    const u32* e = table + (row * cols + k) * MK_ENTRY_WORDS;
#if defined(__HIP_DEVICE_COMPILE__)
    MK_UNROLL
    for (int q = 0; q < MK_STAGE_CHUNKS; q++)
        __builtin_amdgcn_global_load_lds((const void*)(e + 4 * q), (void*)(stage + (b * MK_STAGE_CHUNKS + q) * 64),
                                         16, 0, 0);
#else
    (void)e;
    (void)stage;
    (void)b;
#endif
}

#if MK_PACKED
MK_DEV mk_niels mk_staged_niels(const uint4* stage, u32 b, u32 lane, i32 d) {
    // chunk q of this lane: s[q * 64]. A negative digit reads y-x's chunks for y+x and
    // the other way round, so the swap costs no moves; the words become limbs here.
    const bool neg = d < 0;
    const uint4* s = stage + b * MK_STAGE_CHUNKS * 64 + lane;
    const uint4* pa = s + (neg ? 2 : 0) * 64;
    const uint4* pb = s + (neg ? 0 : 2) * 64;
    const uint4 a0 = pa[0], a1 = pa[64], b0 = pb[0], b1 = pb[64], c0 = s[4 * 64], c1 = s[5 * 64];
    const u32 wa[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
    const u32 wb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
    const u32 wc[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
    mk_niels n;
    n.ypx = fe_from_words(wa);
    n.ymx = fe_from_words(wb);
    n.xyd = fe_from_words(wc);
    if (neg) n.xyd = fe_neg(n.xyd);
    return n;
}
#else
MK_DEV mk_niels mk_staged_niels(const uint4* stage, u32 b, u32 lane, i32 d) {
    // chunk q of this lane: s[q * 64]. A negative digit reads y-x's chunks for y+x and
    // the other way round, and the other half of chunk 6, so the swap costs no moves.
    const bool neg = d < 0;
    const uint4* s = stage + b * MK_STAGE_CHUNKS * 64 + lane;
    const uint4* pa = s + (neg ? 2 : 0) * 64;
    const uint4* pb = s + (neg ? 0 : 2) * 64;
    const uint4 a0 = pa[0], a1 = pa[64], b0 = pb[0], b1 = pb[64], c0 = s[4 * 64], c1 = s[5 * 64];
    const u32* w6 = (const u32*)(s + 6 * 64);
    const uint2 a2 = *(const uint2*)(w6 + (neg ? 2 : 0)), b2 = *(const uint2*)(w6 + (neg ? 0 : 2));
    const uint2 c2 = *(const uint2*)(s + 7 * 64);
    mk_niels n;
    const u32 ya[10] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w, a2.x, a2.y};
    const u32 yb[10] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y};
    const u32 yc[10] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w, c2.x, c2.y};
    MK_UNROLL
    for (int i = 0; i < 10; i++) {
        n.ypx.v[i] = ya[i];
        n.ymx.v[i] = yb[i];
        n.xyd.v[i] = yc[i];
    }
    if (neg) n.xyd = fe_neg(n.xyd);
    return n;
}
#endif

// mk_mul_base_frac with the gathers staged through LDS: the entry of row r + 1 is in
// flight while row r is added. Every wave of the kernel calls this with all 64 lanes.
// row_sync() runs once at the top of each of the rows - 2 additions (the search kernel's
// lockstep variant puts a workgroup barrier there).
template <class R>
MK_DEV void mk_mul_base_frac_staged(const u32* __restrict__ table, u32 w, const u32 s[8], uint4* stage, u32 lane,
                                    fe& yn, fe& yd, fe& xn, fe& xd, R row_sync) {
    mk_digits g = mk_digits_of(s, w);
    const u32 rows = g.rows, cols = mk_cols(w);
    i32 d = mk_next_digit(g);
    mk_stage_entry(table, 0, d, cols, stage, 0);
    i32 dn = mk_next_digit(g);
#if MK_STAGE_BUFFERS == 1
    mk_wait_all();
    const mk_niels first = mk_staged_niels(stage, 0, lane, d);
    mk_wait_lds();  // row 0's entry is in registers: its buffer takes row 1's
    mk_stage_entry(table, 1, dn, cols, stage, 0);
    mk_ext acc = mk_row0_extended(first);
    MK_NO_UNROLL
    for (u32 row = 1; row + 1 < rows; row++) {
        row_sync();
        mk_wait_all();
        const mk_niels cur = mk_staged_niels(stage, 0, lane, dn);
        dn = mk_next_digit(g);
        mk_wait_lds();
        mk_stage_entry(table, row + 1, dn, cols, stage, 0);
        acc = mk_add_niels(acc, cur);
    }
    mk_wait_all();
    mk_add_niels_frac(acc, mk_staged_niels(stage, 0, lane, dn), yn, yd, xn, xd);
#else
    mk_stage_entry(table, 1, dn, cols, stage, 1);
    mk_wait_all();
    mk_ext acc = mk_row0_extended(mk_staged_niels(stage, 0, lane, d));
    MK_NO_UNROLL
    for (u32 row = 1; row + 1 < rows; row++) {
        row_sync();
        // the copy of this row's entry (issued one row ago) must have landed
        mk_wait_all();
        const mk_niels cur = mk_staged_niels(stage, row & 1, lane, dn);
        dn = mk_next_digit(g);
        // the other buffer was read into registers during the previous row
        mk_stage_entry(table, row + 1, dn, cols, stage, (row + 1) & 1);
        acc = mk_add_niels(acc, cur);
    }
    mk_wait_all();
    mk_add_niels_frac(acc, mk_staged_niels(stage, (rows - 1) & 1, lane, dn), yn, yd, xn, xd);
#endif
}

MK_DEV void mk_mul_base_frac_staged(const u32* __restrict__ table, u32 w, const u32 s[8], uint4* stage, u32 lane,
                                    fe& yn, fe& yd, fe& xn, fe& xd) {
    mk_mul_base_frac_staged(table, w, s, stage, lane, yn, yd, xn, xd, [] {});
}
#endif

// Full compressed public key (y with the x sign) of one scalar; one inversion.
MK_DEV void mk_full_pk(const u32* __restrict__ table, u32 w, const u32 s[8], u32 pk[8]) {
    fe yn, yd, xn, xd;
    mk_mul_base_frac(table, w, s, yn, yd, xn, xd);
    const fe inv = fe_invert(fe_mul(yd, xd));
    const fe y = fe_mul(fe_mul(yn, xd), inv);
    const fe x = fe_mul(fe_mul(yd, xn), inv);
    u32 xw[8];
    fe_carried_to_words(y, pk);
    fe_carried_to_words(x, xw);
    pk[7] |= (xw[0] & 1u) << 31;
}

// ------------------------------------------------------------------ constants

// Base point B and 2d as little-endian words (RFC 8032 s5.1).
MK_CONST u32 MK_BX[8] = {0x8f25d51au, 0xc9562d60u, 0x9525a7b2u, 0x692cc760u,
                         0xfdd6dc5cu, 0xc0a4e231u, 0xcd6e53feu, 0x216936d3u};
MK_CONST u32 MK_BY[8] = {0x66666658u, 0x66666666u, 0x66666666u, 0x66666666u,
                         0x66666666u, 0x66666666u, 0x66666666u, 0x66666666u};
MK_CONST u32 MK_D2[8] = {0x26b2f159u, 0xebd69b94u, 0x8283b156u, 0x00e0149au,
                         0xeef3d130u, 0x198e80f2u, 0x56dffce7u, 0x2406d9dcu};
// 1/2 mod p = (p + 1) / 2 = 2^254 - 9.
MK_CONST u32 MK_INV2[8] = {0xfffffff7u, 0xffffffffu, 0xffffffffu, 0xffffffffu,
                           0xffffffffu, 0xffffffffu, 0xffffffffu, 0x3fffffffu};

MK_DEV fe mk_const(const u32* w) {
    u32 t[8];
    MK_UNROLL
    for (int i = 0; i < 8; i++) t[i] = w[i];
    return fe_from_words(t);
}

// Unified addition in extended coordinates (RFC 8032 s5.1.4), complete on
// edwards25519, used for doubling as well. Table build only.
MK_DEV mk_ext mk_ext_add(const mk_ext& p, const mk_ext& q, const fe& d2) {
    const fe a = fe_mul(fe_sub(p.y, p.x), fe_sub(q.y, q.x));
    const fe b = fe_mul(fe_add(p.y, p.x), fe_add(q.y, q.x));
    const fe c = fe_mul(fe_mul(p.t, d2), q.t);
    const fe d = fe_mul(fe_add(p.z, p.z), q.z);
    const fe e = fe_sub(b, a), f = fe_sub(d, c), g = fe_add(d, c), h = fe_add(b, a);
    mk_ext r;
    r.x = fe_mul(e, f);
    r.y = fe_mul(g, h);
    r.t = fe_mul(e, h);
    r.z = fe_mul(f, g);
    return r;
}

MK_DEV mk_ext mk_ext_identity() {
    mk_ext p;
    p.x = fe_small(0);
    p.y = fe_small(1);
    p.z = fe_small(1);
    p.t = fe_small(0);
    return p;
}

// Row bases 2^(w*i) * 8B for every row, plus the top row's offset 2^254 * B, as
// extended points: out[i] for i < rows, out[rows] = the offset.
MK_DEV void mk_row_bases(u32 w, mk_ext* out) {
    const fe d2 = mk_const(MK_D2);
    mk_ext p;
    p.x = mk_const(MK_BX);
    p.y = mk_const(MK_BY);
    p.z = fe_small(1);
    p.t = fe_mul(p.x, p.y);
    for (int i = 0; i < 3; i++) p = mk_ext_add(p, p, d2);
    const u32 rows = mk_rows(w);
    for (u32 i = 0; i < rows; i++) {
        out[i] = p;
        if (i + 1 == rows) {
            // 2^c times the top row's base is 2^254 * B, c = 251 - w * (rows - 1)
            mk_ext t = p;
            for (u32 c = 0; c < MK_M_BITS - w * (rows - 1); c++) t = mk_ext_add(t, t, d2);
            out[rows] = t;
        }
        for (u32 c = 0; c < w; c++) p = mk_ext_add(p, p, d2);
    }
}

// Entry (row, k) = k * base_row (+ 2^254 * B on the top row) as ((y+x)/2, (y-x)/2, dxy),
// row 0 with xy as its third field, each field canonical: 32 words of limbs laid out by
// mk_entry_word (2 pad words zero), or packed as 3 x 8 words.
MK_DEV void mk_table_entry(u32 w, const mk_ext* bases, u32 row, u32 k, u32* out) {
    const u32 rows = mk_rows(w);
    const fe d2 = mk_const(MK_D2);
    const mk_ext base = bases[row];
    mk_ext m = mk_ext_identity();
    for (int b = (int)w - 1; b >= 0; b--) {
        m = mk_ext_add(m, m, d2);
        if ((k >> b) & 1u) m = mk_ext_add(m, base, d2);
    }
    const mk_ext acc = row + 1 == rows ? mk_ext_add(bases[rows], m, d2) : m;
    const fe zi = fe_invert(acc.z);
    const fe x = fe_mul(acc.x, zi), y = fe_mul(acc.y, zi);
    const fe xy = fe_mul(x, y);
    const fe half = mk_const(MK_INV2);
    fe parts[3];
    parts[0] = fe_mul(fe_add(y, x), half);
    parts[1] = fe_mul(fe_sub(y, x), half);
    parts[2] = row == 0 ? xy : fe_mul(xy, fe_mul(d2, half));
    for (int f = 0; f < 3; f++) {
        u32 wds[8];
        fe_carried_to_words(parts[f], wds);
#if MK_PACKED
        for (int i = 0; i < 8; i++) out[8 * f + i] = wds[i];
#else
        const fe c = fe_from_words(wds);  // canonical limbs
        for (int i = 0; i < 10; i++) out[mk_entry_word(f, i)] = c.v[i];
#endif
    }
#if !MK_PACKED
    out[30] = 0;
    out[31] = 0;
#endif
}

// s*B by textbook double-and-add over the 255 scalar bits with the unified addition:
// no table, about 380 additions. Used on the host to re-derive emitted keys
// independently of the comb table.
MK_DEV void mk_pk_by_ladder(const u32 s[8], u32 pk[8]) {
    const fe d2 = mk_const(MK_D2);
    mk_ext b;
    b.x = mk_const(MK_BX);
    b.y = mk_const(MK_BY);
    b.z = fe_small(1);
    b.t = fe_mul(b.x, b.y);
    mk_ext acc = mk_ext_identity();
    for (int i = 254; i >= 0; i--) {
        acc = mk_ext_add(acc, acc, d2);
        if ((s[i >> 5] >> (i & 31)) & 1u) acc = mk_ext_add(acc, b, d2);
    }
    const fe zi = fe_invert(acc.z);
    u32 xw[8];
    fe_carried_to_words(fe_mul(acc.y, zi), pk);
    fe_carried_to_words(fe_mul(acc.x, zi), xw);
    pk[7] |= (xw[0] & 1u) << 31;
}
