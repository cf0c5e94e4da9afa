/*
 * CPU oracle for the Ed25519 vanity-key search path — TEST INFRASTRUCTURE ONLY.
 * Loaded by tests/, the smoke check and the benchmark's CPU-baseline leg; never by
 * the product library. Reference paths below are relative to the meikey sources.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   ChaCha20 block ........ RFC 8439 s2.3.2 and Appendix A.1 block vectors.
 *   seed mapping .......... closed form of src/service/gpu/kernel.cu:134-176 checked
 *                           against an independent pure-Python ChaCha20 (RFC 8439) and
 *                           the consecutive-output property of gpu.rs:1685-1692.
 *   counter seed .......... src/model/search.rs:45-50, wrap-around case included.
 *   SHA-512 ............... hashlib.sha512 on random messages of 0..300 bytes.
 *   clamp + s*B ........... the five RFC 8032 s7.1 vectors of tests/data/rfc8032-ed25519.txt
 *                           (seed -> public key), and a pure-Python RFC 8032 s6 textbook
 *                           derivation on random seeds.
 *   comb multiply ......... equality with orc_mul_base_ladder (double-and-add) for every
 *                           width 2..16 on random clamped scalars; table entries against
 *                           the closed form k*2^(w*i+3)*B (+2^254*B on the top row).
 *   digits ................ reconstruction sum d_i*2^(w*i) == m and the digit ranges of
 *                           src/service/derive/comb.rs:410-445.
 *   field ops ............. Python integers on the edge values of gpu.rs:1869-1882.
 *   base64 ................ Python base64 of the blob, and the known line for RFC 8032
 *                           TEST 1 in src/model/key.rs:161-169.
 *   contains/suffix/regex . Python str operations / re.search on the 43-char hay, and the
 *                           fixed cases of src/model/pattern.rs:350-399 and dfa.rs:225-239.
 */
#include "meikey_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef unsigned __int128 u128;
typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;

/* ===================================================================== seeds */

#define ROTL32(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define QR(a, b, c, d)                     \
    a += b; d ^= a; d = ROTL32(d, 16);     \
    c += d; b ^= c; b = ROTL32(b, 12);     \
    a += b; d ^= a; d = ROTL32(d, 8);      \
    c += d; b ^= c; b = ROTL32(b, 7);

/* RFC 8439 block function with a 64-bit block counter in words 12-13 and a zero
 * stream id (rand's ChaCha20Rng layout). Follows kernel.cu:125-148. */
void orc_chacha20_block(const u32 key[8], u64 counter, u32 out[16]) {
    u32 s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                 key[4],      key[5],      key[6],      key[7],      (u32)counter, (u32)(counter >> 32), 0u, 0u};
    u32 x[16];
    memcpy(x, s, sizeof x);
    for (int r = 0; r < 10; r++) {
        QR(x[0], x[4], x[8], x[12]);
        QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]);
        QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]);
        QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]);
        QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

static void key_words(const u8 key[32], u32 w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (u32)key[4 * i] | (u32)key[4 * i + 1] << 8 | (u32)key[4 * i + 2] << 16 | (u32)key[4 * i + 3] << 24;
}

static void put_le32(u8* o, u32 v) {
    o[0] = (u8)v; o[1] = (u8)(v >> 8); o[2] = (u8)(v >> 16); o[3] = (u8)(v >> 24);
}

/* Seed of index i = 32-byte half (i & 1) of block i >> 1 (kernel.cu:160-176,
 * gpu.rs:1344-1351). */
void orc_chacha20_seed(const u8 key[32], u64 index, u8 seed[32]) {
    u32 k[8], b[16];
    key_words(key, k);
    orc_chacha20_block(k, index >> 1, b);
    for (int i = 0; i < 8; i++) put_le32(seed + 4 * i, b[(index & 1) * 8 + i]);
}

/* Low-security seed: bytes 24..31 as LE u64 plus the index, wrapping
 * (src/model/search.rs:45-50, kernel.cu:179-189). */
void orc_counter_seed(const u8 key[32], u64 index, u8 seed[32]) {
    memcpy(seed, key, 32);
    u64 tail = 0;
    for (int b = 7; b >= 0; b--) tail = tail << 8 | key[24 + b];
    tail += index;
    for (int b = 0; b < 8; b++) seed[24 + b] = (u8)(tail >> (8 * b));
}

void orc_seed(const u8 key[32], int source, u64 index, u8 seed[32]) {
    if (source == 1) orc_counter_seed(key, index, seed);
    else orc_chacha20_seed(key, index, seed);
}

/* =================================================================== SHA-512 */

static const u64 K512[80] = {
    0x428a2f98d728ae22ull, 0x7137449123ef65cdull, 0xb5c0fbcfec4d3b2full, 0xe9b5dba58189dbbcull, 0x3956c25bf348b538ull,
    0x59f111f1b605d019ull, 0x923f82a4af194f9bull, 0xab1c5ed5da6d8118ull, 0xd807aa98a3030242ull, 0x12835b0145706fbeull,
    0x243185be4ee4b28cull, 0x550c7dc3d5ffb4e2ull, 0x72be5d74f27b896full, 0x80deb1fe3b1696b1ull, 0x9bdc06a725c71235ull,
    0xc19bf174cf692694ull, 0xe49b69c19ef14ad2ull, 0xefbe4786384f25e3ull, 0x0fc19dc68b8cd5b5ull, 0x240ca1cc77ac9c65ull,
    0x2de92c6f592b0275ull, 0x4a7484aa6ea6e483ull, 0x5cb0a9dcbd41fbd4ull, 0x76f988da831153b5ull, 0x983e5152ee66dfabull,
    0xa831c66d2db43210ull, 0xb00327c898fb213full, 0xbf597fc7beef0ee4ull, 0xc6e00bf33da88fc2ull, 0xd5a79147930aa725ull,
    0x06ca6351e003826full, 0x142929670a0e6e70ull, 0x27b70a8546d22ffcull, 0x2e1b21385c26c926ull, 0x4d2c6dfc5ac42aedull,
    0x53380d139d95b3dfull, 0x650a73548baf63deull, 0x766a0abb3c77b2a8ull, 0x81c2c92e47edaee6ull, 0x92722c851482353bull,
    0xa2bfe8a14cf10364ull, 0xa81a664bbc423001ull, 0xc24b8b70d0f89791ull, 0xc76c51a30654be30ull, 0xd192e819d6ef5218ull,
    0xd69906245565a910ull, 0xf40e35855771202aull, 0x106aa07032bbd1b8ull, 0x19a4c116b8d2d0c8ull, 0x1e376c085141ab53ull,
    0x2748774cdf8eeb99ull, 0x34b0bcb5e19b48a8ull, 0x391c0cb3c5c95a63ull, 0x4ed8aa4ae3418acbull, 0x5b9cca4f7763e373ull,
    0x682e6ff3d6b2b8a3ull, 0x748f82ee5defb2fcull, 0x78a5636f43172f60ull, 0x84c87814a1f0ab72ull, 0x8cc702081a6439ecull,
    0x90befffa23631e28ull, 0xa4506cebde82bde9ull, 0xbef9a3f7b2c67915ull, 0xc67178f2e372532bull, 0xca273eceea26619cull,
    0xd186b8c721c0c207ull, 0xeada7dd6cde0eb1eull, 0xf57d4f7fee6ed178ull, 0x06f067aa72176fbaull, 0x0a637dc5a2c898a6ull,
    0x113f9804bef90daeull, 0x1b710b35131c471bull, 0x28db77f523047d84ull, 0x32caab7b40c72493ull, 0x3c9ebe0a15c9bebcull,
    0x431d67c49c100d4cull, 0x4cc5d4becb3e42b6ull, 0x597f299cfc657e2aull, 0x5fcb6fab3ad6faecull, 0x6c44198c4a475817ull};
static const u64 IV512[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull, 0xa54ff53a5f1d36f1ull,
                             0x510e527fade682d1ull, 0x9b05688c2b3e6c1full, 0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};

#define ROTR64(x, n) (((x) >> (n)) | ((x) << (64 - (n))))

static void sha512_block(u64 h[8], const u8 blk[128]) {
    u64 w[80];
    for (int t = 0; t < 16; t++) {
        u64 v = 0;
        for (int b = 0; b < 8; b++) v = v << 8 | blk[8 * t + b];
        w[t] = v;
    }
    for (int t = 16; t < 80; t++) {
        u64 s0 = ROTR64(w[t - 15], 1) ^ ROTR64(w[t - 15], 8) ^ (w[t - 15] >> 7);
        u64 s1 = ROTR64(w[t - 2], 19) ^ ROTR64(w[t - 2], 61) ^ (w[t - 2] >> 6);
        w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    u64 a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int t = 0; t < 80; t++) {
        u64 S1 = ROTR64(e, 14) ^ ROTR64(e, 18) ^ ROTR64(e, 41);
        u64 ch = (e & f) ^ (~e & g);
        u64 t1 = hh + S1 + ch + K512[t] + w[t];
        u64 S0 = ROTR64(a, 28) ^ ROTR64(a, 34) ^ ROTR64(a, 39);
        u64 mj = (a & b) ^ (a & c) ^ (b & c);
        u64 t2 = S0 + mj;
        hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
}

/* FIPS 180-4 SHA-512 of an arbitrary message (kernel.cu:195-303 specialises it to
 * one 32-byte block). */
void orc_sha512(const u8* msg, size_t len, u8 out[64]) {
    u64 h[8];
    memcpy(h, IV512, sizeof h);
    size_t off = 0;
    for (; off + 128 <= len; off += 128) sha512_block(h, msg + off);
    u8 last[256];
    size_t rem = len - off;
    memset(last, 0, sizeof last);
    if (rem) memcpy(last, msg + off, rem);
    last[rem] = 0x80;
    size_t nb = (rem + 1 + 16 <= 128) ? 1 : 2;
    u64 bits = (u64)len * 8;
    for (int b = 0; b < 8; b++) last[nb * 128 - 1 - b] = (u8)(bits >> (8 * b));
    sha512_block(h, last);
    if (nb == 2) sha512_block(h, last + 128);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 8; b++) out[8 * i + b] = (u8)(h[i] >> (56 - 8 * b));
}

/* clamp(SHA-512(seed)[..32]) (RFC 8032 s5.1.5 steps 1-2; kernel.cu:240-250, 288-296). */
void orc_secret_scalar(const u8 seed[32], u8 s[32]) {
    u8 d[64];
    orc_sha512(seed, 32, d);
    memcpy(s, d, 32);
    s[0] &= 248;
    s[31] &= 127;
    s[31] |= 64;
}

/* ============================================================ GF(2^255 - 19) */
/* 5 x 51-bit limbs with 128-bit products: an independent representation from the
 * reference's GPU field (src/service/derive/field.rs:19-218 uses the same radix). */

typedef struct { u64 v[5]; } fe;
#define M51 ((1ull << 51) - 1)

static fe fe_from_bytes(const u8 b[32]) {
    u64 w[4];
    for (int i = 0; i < 4; i++) {
        u64 v = 0;
        for (int k = 7; k >= 0; k--) v = v << 8 | b[8 * i + k];
        w[i] = v;
    }
    fe r;
    r.v[0] = w[0] & M51;
    r.v[1] = (w[0] >> 51 | w[1] << 13) & M51;
    r.v[2] = (w[1] >> 38 | w[2] << 26) & M51;
    r.v[3] = (w[2] >> 25 | w[3] << 39) & M51;
    r.v[4] = (w[3] >> 12) & M51;
    r.v[0] += 19 * (w[3] >> 63); /* 2^255 = 19 (mod p) */
    return r;
}

static fe fe_carry(fe a) {
    u64 c;
    for (int i = 0; i < 4; i++) { c = a.v[i] >> 51; a.v[i] &= M51; a.v[i + 1] += c; }
    c = a.v[4] >> 51; a.v[4] &= M51; a.v[0] += 19 * c;
    c = a.v[0] >> 51; a.v[0] &= M51; a.v[1] += c;
    return a;
}

static void fe_to_bytes(fe a, u8 out[32]) {
    a = fe_carry(fe_carry(a));
    /* a < 2^255 + small; subtract p when a >= p */
    u64 q = (a.v[0] + 19) >> 51;
    q = (a.v[1] + q) >> 51;
    q = (a.v[2] + q) >> 51;
    q = (a.v[3] + q) >> 51;
    q = (a.v[4] + q) >> 51;
    a.v[0] += 19 * q;
    u64 c;
    for (int i = 0; i < 4; i++) { c = a.v[i] >> 51; a.v[i] &= M51; a.v[i + 1] += c; }
    a.v[4] &= M51;
    u64 w0 = a.v[0] | a.v[1] << 51, w1 = a.v[1] >> 13 | a.v[2] << 38, w2 = a.v[2] >> 26 | a.v[3] << 25,
        w3 = a.v[3] >> 39 | a.v[4] << 12;
    u64 w[4] = {w0, w1, w2, w3};
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 8; k++) out[8 * i + k] = (u8)(w[i] >> (8 * k));
}

static fe fe_small(u64 x) { fe r = {{x, 0, 0, 0, 0}}; return r; }
static fe fe_add(fe a, fe b) { for (int i = 0; i < 5; i++) a.v[i] += b.v[i]; return fe_carry(a); }
static fe fe_sub(fe a, fe b) {
    b = fe_carry(b);
    /* a + 4p - b keeps every limb non-negative */
    static const u64 p4[5] = {0x1fffffffffffb4ull, 0x1ffffffffffffcull, 0x1ffffffffffffcull, 0x1ffffffffffffcull,
                              0x1ffffffffffffcull};
    for (int i = 0; i < 5; i++) a.v[i] = a.v[i] + p4[i] - b.v[i];
    return fe_carry(a);
}
static fe fe_neg(fe a) { return fe_sub(fe_small(0), a); }

static fe fe_mul(fe a, fe b) {
    a = fe_carry(a); b = fe_carry(b);
    u64 a0 = a.v[0], a1 = a.v[1], a2 = a.v[2], a3 = a.v[3], a4 = a.v[4];
    u64 b0 = b.v[0], b1 = b.v[1], b2 = b.v[2], b3 = b.v[3], b4 = b.v[4];
    u64 b1_19 = 19 * b1, b2_19 = 19 * b2, b3_19 = 19 * b3, b4_19 = 19 * b4;
    u128 r0 = (u128)a0 * b0 + (u128)a1 * b4_19 + (u128)a2 * b3_19 + (u128)a3 * b2_19 + (u128)a4 * b1_19;
    u128 r1 = (u128)a0 * b1 + (u128)a1 * b0 + (u128)a2 * b4_19 + (u128)a3 * b3_19 + (u128)a4 * b2_19;
    u128 r2 = (u128)a0 * b2 + (u128)a1 * b1 + (u128)a2 * b0 + (u128)a3 * b4_19 + (u128)a4 * b3_19;
    u128 r3 = (u128)a0 * b3 + (u128)a1 * b2 + (u128)a2 * b1 + (u128)a3 * b0 + (u128)a4 * b4_19;
    u128 r4 = (u128)a0 * b4 + (u128)a1 * b3 + (u128)a2 * b2 + (u128)a3 * b1 + (u128)a4 * b0;
    fe r;
    r1 += (u64)(r0 >> 51); r.v[0] = (u64)r0 & M51;
    r2 += (u64)(r1 >> 51); r.v[1] = (u64)r1 & M51;
    r3 += (u64)(r2 >> 51); r.v[2] = (u64)r2 & M51;
    r4 += (u64)(r3 >> 51); r.v[3] = (u64)r3 & M51;
    u64 c = (u64)(r4 >> 51); r.v[4] = (u64)r4 & M51;
    r.v[0] += c * 19;
    return fe_carry(r);
}
static fe fe_sq(fe a) { return fe_mul(a, a); }

static fe fe_pow2k(fe a, int k) { while (k-- > 0) a = fe_sq(a); return a; }

/* a^(p-2) by the addition chain of src/service/derive/field.rs:108-124 (kernel.cu:1499-1512). */
static fe fe_invert(fe a) {
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

void orc_fe_op(int op, const u8 a[32], const u8 b[32], u8 out[32]) {
    fe x = fe_from_bytes(a), y = fe_from_bytes(b), r;
    switch (op) {
        case 0: r = fe_add(x, y); break;
        case 1: r = fe_sub(x, y); break;
        case 2: r = fe_mul(x, y); break;
        case 3: r = fe_sq(x); break;
        default: r = fe_invert(x); break;
    }
    fe_to_bytes(r, out);
}

/* ============================================================ curve points */

static const u8 D2_BYTES[32] = {0x59, 0xf1, 0xb2, 0x26, 0x94, 0x9b, 0xd6, 0xeb, 0x56, 0xb1, 0x83,
                                0x82, 0x9a, 0x14, 0xe0, 0x00, 0x30, 0xd1, 0xf3, 0xee, 0xf2, 0x80,
                                0x8e, 0x19, 0xe7, 0xfc, 0xdf, 0x56, 0xdc, 0xd9, 0x06, 0x24};
static const u8 BX_BYTES[32] = {0x1a, 0xd5, 0x25, 0x8f, 0x60, 0x2d, 0x56, 0xc9, 0xb2, 0xa7, 0x25,
                                0x95, 0x60, 0xc7, 0x2c, 0x69, 0x5c, 0xdc, 0xd6, 0xfd, 0x31, 0xe2,
                                0xa4, 0xc0, 0xfe, 0x53, 0x6e, 0xcd, 0xd3, 0x36, 0x69, 0x21};
static const u8 BY_BYTES[32] = {0x58, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66,
                                0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66,
                                0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66, 0x66};

/* Twisted Edwards (a = -1) extended coordinates x = X/Z, y = Y/Z, xy = T/Z. */
typedef struct { fe x, y, z, t; } ge;
/* Affine point prepared for addition: (y+x, y-x, 2dxy) (comb.rs:84-100). */
typedef struct { fe ypx, ymx, xy2d; } niels;

static ge ge_identity(void) { ge p = {fe_small(0), fe_small(1), fe_small(1), fe_small(0)}; return p; }

static ge ge_from_affine(fe x, fe y) { ge p = {x, y, fe_small(1), fe_mul(x, y)}; return p; }

/* Unified addition, RFC 8032 s5.1.4 (also valid for doubling). */
static ge ge_add(ge p, ge q) {
    fe d2 = fe_from_bytes(D2_BYTES);
    fe a = fe_mul(fe_sub(p.y, p.x), fe_sub(q.y, q.x));
    fe b = fe_mul(fe_add(p.y, p.x), fe_add(q.y, q.x));
    fe c = fe_mul(fe_mul(p.t, d2), q.t);
    fe d = fe_mul(fe_add(p.z, p.z), q.z);
    fe e = fe_sub(b, a), f = fe_sub(d, c), g = fe_add(d, c), h = fe_add(b, a);
    ge r = {fe_mul(e, f), fe_mul(g, h), fe_mul(f, g), fe_mul(e, h)};
    return r;
}

/* Extended + affine Niels point, 7 multiplications (kernel.cu:1598-1611, comb.rs:102-178). */
static ge ge_add_niels(ge p, niels q) {
    fe pp = fe_mul(fe_add(p.y, p.x), q.ypx);
    fe mm = fe_mul(fe_sub(p.y, p.x), q.ymx);
    fe txy = fe_mul(p.t, q.xy2d);
    fe z2 = fe_add(p.z, p.z);
    fe x = fe_sub(pp, mm), y = fe_add(pp, mm);
    fe z = fe_add(z2, txy), t = fe_sub(z2, txy);
    ge r = {fe_mul(t, x), fe_mul(y, z), fe_mul(t, z), fe_mul(x, y)};
    return r;
}

static niels niels_from_affine(fe x, fe y) {
    niels n = {fe_add(y, x), fe_sub(y, x), fe_mul(fe_mul(x, y), fe_from_bytes(D2_BYTES))};
    return n;
}

static niels niels_neg(niels n) { niels r = {n.ymx, n.ypx, fe_neg(n.xy2d)}; return r; }

static ge ge_base(void) { return ge_from_affine(fe_from_bytes(BX_BYTES), fe_from_bytes(BY_BYTES)); }

static void ge_to_affine(ge p, fe* x, fe* y) {
    fe zi = fe_invert(p.z);
    *x = fe_mul(p.x, zi);
    *y = fe_mul(p.y, zi);
}

/* Compressed encoding: y little-endian with the parity of x in bit 255
 * (RFC 8032 s5.1.2; kernel.cu:1980-1983). */
static void encode_xy(fe x, fe y, u8 pk[32]) {
    u8 xb[32];
    fe_to_bytes(y, pk);
    fe_to_bytes(x, xb);
    pk[31] |= (u8)((xb[0] & 1) << 7);
}

static void ge_encode(ge p, u8 pk[32]) {
    fe x, y;
    ge_to_affine(p, &x, &y);
    encode_xy(x, y, pk);
}

/* Textbook double-and-add over all 256 scalar bits (RFC 8032 s6 point_mul). */
void orc_mul_base_ladder(const u8 s[32], u8 pk[32]) {
    ge b = ge_base(), r = ge_identity();
    for (int i = 255; i >= 0; i--) {
        r = ge_add(r, r);
        if ((s[i >> 3] >> (i & 7)) & 1) r = ge_add(r, b);
    }
    ge_encode(r, pk);
}

/* ======================================================================= comb */

#define M_BITS 251
#define MIN_W 2
#define MAX_ORC_W 16

/* Rows = ceil(251 / w), cols = 2^(w-1) + 1 (comb.rs:285-287). */
static int comb_rows(int w) { return (M_BITS + w - 1) / w; }
static u32 comb_cols(int w) { return (1u << (w - 1)) + 1; }

/* Signed base-2^w digits of m, where the clamped s = 8 * (2^251 + m): digits in
 * [-2^(w-1), 2^(w-1)), the top one in [0, 2^(w-1)] (comb.rs:414-445). */
int orc_digits(const u8 s[32], int w, int32_t* out) {
    u64 words[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < 4; i++) {
        u64 v = 0;
        for (int b = 7; b >= 0; b--) v = v << 8 | s[8 * i + b];
        words[i] = v;
    }
    for (int i = 0; i < 4; i++) words[i] = words[i] >> 3 | words[i + 1] << 61;
    words[3] &= (1ull << (M_BITS - 192)) - 1;
    words[4] = 0;
    int rows = comb_rows(w);
    int32_t half = 1 << (w - 1), carry = 0;
    for (int i = 0; i < rows; i++) {
        int off = i * w, q = off / 64, r = off % 64;
        u64 bits = words[q] >> r;
        if (r + w > 64 && q + 1 < 5) bits |= words[q + 1] << (64 - r);
        int32_t v = (int32_t)(bits & ((1ull << w) - 1)) + carry;
        if (i == rows - 1) {
            out[i] = v;
        } else {
            carry = (v + half) >> w;
            out[i] = v - (carry << w);
        }
    }
    return rows;
}

/* One table per width, built on first use: entry (i, k) = k * 2^(w*i) * 8B for the
 * lower rows and (k + 2^c) * 2^(w*(R-1)) * 8B on the top row, c = 251 - w*(R-1),
 * so the top row also carries 2^254 * B (comb.rs:289-339). */
typedef struct {
    int built;
    niels* e;
} comb_table;
static comb_table TABLES[MAX_ORC_W + 1];
static pthread_mutex_t TABLE_LOCK = PTHREAD_MUTEX_INITIALIZER;

static ge ge_double_n(ge p, int n) { while (n-- > 0) p = ge_add(p, p); return p; }

static const niels* comb_get(int w) {
    pthread_mutex_lock(&TABLE_LOCK);
    comb_table* t = &TABLES[w];
    if (!t->built) {
        int rows = comb_rows(w);
        u32 cols = comb_cols(w);
        t->e = (niels*)malloc(sizeof(niels) * (size_t)rows * cols);
        ge base = ge_double_n(ge_base(), 3); /* 8B */
        for (int i = 0; i < rows; i++) {
            ge acc = ge_identity();
            if (i == rows - 1) acc = ge_double_n(base, M_BITS - w * (rows - 1));
            for (u32 k = 0; k < cols; k++) {
                fe x, y;
                ge_to_affine(acc, &x, &y);
                t->e[(size_t)i * cols + k] = niels_from_affine(x, y);
                acc = ge_add(acc, base);
            }
            base = ge_double_n(base, w);
        }
        t->built = 1;
    }
    pthread_mutex_unlock(&TABLE_LOCK);
    return t->e;
}

static ge mul_base_comb_point(const u8 s[32], int w) {
    const niels* tab = comb_get(w);
    int32_t d[(M_BITS + MIN_W - 1) / MIN_W];
    int rows = orc_digits(s, w, d);
    u32 cols = comb_cols(w);
    ge acc = ge_identity();
    for (int i = 0; i < rows; i++) {
        int32_t di = d[i];
        niels n = tab[(size_t)i * cols + (u32)(di < 0 ? -di : di)];
        acc = ge_add_niels(acc, di < 0 ? niels_neg(n) : n);
    }
    return acc;
}

int orc_mul_base_comb(const u8 s[32], int w, u8 pk[32]) {
    if (w < MIN_W || w > MAX_ORC_W) return -1;
    ge_encode(mul_base_comb_point(s, w), pk);
    return 0;
}

int orc_comb_entry(int w, int row, u32 k, u8 pk[32]) {
    if (w < MIN_W || w > 30 || row < 0 || row >= comb_rows(w) || k >= comb_cols(w)) return -1;
    /* closed form by double-and-add: (k [+ 2^c]) * 2^(w*row + 3) * B */
    int rows = comb_rows(w);
    u8 sc[40];
    memset(sc, 0, sizeof sc);
    /* scalar = k << (w*row + 3), plus 2^254 on the top row */
    int sh = w * row + 3;
    for (int b = 0; b < 32; b++)
        if ((k >> b) & 1) {
            int bit = sh + b;
            if (bit < 320) sc[bit >> 3] |= (u8)(1u << (bit & 7));
        }
    ge r = ge_identity(), base = ge_base();
    if (row == rows - 1) r = ge_double_n(base, 254);
    ge acc = ge_identity();
    for (int i = 319; i >= 0; i--) {
        acc = ge_add(acc, acc);
        if ((sc[i >> 3] >> (i & 7)) & 1) acc = ge_add(acc, base);
    }
    ge_encode(ge_add(acc, r), pk);
    return 0;
}

int orc_comb_entry_niels(int w, int row, u32 k, u8 out[96]) {
    if (w < MIN_W || w > MAX_ORC_W || row < 0 || row >= comb_rows(w) || k >= comb_cols(w)) return -1;
    const niels* tab = comb_get(w);
    niels n = tab[(size_t)row * comb_cols(w) + k];
    fe_to_bytes(n.ypx, out);
    fe_to_bytes(n.ymx, out + 32);
    fe_to_bytes(n.xy2d, out + 64);
    return 0;
}

/* ======================================================================= keys */

/* The comb width of the reference's CPU path (comb.rs:17). */
#define ORC_W 12

void orc_public_key(const u8 seed[32], u8 pk[32]) {
    u8 s[32];
    orc_secret_scalar(seed, s);
    orc_mul_base_comb(s, ORC_W, pk);
}

static const char B64[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";

/* base64 of the 51-byte blob string("ssh-ed25519") || string(pk) (src/model/key.rs:9-52). */
void orc_public_b64(const u8 pk[32], char out[68]) {
    u8 blob[51];
    memset(blob, 0, 19);
    blob[3] = 11;
    memcpy(blob + 4, "ssh-ed25519", 11);
    blob[18] = 32;
    memcpy(blob + 19, pk, 32);
    for (int g = 0; g < 17; g++) {
        u32 n = (u32)blob[3 * g] << 16 | (u32)blob[3 * g + 1] << 8 | blob[3 * g + 2];
        for (int k = 0; k < 4; k++) out[4 * g + k] = B64[(n >> (18 - 6 * k)) & 63];
    }
}

/* Montgomery batch inversion over a chunk of keys (src/service/derive.rs:51 uses 256). */
#define DERIVE_CHUNK 64

static void derive_chunk(const u8 key[32], int source, u64 first, u64 n, u8* pks) {
    ge pts[DERIVE_CHUNK];
    fe pre[DERIVE_CHUNK];
    for (u64 j = 0; j < n; j++) {
        u8 seed[32], s[32];
        orc_seed(key, source, first + j, seed);
        orc_secret_scalar(seed, s);
        pts[j] = mul_base_comb_point(s, ORC_W);
        pre[j] = j == 0 ? pts[j].z : fe_mul(pre[j - 1], pts[j].z);
    }
    fe inv = fe_invert(pre[n - 1]);
    for (u64 j = n; j-- > 0;) {
        fe zi = j == 0 ? inv : fe_mul(inv, pre[j - 1]);
        inv = fe_mul(inv, pts[j].z);
        encode_xy(fe_mul(pts[j].x, zi), fe_mul(pts[j].y, zi), pks + 32 * j);
    }
}

typedef struct {
    const u8* key;
    int source;
    u64 first, n;
    u8* pks;
    /* search mode */
    const u8* needle;
    u32 len;
    int cs, mode;
    u64* hits;
    u64 cap, count;
} job;

static int auto_threads(int threads, u64 n) {
    if (threads <= 0) {
        long c = sysconf(_SC_NPROCESSORS_ONLN);
        threads = c > 0 ? (int)c : 1;
        if (threads > 16) threads = 16;
    }
    u64 maxt = (n + 1023) / 1024;
    if (maxt < 1) maxt = 1;
    if ((u64)threads > maxt) threads = (int)maxt;
    return threads;
}

static void* derive_worker(void* arg) {
    job* j = (job*)arg;
    for (u64 off = 0; off < j->n; off += DERIVE_CHUNK) {
        u64 m = j->n - off < DERIVE_CHUNK ? j->n - off : DERIVE_CHUNK;
        derive_chunk(j->key, j->source, j->first + off, m, j->pks + 32 * off);
    }
    return NULL;
}

void orc_derive(const u8 key[32], int source, u64 first, u64 n, u8* pks, int threads) {
    if (n == 0) return;
    comb_get(ORC_W);
    int t = auto_threads(threads, n);
    pthread_t th[64];
    job jobs[64];
    u64 per = (n + t - 1) / t;
    int started = 0;
    for (int i = 0; i < t; i++) {
        u64 a = per * i;
        if (a >= n) break;
        u64 m = n - a < per ? n - a : per;
        jobs[i] = (job){key, source, first + a, m, pks + 32 * a, 0, 0, 0, 0, 0, 0, 0};
        pthread_create(&th[i], NULL, derive_worker, &jobs[i]);
        started++;
    }
    for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
}

/* ================================================================== matching */

static int b64_value(u8 c) {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return 64;
}

static u8 lower(u8 c) { return (c >= 'A' && c <= 'Z') ? (u8)(c + 32) : c; }

static int same(const char* w, const u8* needle, u32 len, int cs) {
    for (u32 i = 0; i < len; i++) {
        u8 a = (u8)w[i], b = needle[i];
        if (cs ? a != b : lower(a) != lower(b)) return 0;
    }
    return 1;
}

#define SEARCH_START 25
#define B64_LEN 68

static u32 dfa_next(const u8* table, u32 s, u32 i) {
    u32 at = (s * 65u + i) * 2u;
    return (u32)table[at] | (u32)table[at + 1] << 8;
}

/* Pattern::find semantics on the hay b64[25..68] (src/model/pattern.rs:192-220);
 * the DFA walk follows Dfa::is_match_values (src/model/pattern/dfa.rs:117-137):
 * state 0 = ACCEPT (absorbing), 1 = START, column 64 = end-of-input flag. */
int orc_match(const char b64[68], const u8* needle, u32 len, int cs, int mode) {
    if (mode == ORC_REGEX) {
        u32 s = 1;
        for (int i = SEARCH_START; i < B64_LEN; i++) {
            s = dfa_next(needle, s, (u32)b64_value((u8)b64[i]) & 63);
            if (s == 0) return 1;
        }
        return dfa_next(needle, s, 64) != 0;
    }
    if (len == 0 || len > B64_LEN - SEARCH_START) return 0;
    if (mode == ORC_SUFFIX) return same(b64 + B64_LEN - len, needle, len, cs);
    for (u32 st = SEARCH_START; st + len <= B64_LEN; st++)
        if (same(b64 + st, needle, len, cs)) return 1;
    return 0;
}

int orc_match_unsigned(const u8 pk[32], const u8* needle, u32 len, int cs, int mode) {
    u8 k[32];
    char b64[68];
    memcpy(k, pk, 32);
    k[31] &= 0x7f;
    orc_public_b64(k, b64);
    if (orc_match(b64, needle, len, cs, mode)) return 1;
    k[31] |= 0x80;
    orc_public_b64(k, b64);
    return orc_match(b64, needle, len, cs, mode);
}

static void* search_worker(void* arg) {
    job* j = (job*)arg;
    u8 pks[32 * DERIVE_CHUNK];
    j->count = 0;
    for (u64 off = 0; off < j->n; off += DERIVE_CHUNK) {
        u64 m = j->n - off < DERIVE_CHUNK ? j->n - off : DERIVE_CHUNK;
        derive_chunk(j->key, j->source, j->first + off, m, pks);
        for (u64 i = 0; i < m; i++) {
            char b64[68];
            orc_public_b64(pks + 32 * i, b64);
            if (orc_match(b64, j->needle, j->len, j->cs, j->mode)) {
                if (j->count < j->cap) j->hits[j->count] = j->first + off + i;
                j->count++;
            }
        }
    }
    return NULL;
}

u64 orc_search(const u8 key[32], int source, u64 first, u64 n, const u8* needle, u32 len, int cs, int mode,
               u64* hits, u64 cap, int threads) {
    if (n == 0) return 0;
    comb_get(ORC_W);
    int t = auto_threads(threads, n);
    pthread_t th[64];
    job jobs[64];
    u64 per = (n + t - 1) / t;
    int started = 0;
    for (int i = 0; i < t; i++) {
        u64 a = per * i;
        if (a >= n) break;
        u64 m = n - a < per ? n - a : per;
        jobs[i] = (job){key, source, first + a, m, NULL, needle, len, cs, mode, NULL, 0, 0};
        jobs[i].hits = (u64*)malloc(sizeof(u64) * (cap ? cap : 1));
        jobs[i].cap = cap;
        pthread_create(&th[i], NULL, search_worker, &jobs[i]);
        started++;
    }
    u64 total = 0;
    for (int i = 0; i < started; i++) {
        pthread_join(th[i], NULL);
        for (u64 h = 0; h < jobs[i].count && h < jobs[i].cap; h++)
            if (total + h < cap) hits[total + h] = jobs[i].hits[h];
        total += jobs[i].count;
        free(jobs[i].hits);
    }
    return total;
}
