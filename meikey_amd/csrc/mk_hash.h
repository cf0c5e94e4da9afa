// Seed generation and the secret scalar: ChaCha20 (RFC 8439 block function with a
// 64-bit block counter) and SHA-512 of one 32-byte seed, clamped (RFC 8032 s5.1.5).
//
// Seed of key index i (reference src/service/gpu/kernel.cu:125-176, gpu.rs:1344-1351):
// ChaCha20 block i >> 1 under the search's 32-byte key, first 32 bytes for even i,
// last 32 bytes for odd i. One block therefore serves two consecutive indices.
// Low-security seeds (kernel.cu:179-189, src/model/search.rs:45-50): the key with its
// last 8 bytes read as a little-endian u64 plus i (wrapping).
//
// Words are little-endian u32 throughout, the byte order the seed and the scalar
// have in memory; SHA-512's big-endian message words are formed with byte swaps.
#pragma once

#include "mk_common.h"

#define MK_ROTL32(x, n) (((x) << (n)) | ((x) >> (32 - (n))))
#define MK_QR(a, b, c, d)                        \
    a += b; d ^= a; d = MK_ROTL32(d, 16);        \
    c += d; b ^= c; b = MK_ROTL32(b, 12);        \
    a += b; d ^= a; d = MK_ROTL32(d, 8);         \
    c += d; b ^= c; b = MK_ROTL32(b, 7);

MK_DEV void mk_chacha20_block(const u32 key[8], u64 counter, u32 out[16]) {
    u32 x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                 key[4],      key[5],      key[6],      key[7],      (u32)counter, (u32)(counter >> 32), 0u, 0u};
    MK_UNROLL
    for (int r = 0; r < 10; r++) {
        MK_QR(x[0], x[4], x[8], x[12]);
        MK_QR(x[1], x[5], x[9], x[13]);
        MK_QR(x[2], x[6], x[10], x[14]);
        MK_QR(x[3], x[7], x[11], x[15]);
        MK_QR(x[0], x[5], x[10], x[15]);
        MK_QR(x[1], x[6], x[11], x[12]);
        MK_QR(x[2], x[7], x[8], x[13]);
        MK_QR(x[3], x[4], x[9], x[14]);
    }
    out[0] = x[0] + 0x61707865u;
    out[1] = x[1] + 0x3320646eu;
    out[2] = x[2] + 0x79622d32u;
    out[3] = x[3] + 0x6b206574u;
    MK_UNROLL
    for (int i = 0; i < 8; i++) out[4 + i] = x[4 + i] + key[i];
    out[12] = x[12] + (u32)counter;
    out[13] = x[13] + (u32)(counter >> 32);
    out[14] = x[14];
    out[15] = x[15];
}

MK_DEV void mk_counter_seed(const u32 key[8], u64 index, u32 seed[8]) {
    u64 tail = ((u64)key[7] << 32 | key[6]) + index;
    MK_UNROLL
    for (int i = 0; i < 6; i++) seed[i] = key[i];
    seed[6] = (u32)tail;
    seed[7] = (u32)(tail >> 32);
}

MK_DEV u32 mk_bswap32(u32 x) { return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24); }

// SHA-512 word operations on the two 32-bit halves of a u64: every rotate is two funnel
// shifts (v_alignbit_b32), three-input XOR, choose and majority are one v_bitop3_b32 per
// half on gfx950, and 64-bit additions stay single v_lshl_add_u64. The halves are taken
// and rejoined through a <2 x u32> bit cast, which the compiler keeps as plain register
// pairs (building the u64 with shift-and-or instead lets it split the later 64-bit sums
// into 32-bit pieces with extra moves).
#if defined(__HIP_DEVICE_COMPILE__)
typedef u32 mk_u32x2 __attribute__((ext_vector_type(2)));
MK_DEV u32 mk_lo(u64 x) { return __builtin_bit_cast(mk_u32x2, x).x; }
MK_DEV u32 mk_hi(u64 x) { return __builtin_bit_cast(mk_u32x2, x).y; }
MK_DEV u64 mk_join(u32 lo, u32 hi) {
    mk_u32x2 v;
    v.x = lo;
    v.y = hi;
    return __builtin_bit_cast(u64, v);
}
MK_DEV u32 mk_funnel(u32 hi, u32 lo, u32 n) { return __builtin_amdgcn_alignbit(hi, lo, n); }
// v_bitop3_b32 truth tables index the inputs as a = 0xF0, b = 0xCC, c = 0xAA:
// a ^ b ^ c = 0x96, maj(a, b, c) = 0xE8, a ? b : c = 0xCA.
MK_DEV u32 mk_xor3(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }
MK_DEV u32 mk_maj(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8); }
MK_DEV u32 mk_choose(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0xCA); }
#else
MK_DEV u32 mk_lo(u64 x) { return (u32)x; }
MK_DEV u32 mk_hi(u64 x) { return (u32)(x >> 32); }
MK_DEV u64 mk_join(u32 lo, u32 hi) { return (u64)hi << 32 | lo; }
// (hi:lo) >> n, low 32 bits, for n in 0..31
MK_DEV u32 mk_funnel(u32 hi, u32 lo, u32 n) { return (u32)((((u64)hi << 32) | lo) >> (n & 31)); }
MK_DEV u32 mk_xor3(u32 a, u32 b, u32 c) { return a ^ b ^ c; }
MK_DEV u32 mk_maj(u32 a, u32 b, u32 c) { return (a & b) | (a & c) | (b & c); }
MK_DEV u32 mk_choose(u32 a, u32 b, u32 c) { return (a & b) | (~a & c); }
#endif

// Halves of x rotated right by a constant n in 1..63.
MK_DEV u32 mk_rotr_lo(u64 x, int n) {
    return n < 32 ? mk_funnel(mk_hi(x), mk_lo(x), (u32)n) : mk_funnel(mk_lo(x), mk_hi(x), (u32)(n - 32));
}
MK_DEV u32 mk_rotr_hi(u64 x, int n) {
    return n < 32 ? mk_funnel(mk_lo(x), mk_hi(x), (u32)n) : mk_funnel(mk_hi(x), mk_lo(x), (u32)(n - 32));
}

// rotr(x, a) ^ rotr(x, b) ^ rotr(x, c): the round functions Sigma0 / Sigma1
MK_DEV u64 mk_rot3(u64 x, int a, int b, int c) {
    return mk_join(mk_xor3(mk_rotr_lo(x, a), mk_rotr_lo(x, b), mk_rotr_lo(x, c)),
                   mk_xor3(mk_rotr_hi(x, a), mk_rotr_hi(x, b), mk_rotr_hi(x, c)));
}

// rotr(x, a) ^ rotr(x, b) ^ (x >> c), c in 1..31: the schedule functions sigma0 / sigma1
MK_DEV u64 mk_rot2_shr(u64 x, int a, int b, int c) {
    return mk_join(mk_xor3(mk_rotr_lo(x, a), mk_rotr_lo(x, b), mk_funnel(mk_hi(x), mk_lo(x), (u32)c)),
                   mk_xor3(mk_rotr_hi(x, a), mk_rotr_hi(x, b), mk_hi(x) >> c));
}

MK_DEV u64 mk_ch64(u64 e, u64 f, u64 g) {
    return mk_join(mk_choose(mk_lo(e), mk_lo(f), mk_lo(g)), mk_choose(mk_hi(e), mk_hi(f), mk_hi(g)));
}

MK_DEV u64 mk_maj64(u64 a, u64 b, u64 c) {
    return mk_join(mk_maj(mk_lo(a), mk_lo(b), mk_lo(c)), mk_maj(mk_hi(a), mk_hi(b), mk_hi(c)));
}

// FIPS 180-4 round constants.
#define MK_SHA512_K                                                                                                 \
    {0x428a2f98d728ae22ull, 0x7137449123ef65cdull, 0xb5c0fbcfec4d3b2full, 0xe9b5dba58189dbbcull, 0x3956c25bf348b538ull, \
     0x59f111f1b605d019ull, 0x923f82a4af194f9bull, 0xab1c5ed5da6d8118ull, 0xd807aa98a3030242ull, 0x12835b0145706fbeull, \
     0x243185be4ee4b28cull, 0x550c7dc3d5ffb4e2ull, 0x72be5d74f27b896full, 0x80deb1fe3b1696b1ull, 0x9bdc06a725c71235ull, \
     0xc19bf174cf692694ull, 0xe49b69c19ef14ad2ull, 0xefbe4786384f25e3ull, 0x0fc19dc68b8cd5b5ull, 0x240ca1cc77ac9c65ull, \
     0x2de92c6f592b0275ull, 0x4a7484aa6ea6e483ull, 0x5cb0a9dcbd41fbd4ull, 0x76f988da831153b5ull, 0x983e5152ee66dfabull, \
     0xa831c66d2db43210ull, 0xb00327c898fb213full, 0xbf597fc7beef0ee4ull, 0xc6e00bf33da88fc2ull, 0xd5a79147930aa725ull, \
     0x06ca6351e003826full, 0x142929670a0e6e70ull, 0x27b70a8546d22ffcull, 0x2e1b21385c26c926ull, 0x4d2c6dfc5ac42aedull, \
     0x53380d139d95b3dfull, 0x650a73548baf63deull, 0x766a0abb3c77b2a8ull, 0x81c2c92e47edaee6ull, 0x92722c851482353bull, \
     0xa2bfe8a14cf10364ull, 0xa81a664bbc423001ull, 0xc24b8b70d0f89791ull, 0xc76c51a30654be30ull, 0xd192e819d6ef5218ull, \
     0xd69906245565a910ull, 0xf40e35855771202aull, 0x106aa07032bbd1b8ull, 0x19a4c116b8d2d0c8ull, 0x1e376c085141ab53ull, \
     0x2748774cdf8eeb99ull, 0x34b0bcb5e19b48a8ull, 0x391c0cb3c5c95a63ull, 0x4ed8aa4ae3418acbull, 0x5b9cca4f7763e373ull, \
     0x682e6ff3d6b2b8a3ull, 0x748f82ee5defb2fcull, 0x78a5636f43172f60ull, 0x84c87814a1f0ab72ull, 0x8cc702081a6439ecull, \
     0x90befffa23631e28ull, 0xa4506cebde82bde9ull, 0xbef9a3f7b2c67915ull, 0xc67178f2e372532bull, 0xca273eceea26619cull, \
     0xd186b8c721c0c207ull, 0xeada7dd6cde0eb1eull, 0xf57d4f7fee6ed178ull, 0x06f067aa72176fbaull, 0x0a637dc5a2c898a6ull, \
     0x113f9804bef90daeull, 0x1b710b35131c471bull, 0x28db77f523047d84ull, 0x32caab7b40c72493ull, 0x3c9ebe0a15c9bebcull, \
     0x431d67c49c100d4cull, 0x4cc5d4becb3e42b6ull, 0x597f299cfc657e2aull, 0x5fcb6fab3ad6faecull, 0x6c44198c4a475817ull}

#define MK_SHA512_IV                                                                                   \
    {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull, 0xa54ff53a5f1d36f1ull,         \
     0x510e527fade682d1ull, 0x9b05688c2b3e6c1full, 0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull}

MK_CONST u64 MK_K512[80] = MK_SHA512_K;

// One SHA-512 round t (compile-time), message schedule included. The eight working
// variables rotate through the call sites instead of being moved, and w[] is indexed
// only by constants, so the 80 rounds become straight-line code with no register
// indexing.
// Mid = true: W16 and W17 are already in w[0] and w[1] (mk_sha_mid).
template <int T, bool Mid = false>
MK_DEV void mk_sha_round(u64 w[16], u64 a, u64 b, u64 c, u64& d, u64 e, u64 f, u64 g, u64& h) {
    u64 wt;
    if (T < 16 || (Mid && T < 18)) {
        wt = w[T & 15];
    } else {
        const u64 w15 = w[(T + 1) & 15], w2 = w[(T + 14) & 15];
        wt = w[T & 15] + mk_rot2_shr(w15, 1, 8, 7) + w[(T + 9) & 15] + mk_rot2_shr(w2, 19, 61, 6);
        w[T & 15] = wt;
    }
    const u64 t1 = h + mk_rot3(e, 14, 18, 41) + mk_ch64(e, f, g) + (MK_K512[T] + wt);
    d += t1;                                                  // the new e
    h = t1 + mk_rot3(a, 28, 34, 39) + mk_maj64(a, b, c);      // the new a
}

// Eight rounds from T; afterwards the variables are back in their original roles.
template <int T, bool Mid = false>
MK_DEV void mk_sha_8rounds(u64 w[16], u64& a, u64& b, u64& c, u64& d, u64& e, u64& f, u64& g, u64& h) {
    mk_sha_round<T + 0, Mid>(w, a, b, c, d, e, f, g, h);
    mk_sha_round<T + 1, Mid>(w, h, a, b, c, d, e, f, g);
    mk_sha_round<T + 2, Mid>(w, g, h, a, b, c, d, e, f);
    mk_sha_round<T + 3, Mid>(w, f, g, h, a, b, c, d, e);
    mk_sha_round<T + 4, Mid>(w, e, f, g, h, a, b, c, d);
    mk_sha_round<T + 5, Mid>(w, d, e, f, g, h, a, b, c);
    mk_sha_round<T + 6, Mid>(w, c, d, e, f, g, h, a, b);
    mk_sha_round<T + 7, Mid>(w, b, c, d, e, f, g, h, a);
}

// Big-endian SHA-512 word of two little-endian seed words.
MK_DEV u64 mk_be_word(u32 lo_word, u32 hi_word) { return (u64)mk_bswap32(lo_word) << 32 | mk_bswap32(hi_word); }

// H0..H3 of the final state as the clamped scalar (8 LE words).
MK_DEV void mk_sha_finish(u64 a, u64 b, u64 c, u64 d, const u64 iv[8], u32 s_out[8]) {
    const u64 H[4] = {a + iv[0], b + iv[1], c + iv[2], d + iv[3]};
    MK_UNROLL
    for (int i = 0; i < 4; i++) {
        s_out[2 * i] = mk_bswap32((u32)(H[i] >> 32));
        s_out[2 * i + 1] = mk_bswap32((u32)H[i]);
    }
    s_out[0] &= ~7u;
    s_out[7] = (s_out[7] & 0x7fffffffu) | 0x40000000u;
}

// clamp(SHA-512(seed)[..32]) for a 32-byte seed given as 8 LE words; the scalar is
// returned as 8 LE words. The message is one block: the 4 seed words, 0x80 padding
// and the bit length 256, so w[5..14] are zero and fold away.
MK_DEV void mk_secret_scalar(const u32 seed[8], u32 s_out[8]) {
    const u64 iv[8] = MK_SHA512_IV;
    u64 w[16];
    MK_UNROLL
    for (int i = 0; i < 4; i++) w[i] = mk_be_word(seed[2 * i], seed[2 * i + 1]);
    w[4] = 0x8000000000000000ull;
    MK_UNROLL
    for (int i = 5; i < 15; i++) w[i] = 0;
    w[15] = 256;
    u64 a = iv[0], b = iv[1], c = iv[2], d = iv[3], e = iv[4], f = iv[5], g = iv[6], h = iv[7];
    mk_sha_8rounds<0>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<8>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<16>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<24>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<32>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<40>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<48>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<56>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<64>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<72>(w, a, b, c, d, e, f, g, h);
    mk_sha_finish(a, b, c, d, iv, s_out);
}

// Low-security (counter) seeds share their first 24 bytes, so SHA-512 words W0..W2 are
// the same for every index of a search and only W3 (bytes 24..31) varies (reference
// kernel.cu:308-362, ShaMid). The mid-state holds what depends on W0..W2 alone: the
// working variables after rounds 0-2 with round 3's constant part folded into the two
// variables it writes (so round 3 costs two additions per key), the schedule words
// W16 and W17, and W2 for W18.
struct mk_sha_mid {
    u64 v[8];   // a..h after round 2; v[0] and v[4] hold round 3's constant parts
    u64 w2, w16, w17;
};

MK_DEV void mk_sha_mid_init(const u32 key[8], mk_sha_mid& m) {
    const u64 iv[8] = MK_SHA512_IV;
    u64 w[16];
    MK_UNROLL
    for (int i = 0; i < 3; i++) w[i] = mk_be_word(key[2 * i], key[2 * i + 1]);
    w[3] = 0;  // not read by rounds 0-2
    w[4] = 0x8000000000000000ull;
    MK_UNROLL
    for (int i = 5; i < 15; i++) w[i] = 0;
    w[15] = 256;
    u64 a = iv[0], b = iv[1], c = iv[2], d = iv[3], e = iv[4], f = iv[5], g = iv[6], h = iv[7];
    mk_sha_round<0>(w, a, b, c, d, e, f, g, h);
    mk_sha_round<1>(w, h, a, b, c, d, e, f, g);
    mk_sha_round<2>(w, g, h, a, b, c, d, e, f);
    // round 3 runs as mk_sha_round<3>(w, f, g, h, a, b, c, d, e): it adds t1 = P + W3 to
    // the variable in slot a and sets slot e to t1 + T2
    const u64 P = e + mk_rot3(b, 14, 18, 41) + mk_ch64(b, c, d) + MK_K512[3];
    const u64 T2 = mk_rot3(f, 28, 34, 39) + mk_maj64(f, g, h);
    m.v[0] = a + P;
    m.v[1] = b;
    m.v[2] = c;
    m.v[3] = d;
    m.v[4] = P + T2;
    m.v[5] = f;
    m.v[6] = g;
    m.v[7] = h;
    m.w2 = w[2];
    m.w16 = w[0] + mk_rot2_shr(w[1], 1, 8, 7) + w[9] + mk_rot2_shr(w[14], 19, 61, 6);
    m.w17 = w[1] + mk_rot2_shr(w[2], 1, 8, 7) + w[10] + mk_rot2_shr(w[15], 19, 61, 6);
}

// clamp(SHA-512(seed)[..32]) of a counter seed whose bytes 24..31 are the big-endian
// word w3, from the search's mid-state. Equal to mk_secret_scalar of the full seed.
MK_DEV void mk_secret_scalar_mid(const mk_sha_mid& m, u64 w3, u32 s_out[8]) {
    const u64 iv[8] = MK_SHA512_IV;
    u64 w[16];
    // slots 0 and 1 carry W16 and W17, which rounds 16 and 17 then only read
    w[0] = m.w16;
    w[1] = m.w17;
    w[2] = m.w2;
    w[3] = w3;
    w[4] = 0x8000000000000000ull;
    MK_UNROLL
    for (int i = 5; i < 15; i++) w[i] = 0;
    w[15] = 256;
    u64 a = m.v[0] + w3, b = m.v[1], c = m.v[2], d = m.v[3], e = m.v[4] + w3, f = m.v[5], g = m.v[6], h = m.v[7];
    mk_sha_round<4, true>(w, e, f, g, h, a, b, c, d);
    mk_sha_round<5, true>(w, d, e, f, g, h, a, b, c);
    mk_sha_round<6, true>(w, c, d, e, f, g, h, a, b);
    mk_sha_round<7, true>(w, b, c, d, e, f, g, h, a);
    mk_sha_8rounds<8, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<16, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<24, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<32, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<40, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<48, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<56, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<64, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_8rounds<72, true>(w, a, b, c, d, e, f, g, h);
    mk_sha_finish(a, b, c, d, iv, s_out);
}

// Seed of index i for source 0 (ChaCha20) or 1 (counter). `upper` keeps the second
// half of the last ChaCha20 block computed, which is the seed of the odd index that
// follows an even one; `fresh` forces a new block.
MK_DEV void mk_seed_at(const u32 key[8], u32 source, u64 index, u32 upper[8], bool fresh, u32 seed[8]) {
    if (source == 1) {
        mk_counter_seed(key, index, seed);
        return;
    }
    if (fresh || (index & 1) == 0) {
        u32 block[16];
        mk_chacha20_block(key, index >> 1, block);
        const bool odd = (index & 1) != 0;
        MK_UNROLL
        for (int i = 0; i < 8; i++) {
            seed[i] = odd ? block[8 + i] : block[i];
            upper[i] = block[8 + i];
        }
        return;
    }
    MK_UNROLL
    for (int i = 0; i < 8; i++) seed[i] = upper[i];
}
