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

// 64-bit values as (hi, lo) 32-bit halves so that every rotate is two funnel shifts
// (v_alignbit_b32) and three-input XOR / majority are one v_bitop3_b32 each on gfx950.
struct mk_u64p {
    u32 lo, hi;
};

#if defined(__HIP_DEVICE_COMPILE__)
MK_DEV u32 mk_funnel(u32 hi, u32 lo, u32 n) { return __builtin_amdgcn_alignbit(hi, lo, n); }
MK_DEV u32 mk_xor3(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96); }
MK_DEV u32 mk_maj(u32 a, u32 b, u32 c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0xE8); }
#else
// (hi:lo) >> n, low 32 bits, for n in 0..31
MK_DEV u32 mk_funnel(u32 hi, u32 lo, u32 n) { return (u32)((((u64)hi << 32) | lo) >> (n & 31)); }
MK_DEV u32 mk_xor3(u32 a, u32 b, u32 c) { return a ^ b ^ c; }
MK_DEV u32 mk_maj(u32 a, u32 b, u32 c) { return (a & b) | (a & c) | (b & c); }
#endif

// rotate right by a constant n in 1..63
MK_DEV mk_u64p mk_rotr(mk_u64p x, int n) {
    mk_u64p r;
    if (n < 32) {
        r.lo = mk_funnel(x.hi, x.lo, (u32)n);
        r.hi = mk_funnel(x.lo, x.hi, (u32)n);
    } else {
        r.lo = mk_funnel(x.lo, x.hi, (u32)(n - 32));
        r.hi = mk_funnel(x.hi, x.lo, (u32)(n - 32));
    }
    return r;
}

// shift right by a constant n in 1..31
MK_DEV mk_u64p mk_shr(mk_u64p x, int n) {
    mk_u64p r;
    r.lo = mk_funnel(x.hi, x.lo, (u32)n);
    r.hi = x.hi >> n;
    return r;
}

MK_DEV mk_u64p mk_x3(mk_u64p a, mk_u64p b, mk_u64p c) {
    mk_u64p r;
    r.lo = mk_xor3(a.lo, b.lo, c.lo);
    r.hi = mk_xor3(a.hi, b.hi, c.hi);
    return r;
}

MK_DEV u64 mk_v(mk_u64p x) { return (u64)x.hi << 32 | x.lo; }

MK_DEV mk_u64p mk_p(u64 v) {
    mk_u64p r;
    r.lo = (u32)v;
    r.hi = (u32)(v >> 32);
    return r;
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
template <int T>
MK_DEV void mk_sha_round(u64 w[16], u64 a, u64 b, u64 c, u64& d, u64 e, u64 f, u64 g, u64& h) {
    u64 wt;
    if (T < 16) {
        wt = w[T];
    } else {
        const mk_u64p w15 = mk_p(w[(T + 1) & 15]), w2 = mk_p(w[(T + 14) & 15]);
        const u64 s0 = mk_v(mk_x3(mk_rotr(w15, 1), mk_rotr(w15, 8), mk_shr(w15, 7)));
        const u64 s1 = mk_v(mk_x3(mk_rotr(w2, 19), mk_rotr(w2, 61), mk_shr(w2, 6)));
        wt = w[T & 15] + s0 + w[(T + 9) & 15] + s1;
        w[T & 15] = wt;
    }
    const mk_u64p ep = mk_p(e), ap = mk_p(a), bp = mk_p(b), cp = mk_p(c);
    const u64 S1 = mk_v(mk_x3(mk_rotr(ep, 14), mk_rotr(ep, 18), mk_rotr(ep, 41)));
    const u64 ch = (e & f) ^ (~e & g);
    const u64 t1 = h + S1 + ch + (MK_K512[T] + wt);
    const u64 S0 = mk_v(mk_x3(mk_rotr(ap, 28), mk_rotr(ap, 34), mk_rotr(ap, 39)));
    mk_u64p mj;
    mj.lo = mk_maj(ap.lo, bp.lo, cp.lo);
    mj.hi = mk_maj(ap.hi, bp.hi, cp.hi);
    d += t1;                // the new e
    h = t1 + S0 + mk_v(mj);  // the new a
}

// Eight rounds from T; afterwards the variables are back in their original roles.
template <int T>
MK_DEV void mk_sha_8rounds(u64 w[16], u64& a, u64& b, u64& c, u64& d, u64& e, u64& f, u64& g, u64& h) {
    mk_sha_round<T + 0>(w, a, b, c, d, e, f, g, h);
    mk_sha_round<T + 1>(w, h, a, b, c, d, e, f, g);
    mk_sha_round<T + 2>(w, g, h, a, b, c, d, e, f);
    mk_sha_round<T + 3>(w, f, g, h, a, b, c, d, e);
    mk_sha_round<T + 4>(w, e, f, g, h, a, b, c, d);
    mk_sha_round<T + 5>(w, d, e, f, g, h, a, b, c);
    mk_sha_round<T + 6>(w, c, d, e, f, g, h, a, b);
    mk_sha_round<T + 7>(w, b, c, d, e, f, g, h, a);
}

// clamp(SHA-512(seed)[..32]) for a 32-byte seed given as 8 LE words; the scalar is
// returned as 8 LE words. The message is one block: the 4 seed words, 0x80 padding
// and the bit length 256, so w[5..14] are zero and fold away.
MK_DEV void mk_secret_scalar(const u32 seed[8], u32 s_out[8]) {
    const u64 iv[8] = MK_SHA512_IV;
    u64 w[16];
    MK_UNROLL
    for (int i = 0; i < 4; i++) w[i] = (u64)mk_bswap32(seed[2 * i]) << 32 | mk_bswap32(seed[2 * i + 1]);
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
    const u64 H[4] = {a + iv[0], b + iv[1], c + iv[2], d + iv[3]};
    MK_UNROLL
    for (int i = 0; i < 4; i++) {
        s_out[2 * i] = mk_bswap32((u32)(H[i] >> 32));
        s_out[2 * i + 1] = mk_bswap32((u32)H[i]);
    }
    s_out[0] &= ~7u;
    s_out[7] = (s_out[7] & 0x7fffffffu) | 0x40000000u;
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
