// One instruction set's engine of the SIMD CPU baseline: 256 keys per call, 8 keys in the
// 8 lanes of each vector. meikey_simd.cpp includes this file twice, inside GCC target
// regions: MKS_ISA 512 (AVX-512F + IFMA, namespace mks512) and MKS_ISA 2 (AVX2, mks2).
// Test and measurement infrastructure only.
//
// A batch of 256 keys, in the order of the reference's batch path
// (src/service/derive/avx512.rs:423-543):
//  1. seeds (ChaCha20 blocks 8 at a time, or counter seeds), SHA-512 x 8 and clamping;
//  2. the W = 12 signed digits of every key;
//  3. row by row over the 32 groups of 8 keys: row 0 is the table entry itself in
//     extended coordinates, rows 1-19 are mixed additions (7 multiplications), and the
//     last row stops at the fraction form y = H/F, x = E/G (3 multiplications);
//  4. one inversion of the product of the 32 groups' F (per lane), then 1/F and y per
//     group, back to front;
//  5. unsigned-first matching per key; only a maybe-hit has its x sign computed.

#if MKS_ISA == 512
#define MKS_NS mks512
#else
#define MKS_NS mks2
#endif

namespace MKS_NS {

typedef u32 v8u32 __attribute__((vector_size(32)));

#if MKS_ISA == 512
// 8 lanes of u64: one zmm register.
typedef u64 v8 __attribute__((vector_size(64)));

static inline v8 splat(u64 x) {
    v8 r = {x, x, x, x, x, x, x, x};
    return r;
}
static inline u64 lane(const v8& v, int l) { return v[l]; }
static inline void set_lane(v8& v, int l, u64 x) { v[l] = x; }
#else
// 8 lanes of u64 as two ymm registers. (A 64-byte vector type has no register of its own
// under AVX2, and GCC then keeps it in memory.)
typedef u64 v4 __attribute__((vector_size(32)));
struct v8 {
    v4 h[2];
};
static inline v8 splat(u64 x) {
    const v4 q = {x, x, x, x};
    return v8{{q, q}};
}
static inline u64 lane(const v8& v, int l) { return v.h[l >> 2][l & 3]; }
static inline void set_lane(v8& v, int l, u64 x) { v.h[l >> 2][l & 3] = x; }
#define MKS_BINOP(op)                                                                          \
    static inline v8 operator op(const v8& a, const v8& b) { return v8{{a.h[0] op b.h[0], a.h[1] op b.h[1]}}; } \
    static inline v8 operator op(const v8& a, u64 b) { return v8{{a.h[0] op b, a.h[1] op b}}; }              \
    static inline v8 operator op(u64 a, const v8& b) { return v8{{a op b.h[0], a op b.h[1]}}; }              \
    static inline v8& operator op##=(v8& a, const v8& b) { a = a op b; return a; }                           \
    static inline v8& operator op##=(v8& a, u64 b) { a = a op b; return a; }
MKS_BINOP(+)
MKS_BINOP(-)
MKS_BINOP(&)
MKS_BINOP(|)
MKS_BINOP(^)
#undef MKS_BINOP
static inline v8 operator~(const v8& a) { return v8{{~a.h[0], ~a.h[1]}}; }
static inline v8 operator<<(const v8& a, int n) { return v8{{a.h[0] << n, a.h[1] << n}}; }
static inline v8 operator>>(const v8& a, int n) { return v8{{a.h[0] >> n, a.h[1] >> n}}; }
#endif

// ------------------------------------------------------------------ seeds

#define MKS_QR(a, b, c, d)                      \
    a += b; d ^= a; d = (d << 16) | (d >> 16); \
    c += d; b ^= c; b = (b << 12) | (b >> 20); \
    a += b; d ^= a; d = (d << 8) | (d >> 24);  \
    c += d; b ^= c; b = (b << 7) | (b >> 25);

// ChaCha20 blocks ctr[0..7] (RFC 8439 rounds, 64-bit counter in words 12-13, zero stream
// id): out[b] = the 16 words of block ctr[b].
static void chacha8(const u32 key[8], const u64 ctr[8], u32 out[][16]) {
    static const u32 sigma[4] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    v8u32 s[16], x[16];
    for (int i = 0; i < 4; i++) s[i] = (v8u32){} + sigma[i];
    for (int i = 0; i < 8; i++) s[4 + i] = (v8u32){} + key[i];
    for (int b = 0; b < 8; b++) {
        s[12][b] = (u32)ctr[b];
        s[13][b] = (u32)(ctr[b] >> 32);
    }
    s[14] = (v8u32){};
    s[15] = (v8u32){};
    for (int i = 0; i < 16; i++) x[i] = s[i];
    for (int r = 0; r < 10; r++) {
        MKS_QR(x[0], x[4], x[8], x[12]) MKS_QR(x[1], x[5], x[9], x[13])
        MKS_QR(x[2], x[6], x[10], x[14]) MKS_QR(x[3], x[7], x[11], x[15])
        MKS_QR(x[0], x[5], x[10], x[15]) MKS_QR(x[1], x[6], x[11], x[12])
        MKS_QR(x[2], x[7], x[8], x[13]) MKS_QR(x[3], x[4], x[9], x[14])
    }
    for (int i = 0; i < 16; i++) {
        x[i] += s[i];
        for (int b = 0; b < 8; b++) out[b][i] = x[i][b];
    }
}

// Seeds of indices first .. first + 255. Source 0: index i is half (i & 1) of ChaCha20
// block i >> 1 (src/service/gpu.rs:1344-1351). Source 1: the key with its bytes 24..31,
// read as a little-endian u64, plus i (src/model/search.rs:45-50).
static void seeds256(const u8 key[32], int source, u64 first, u8 seeds[MKS_BATCH][32]) {
    if (source == 1) {
        u64 tail;
        memcpy(&tail, key + 24, 8);
        for (int j = 0; j < MKS_BATCH; j++) {
            memcpy(seeds[j], key, 24);
            const u64 t = tail + first + (u64)j;
            memcpy(seeds[j] + 24, &t, 8);
        }
        return;
    }
    // the distinct blocks of the 256 indices in order (the index wraps at 2^64), 8 at a time
    u32 kw[8];
    memcpy(kw, key, 32);
    u64 ctr[MKS_BATCH / 2 + 16];
    u8 which[MKS_BATCH];
    int nb = 0;
    for (int j = 0; j < MKS_BATCH; j++) {
        const u64 b = (first + (u64)j) >> 1;
        if (nb == 0 || ctr[nb - 1] != b) ctr[nb++] = b;
        which[j] = (u8)(nb - 1);
    }
    const int padded = (nb + 7) & ~7;
    for (int k = nb; k < padded; k++) ctr[k] = ctr[nb - 1];
    u32 blk[MKS_BATCH / 2 + 16][16];
    for (int c = 0; c < padded; c += 8) chacha8(kw, ctr + c, blk + c);
    for (int j = 0; j < MKS_BATCH; j++) memcpy(seeds[j], &blk[which[j]][8 * ((first + (u64)j) & 1)], 32);
}

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
static const u64 IV512[8] = {0x6a09e667f3bcc908ull, 0xbb67ae8584caa73bull, 0x3c6ef372fe94f82bull,
                             0xa54ff53a5f1d36f1ull, 0x510e527fade682d1ull, 0x9b05688c2b3e6c1full,
                             0x1f83d9abfb41bd6bull, 0x5be0cd19137e2179ull};

static inline v8 rotr(v8 x, int n) { return (x >> n) | (x << (64 - n)); }

// Clamped secret scalars (RFC 8032 5.1.5) of 8 seeds: the first half of SHA-512 of each
// 32-byte seed, one seed per lane (the message is one padded block).
static void sha512_clamp8(const u8 (*seeds)[32], u8 (*out)[32]) {
    v8 w[16];
    for (int i = 0; i < 4; i++)
        for (int l = 0; l < 8; l++) {
            u64 v;
            memcpy(&v, seeds[l] + 8 * i, 8);
            set_lane(w[i], l, __builtin_bswap64(v));
        }
    w[4] = splat(0x8000000000000000ull);
    for (int i = 5; i < 15; i++) w[i] = splat(0);
    w[15] = splat(256);
    v8 a = splat(IV512[0]), b = splat(IV512[1]), c = splat(IV512[2]), d = splat(IV512[3]);
    v8 e = splat(IV512[4]), f = splat(IV512[5]), g = splat(IV512[6]), h = splat(IV512[7]);
    for (int t = 0; t < 80; t++) {
        if (t >= 16) {
            const v8 w2 = w[(t - 2) & 15], w15 = w[(t - 15) & 15];
            w[t & 15] += (rotr(w15, 1) ^ rotr(w15, 8) ^ (w15 >> 7)) + w[(t - 7) & 15] +
                         (rotr(w2, 19) ^ rotr(w2, 61) ^ (w2 >> 6));
        }
        const v8 t1 = h + (rotr(e, 14) ^ rotr(e, 18) ^ rotr(e, 41)) + ((e & f) ^ (~e & g)) + K512[t] + w[t & 15];
        const v8 t2 = (rotr(a, 28) ^ rotr(a, 34) ^ rotr(a, 39)) + ((a & b) ^ (a & c) ^ (b & c));
        h = g;
        g = f;
        f = e;
        e = d + t1;
        d = c;
        c = b;
        b = a;
        a = t1 + t2;
    }
    const v8 hs[4] = {a + IV512[0], b + IV512[1], c + IV512[2], d + IV512[3]};
    for (int l = 0; l < 8; l++) {
        for (int i = 0; i < 4; i++) {
            const u64 v = __builtin_bswap64(lane(hs[i], l));
            memcpy(out[l] + 8 * i, &v, 8);
        }
        out[l][0] &= 248;
        out[l][31] &= 127;
        out[l][31] |= 64;
    }
}

// ------------------------------------------------------------------ field

#if MKS_ISA == 512
// GF(2^255 - 19) in five 51-bit limbs per lane; products by AVX-512 IFMA, whose 52-bit
// multiply-accumulate reads the low 52 bits of each operand: every multiplication input
// has limbs below 2^52 (sums and differences are carried before they are multiplied).
#define NL 5
struct fe {
    v8 v[NL];
};
static const u64 MASK = (1ull << 51) - 1;
static const u64 P2[NL] = {2 * ((1ull << 51) - 19), 2 * MASK, 2 * MASK, 2 * MASK, 2 * MASK};  // 2p
static const int LIMB_BITS[NL] = {51, 51, 51, 51, 51};

static inline v8 madd_lo(v8 acc, v8 a, v8 b) {
    return (v8)_mm512_madd52lo_epu64((__m512i)acc, (__m512i)a, (__m512i)b);
}
static inline v8 madd_hi(v8 acc, v8 a, v8 b) {
    return (v8)_mm512_madd52hi_epu64((__m512i)acc, (__m512i)a, (__m512i)b);
}
static inline v8 times19(v8 x) { return (x << 4) + (x << 1) + x; }

// Limbs below 2^51 except limb 1 (at most 2^51); inputs below 2^62 per limb.
static inline fe fe_carry(fe r) {
    for (int k = 0; k < NL - 1; k++) {
        r.v[k + 1] += r.v[k] >> 51;
        r.v[k] &= MASK;
    }
    const v8 top = r.v[NL - 1] >> 51;
    r.v[NL - 1] &= MASK;
    r.v[0] += times19(top);
    r.v[1] += r.v[0] >> 51;
    r.v[0] &= MASK;
    return r;
}

static inline fe fe_mul(const fe& a, const fe& b) {
    // a_i b_j = lo + hi 2^52 at weight 2^(51 (i + j)): lo in column i + j, 2 hi in the
    // next; columns 5..9 fold by 2^255 = 19
    v8 lo[2 * NL], hi[2 * NL];
    for (int k = 0; k < 2 * NL; k++) lo[k] = hi[k] = splat(0);
    for (int i = 0; i < NL; i++)
        for (int j = 0; j < NL; j++) {
            lo[i + j] = madd_lo(lo[i + j], a.v[i], b.v[j]);
            hi[i + j] = madd_hi(hi[i + j], a.v[i], b.v[j]);
        }
    v8 c[2 * NL];
    c[0] = lo[0];
    for (int k = 1; k < 2 * NL; k++) c[k] = lo[k] + (hi[k - 1] << 1);
    fe r;
    for (int k = 0; k < NL; k++) r.v[k] = c[k] + times19(c[k + NL]);
    return fe_carry(r);
}
#else
// GF(2^255 - 19) in ten limbs of 26 and 25 bits (radix 2^25.5) per lane; products by
// 32 x 32 -> 64-bit multiplies (vpmuludq), two AVX2 registers per 8 lanes.
#define NL 10
struct fe {
    v8 v[NL];
};
static const u64 M26 = (1ull << 26) - 1, M25 = (1ull << 25) - 1;
static const u64 P2[NL] = {2 * ((1ull << 26) - 19), 2 * M25, 2 * M26, 2 * M25, 2 * M26,
                           2 * M25, 2 * M26, 2 * M25, 2 * M26, 2 * M25};
static const int LIMB_BITS[NL] = {26, 25, 26, 25, 26, 25, 26, 25, 26, 25};

static inline v8 mul32(const v8& a, const v8& b) {
    return v8{{(v4)_mm256_mul_epu32((__m256i)a.h[0], (__m256i)b.h[0]),
               (v4)_mm256_mul_epu32((__m256i)a.h[1], (__m256i)b.h[1])}};
}

// Limbs below 2^26 / 2^25 except limb 1 (below 2^25 + 2^16); inputs below 2^63.
static inline fe fe_carry(fe h) {
    v8 c;
#define MKS_CARRY(i)                          \
    c = h.v[i] >> LIMB_BITS[i];               \
    h.v[(i) + 1] += c;                        \
    h.v[i] &= (1ull << LIMB_BITS[i]) - 1;
    MKS_CARRY(0) MKS_CARRY(4) MKS_CARRY(1) MKS_CARRY(5) MKS_CARRY(2) MKS_CARRY(6) MKS_CARRY(3) MKS_CARRY(7)
    MKS_CARRY(4) MKS_CARRY(8)
#undef MKS_CARRY
    c = h.v[9] >> 25;
    h.v[9] &= M25;
    h.v[0] += (c << 4) + (c << 1) + c;
    c = h.v[0] >> 26;
    h.v[1] += c;
    h.v[0] &= M26;
    return h;
}

static inline fe fe_mul(const fe& f, const fe& g) {
    // limb i has weight 2^ceil(25.5 i): f_i g_j lands on limb i + j (times 2 when i and j
    // are both odd), limbs 10..18 fold by 2^255 = 19
    v8 g19[NL], f2[NL], h[NL];
    for (int j = 0; j < NL; j++) g19[j] = (g.v[j] << 4) + (g.v[j] << 1) + g.v[j];
    for (int i = 0; i < NL; i++) f2[i] = (i & 1) ? f.v[i] << 1 : f.v[i];
    for (int k = 0; k < NL; k++) h[k] = splat(0);
    for (int i = 0; i < NL; i++)
        for (int j = 0; j < NL; j++) {
            const int k = i + j;
            const v8 fi = (j & 1) ? f2[i] : f.v[i];
            if (k < NL)
                h[k] += mul32(fi, g.v[j]);
            else
                h[k - NL] += mul32(fi, g19[j]);
        }
    fe r;
    for (int k = 0; k < NL; k++) r.v[k] = h[k];
    return fe_carry(r);
}
#endif

static inline fe fe_small(u64 x) {
    fe r;
    r.v[0] = splat(x);
    for (int k = 1; k < NL; k++) r.v[k] = splat(0);
    return r;
}

static inline fe fe_add(const fe& a, const fe& b) {
    fe r;
    for (int k = 0; k < NL; k++) r.v[k] = a.v[k] + b.v[k];
    return fe_carry(r);
}

// a - b + 2p (b carried, so every limb stays non-negative).
static inline fe fe_sub(const fe& a, const fe& b) {
    fe r;
    for (int k = 0; k < NL; k++) r.v[k] = a.v[k] + P2[k] - b.v[k];
    return fe_carry(r);
}

static inline fe fe_sq(const fe& a) { return fe_mul(a, a); }

// a^(p - 2): 254 squarings and 11 multiplications (src/service/derive/field.rs:108-124).
static fe fe_invert(const fe& a) {
    auto pow2k = [](fe x, int k) {
        for (int i = 0; i < k; i++) x = fe_sq(x);
        return x;
    };
    const fe a2 = fe_sq(a);
    const fe a9 = fe_mul(fe_sq(fe_sq(a2)), a);
    const fe a11 = fe_mul(a9, a2);
    const fe a_5_0 = fe_mul(fe_sq(a11), a9);
    const fe a_10_0 = fe_mul(pow2k(a_5_0, 5), a_5_0);
    const fe a_20_0 = fe_mul(pow2k(a_10_0, 10), a_10_0);
    const fe a_40_0 = fe_mul(pow2k(a_20_0, 20), a_20_0);
    const fe a_50_0 = fe_mul(pow2k(a_40_0, 10), a_10_0);
    const fe a_100_0 = fe_mul(pow2k(a_50_0, 50), a_50_0);
    const fe a_200_0 = fe_mul(pow2k(a_100_0, 100), a_100_0);
    const fe a_250_0 = fe_mul(pow2k(a_200_0, 50), a_50_0);
    return fe_mul(pow2k(a_250_0, 5), a11);
}

// The canonical value (below p) of each lane as four little-endian 64-bit words:
// w[q][lane]. Input: carried limbs.
static void fe_canon(fe h, v8 w[4]) {
    // three carry passes leave every limb below its radix and the value below 2^255
    for (int pass = 0; pass < 3; pass++) {
        for (int k = 0; k < NL - 1; k++) {
            h.v[k + 1] += h.v[k] >> LIMB_BITS[k];
            h.v[k] &= (1ull << LIMB_BITS[k]) - 1;
        }
        const v8 c = h.v[NL - 1] >> LIMB_BITS[NL - 1];
        h.v[NL - 1] &= (1ull << LIMB_BITS[NL - 1]) - 1;
        h.v[0] += (c << 4) + (c << 1) + c;
    }
    // q = 1 exactly when the value is at least p (value + 19 reaches 2^255)
    v8 q = (h.v[0] + 19) >> LIMB_BITS[0];
    for (int k = 1; k < NL; k++) q = (h.v[k] + q) >> LIMB_BITS[k];
    h.v[0] += (q << 4) + (q << 1) + q;
    for (int k = 0; k < NL - 1; k++) {
        h.v[k + 1] += h.v[k] >> LIMB_BITS[k];
        h.v[k] &= (1ull << LIMB_BITS[k]) - 1;
    }
    h.v[NL - 1] &= (1ull << LIMB_BITS[NL - 1]) - 1;
    // pack the limbs into 64-bit words
    for (int q2 = 0; q2 < 4; q2++) w[q2] = splat(0);
    int off = 0;
    for (int k = 0; k < NL; k++) {
        const int word = off >> 6, sh = off & 63;
        w[word] |= h.v[k] << sh;
        if (sh + LIMB_BITS[k] > 64 && word + 1 < 4) w[word + 1] |= h.v[k] >> (64 - sh);
        off += LIMB_BITS[k];
    }
}

// Limbs (u64 each) of a canonical little-endian 256-bit value.
static void limbs_of(const u8 b[32], u64 out[NL]) {
    u64 w[5];
    memcpy(w, b, 32);
    w[4] = 0;
    int off = 0;
    for (int k = 0; k < NL; k++) {
        const int word = off >> 6, sh = off & 63;
        u64 v = w[word] >> sh;
        if (sh + LIMB_BITS[k] > 64) v |= w[word + 1] << (64 - sh);
        out[k] = v & ((1ull << LIMB_BITS[k]) - 1);
        off += LIMB_BITS[k];
    }
}

// ------------------------------------------------------------------ comb table

// Entries of W = 12 (21 rows of 2049): y + x, y - x, 2dxy.
#if MKS_ISA == 512
#define ENTRY_WORDS 16  // 15 u64 limbs and one spare: 128 bytes
typedef u64 entry_word;
#else
#define ENTRY_WORDS 32  // 30 u32 limbs and two spare: 128 bytes
typedef u32 entry_word;
#endif

static entry_word* g_table = nullptr;

static int load_table(const u8* niels) {
    const size_t n = (size_t)MKS_ROWS * MKS_COLS;
    entry_word* t = (entry_word*)aligned_alloc(64, n * ENTRY_WORDS * sizeof(entry_word));
    if (!t) return -1;
    memset(t, 0, n * ENTRY_WORDS * sizeof(entry_word));
    for (size_t e = 0; e < n; e++)
        for (int f = 0; f < 3; f++) {
            u64 l[NL];
            limbs_of(niels + 96 * e + 32 * f, l);
            for (int k = 0; k < NL; k++) t[e * ENTRY_WORDS + NL * f + k] = (entry_word)l[k];
        }
    free(g_table);
    g_table = t;
    return 0;
}

struct niels8 {
    fe ypx, ymx, t2d;
};

#if MKS_ISA == 512
// 8 x 8 transpose of 64-bit words: out[q] lane k = in[k] word q.
static inline void transpose8(const __m512i in[8], __m512i out[8]) {
    __m512i a[8], b[8];
    for (int i = 0; i < 8; i += 2) {
        a[i] = _mm512_unpacklo_epi64(in[i], in[i + 1]);      // words 0 2 4 6 of rows i, i + 1
        a[i + 1] = _mm512_unpackhi_epi64(in[i], in[i + 1]);  // words 1 3 5 7
    }
    for (int h = 0; h < 8; h += 4) {
        b[h + 0] = _mm512_shuffle_i64x2(a[h], a[h + 2], 0x88);      // words 0, 4
        b[h + 1] = _mm512_shuffle_i64x2(a[h + 1], a[h + 3], 0x88);  // words 1, 5
        b[h + 2] = _mm512_shuffle_i64x2(a[h], a[h + 2], 0xdd);      // words 2, 6
        b[h + 3] = _mm512_shuffle_i64x2(a[h + 1], a[h + 3], 0xdd);  // words 3, 7
    }
    for (int q = 0; q < 4; q++) {
        out[q] = _mm512_shuffle_i64x2(b[q], b[4 + q], 0x88);
        out[q + 4] = _mm512_shuffle_i64x2(b[q], b[4 + q], 0xdd);
    }
}

// The entries of row `row` at |d[k]| for the 8 lanes, negated where d[k] < 0
// (-(x, y) = (-x, y): y + x and y - x swap, 2dxy changes sign).
static inline niels8 gather(int row, const int16_t d[8]) {
    const u64* base = g_table + (size_t)row * MKS_COLS * ENTRY_WORDS;
    __m512i in[8], t[16];
    __mmask8 neg = 0;
    for (int k = 0; k < 8; k++) {
        const int a = d[k] < 0 ? -d[k] : d[k];
        neg |= (__mmask8)((d[k] < 0) << k);
        in[k] = _mm512_load_si512((const void*)(base + (size_t)a * ENTRY_WORDS));
    }
    transpose8(in, t);
    for (int k = 0; k < 8; k++) {
        const int a = d[k] < 0 ? -d[k] : d[k];
        in[k] = _mm512_load_si512((const void*)(base + (size_t)a * ENTRY_WORDS + 8));
    }
    transpose8(in, t + 8);
    niels8 n;
    for (int k = 0; k < NL; k++) {
        const __m512i ypx = t[k], ymx = t[NL + k], t2d = t[2 * NL + k];
        n.ypx.v[k] = (v8)_mm512_mask_blend_epi64(neg, ypx, ymx);
        n.ymx.v[k] = (v8)_mm512_mask_blend_epi64(neg, ymx, ypx);
        n.t2d.v[k] = (v8)_mm512_mask_sub_epi64(t2d, neg, _mm512_set1_epi64((long long)P2[k]), t2d);
    }
    return n;
}

static inline void prefetch_entries(int row, const int16_t d[8]) {
    const u64* base = g_table + (size_t)row * MKS_COLS * ENTRY_WORDS;
    for (int k = 0; k < 8; k++) {
        const int a = d[k] < 0 ? -d[k] : d[k];
        __builtin_prefetch(base + (size_t)a * ENTRY_WORDS);
        __builtin_prefetch(base + (size_t)a * ENTRY_WORDS + 8);
    }
}
#else
static inline niels8 gather(int row, const int16_t d[8]) {
    const u32* base = g_table + (size_t)row * MKS_COLS * ENTRY_WORDS;
    alignas(32) int idx[8];
    alignas(64) u64 negm[8];
    for (int k = 0; k < 8; k++) {
        idx[k] = (d[k] < 0 ? -d[k] : d[k]) * ENTRY_WORDS;
        negm[k] = d[k] < 0 ? ~0ull : 0ull;
    }
    const __m256i vidx = _mm256_load_si256((const __m256i*)idx);
    v8 neg;
    memcpy(&neg, negm, 64);
    niels8 n;
    v8 w[3 * NL];
    for (int q = 0; q < 3 * NL; q++) {
        const __m256i g = _mm256_i32gather_epi32((const int*)(base + q), vidx, 4);
        w[q] = v8{{(v4)_mm256_cvtepu32_epi64(_mm256_castsi256_si128(g)),
                   (v4)_mm256_cvtepu32_epi64(_mm256_extracti128_si256(g, 1))}};
    }
    for (int k = 0; k < NL; k++) {
        const v8 ypx = w[k], ymx = w[NL + k], t2d = w[2 * NL + k];
        n.ypx.v[k] = (ypx & ~neg) | (ymx & neg);
        n.ymx.v[k] = (ymx & ~neg) | (ypx & neg);
        n.t2d.v[k] = (t2d & ~neg) | ((P2[k] - t2d) & neg);
    }
    return n;
}

static inline void prefetch_entries(int row, const int16_t d[8]) {
    const u32* base = g_table + (size_t)row * MKS_COLS * ENTRY_WORDS;
    for (int k = 0; k < 8; k++) {
        const int a = d[k] < 0 ? -d[k] : d[k];
        __builtin_prefetch(base + (size_t)a * ENTRY_WORDS);
        __builtin_prefetch(base + (size_t)a * ENTRY_WORDS + 16);
    }
}
#endif

// ------------------------------------------------------------------ batches

struct ext {
    fe x, y, z, t;
};

// Per-thread state of one batch of 256 keys (MKS_GROUPS groups of 8 lanes).
struct state {
    ext acc[MKS_GROUPS];  // rows 0..19: extended X, Y, Z, T; after the last row: E, H, F, G
    fe pre[MKS_GROUPS];   // prefix products of F
    fe ginv[MKS_GROUPS];  // 1/G where needed
    int ginv_ok[MKS_GROUPS];
    v8 yw[MKS_GROUPS][4];  // canonical y (sign bit clear)
    int16_t digits[MKS_ROWS][MKS_BATCH];
    u8 seeds[MKS_BATCH][32];
    u8 scalars[MKS_BATCH][32];
};

static state* new_state() {
    state* s = (state*)aligned_alloc(64, (sizeof(state) + 63) / 64 * 64);
    return s;
}

// Row 0 (the identity plus the entry): E = (y+x) - (y-x) = 2x, H = 2y, so
// (X, Y, Z, T) = (2E, 2H, 4, EH) -- one multiplication.
static inline void first_row(ext& p, const niels8& n) {
    const fe e = fe_sub(n.ypx, n.ymx), h = fe_add(n.ypx, n.ymx);
    p.x = fe_add(e, e);
    p.y = fe_add(h, h);
    p.z = fe_small(4);
    p.t = fe_mul(e, h);
}

// p + n (extended + affine Niels, Hisil-Wong-Carter-Dawson 2008, 7 multiplications).
static inline void add_row(ext& p, const niels8& n) {
    const fe a = fe_mul(fe_sub(p.y, p.x), n.ymx);
    const fe b = fe_mul(fe_add(p.y, p.x), n.ypx);
    const fe c = fe_mul(p.t, n.t2d);
    const fe d = fe_add(p.z, p.z);
    const fe e = fe_sub(b, a), f = fe_sub(d, c), g = fe_add(d, c), h = fe_add(b, a);
    p.x = fe_mul(e, f);
    p.y = fe_mul(g, h);
    p.z = fe_mul(f, g);
    p.t = fe_mul(e, h);
}

// The last row in fraction form (3 multiplications): y = H/F, x = E/G, kept as
// (E, H, F, G) in (x, y, z, t).
static inline void last_row(ext& p, const niels8& n) {
    const fe a = fe_mul(fe_sub(p.y, p.x), n.ymx);
    const fe b = fe_mul(fe_add(p.y, p.x), n.ypx);
    const fe c = fe_mul(p.t, n.t2d);
    const fe d = fe_add(p.z, p.z);
    p.x = fe_sub(b, a);
    p.y = fe_add(b, a);
    p.z = fe_sub(d, c);
    p.t = fe_add(d, c);
}

// Signed base-2^12 digits of m, where the clamped s = 8 (2^251 + m): digits in
// [-2^11, 2^11), the top one in [0, 2^11] (src/service/derive/comb.rs:414-445).
static void digits12(const u8 s[32], int16_t* out, int stride) {
    u64 w[5];
    memcpy(w, s, 32);
    w[4] = 0;
    for (int i = 0; i < 4; i++) w[i] = w[i] >> 3 | w[i + 1] << 61;
    w[3] &= (1ull << (251 - 192)) - 1;
    int carry = 0;
    for (int i = 0; i < MKS_ROWS; i++) {
        const int off = i * MKS_W, q = off >> 6, r = off & 63;
        u64 bits = w[q] >> r;
        if (r + MKS_W > 64) bits |= w[q + 1] << (64 - r);
        const int v = (int)(bits & ((1u << MKS_W) - 1)) + carry;
        if (i == MKS_ROWS - 1) {
            out[i * stride] = (int16_t)v;
        } else {
            carry = (v + (1 << (MKS_W - 1))) >> MKS_W;
            out[i * stride] = (int16_t)(v - (carry << MKS_W));
        }
    }
}

// Points of keys first .. first + 255, up to the fraction form of the last row, and y.
static void points256(state* st, const u8 key[32], int source, u64 first) {
    seeds256(key, source, first, st->seeds);
    for (int g = 0; g < MKS_GROUPS; g++) sha512_clamp8(st->seeds + 8 * g, st->scalars + 8 * g);
    for (int j = 0; j < MKS_BATCH; j++) digits12(st->scalars[j], &st->digits[0][j], MKS_BATCH);
    // row by row: one row of the table (256 KB) serves all 32 groups
    for (int g = 0; g < MKS_GROUPS; g++) {
        if (g + 1 < MKS_GROUPS) prefetch_entries(0, &st->digits[0][8 * (g + 1)]);
        first_row(st->acc[g], gather(0, &st->digits[0][8 * g]));
    }
    for (int row = 1; row < MKS_ROWS - 1; row++)
        for (int g = 0; g < MKS_GROUPS; g++) {
            if (g + 1 < MKS_GROUPS)
                prefetch_entries(row, &st->digits[row][8 * (g + 1)]);
            else
                prefetch_entries(row + 1, &st->digits[row + 1][0]);
            add_row(st->acc[g], gather(row, &st->digits[row][8 * g]));
        }
    for (int g = 0; g < MKS_GROUPS; g++) {
        if (g + 1 < MKS_GROUPS) prefetch_entries(MKS_ROWS - 1, &st->digits[MKS_ROWS - 1][8 * (g + 1)]);
        last_row(st->acc[g], gather(MKS_ROWS - 1, &st->digits[MKS_ROWS - 1][8 * g]));
    }
    // Montgomery's batch inversion of F over the groups, per lane
    for (int g = 0; g < MKS_GROUPS; g++) st->pre[g] = g == 0 ? st->acc[0].z : fe_mul(st->pre[g - 1], st->acc[g].z);
    fe inv = fe_invert(st->pre[MKS_GROUPS - 1]);
    for (int g = MKS_GROUPS - 1; g >= 0; g--) {
        const fe finv = g == 0 ? inv : fe_mul(inv, st->pre[g - 1]);
        if (g > 0) inv = fe_mul(inv, st->acc[g].z);
        fe_canon(fe_mul(st->acc[g].y, finv), st->yw[g]);
        st->ginv_ok[g] = 0;
    }
}

// Bit 0 of the canonical x = E/G of key j.
static int x_sign(state* st, int j) {
    const int g = j >> 3, l = j & 7;
    if (!st->ginv_ok[g]) {
        st->ginv[g] = fe_invert(st->acc[g].t);
        st->ginv_ok[g] = 1;
    }
    v8 xw[4];
    fe_canon(fe_mul(st->acc[g].x, st->ginv[g]), xw);
    return (int)(lane(xw[0], l) & 1);
}

static inline void y_bytes(const state* st, int j, u8 pk[32]) {
    for (int q = 0; q < 4; q++) {
        const u64 v = lane(st->yw[j >> 3][q], j & 7);
        memcpy(pk + 8 * q, &v, 8);
    }
}

// Public keys of keys first .. first + n - 1 (n <= 256), x sign included.
static void derive(state* st, const u8 key[32], int source, u64 first, u32 n, u8* pks) {
    points256(st, key, source, first);
    // 1/G of every group by one more batch inversion
    for (int g = 0; g < MKS_GROUPS; g++) st->pre[g] = g == 0 ? st->acc[0].t : fe_mul(st->pre[g - 1], st->acc[g].t);
    fe inv = fe_invert(st->pre[MKS_GROUPS - 1]);
    for (int g = MKS_GROUPS - 1; g >= 0; g--) {
        st->ginv[g] = g == 0 ? inv : fe_mul(inv, st->pre[g - 1]);
        if (g > 0) inv = fe_mul(inv, st->acc[g].t);
        st->ginv_ok[g] = 1;
    }
    for (int g = 0; g < MKS_GROUPS; g++) {
        v8 xw[4];
        fe_canon(fe_mul(st->acc[g].x, st->ginv[g]), xw);
        for (int l = 0; l < 8; l++) {
            const u32 j = 8 * g + l;
            if (j >= n) return;
            y_bytes(st, (int)j, pks + 32 * j);
            pks[32 * j + 31] |= (u8)((lane(xw[0], l) & 1) << 7);
        }
    }
}

// Exact hits among keys first .. first + n - 1 (n <= 256), unsigned first: the key text
// with a zero sign bit is matched with either sign (one character differs), and only a
// maybe-hit has its sign computed and is matched exactly (src/service/search.rs:273-288,
// src/service/matcher.rs:99-129).
static u32 search(state* st, const u8 key[32], int source, u64 first, u32 n, const matcher& m, u64* hits,
                  u32 cap) {
    points256(st, key, source, first);
    u32 found = 0;
    for (u32 j = 0; j < n; j++) {
        u8 pk[32];
        char text[MKS_B64_LEN];
        y_bytes(st, (int)j, pk);
        const int v66 = encode_tail(pk, text);
        bool maybe = hay_match(text, m);
        if (!maybe) {
            text[MKS_SIGN_CHAR] = B64_ALPHABET[v66 | 2];
            maybe = hay_match(text, m);
        }
        if (!maybe) continue;
        pk[31] |= (u8)(x_sign(st, (int)j) << 7);
        encode_tail(pk, text);
        if (hay_match(text, m)) {
            if (found < cap) hits[found] = first + j;
            found++;
        }
    }
    return found;
}

#undef NL
#undef ENTRY_WORDS
}  // namespace MKS_NS

#undef MKS_NS
#undef MKS_QR
