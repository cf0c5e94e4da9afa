/*
 * CPU oracle for the Ed25519 vanity-key search path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker the HIP kernels are compared against. Only the
 * tests, the smoke check and the benchmark's CPU-baseline leg may load it; the
 * product library (libmeikey_hip.so) never links or calls it.
 *
 * It restates, in plain C, the reference's per-key algorithm:
 *   seed (ChaCha20 or counter) -> clamp(SHA-512(seed)[..32]) -> signed-digit
 *   comb multiply s*B -> compressed key (with the x sign) -> base64 -> match.
 * Every function cites the reference file:line it follows. How each function is
 * pinned (RFC vectors, hashlib, a textbook double-and-add, the reference's own
 * test vectors) is listed at the top of meikey_oracle.c.
 */
#ifndef MEIKEY_ORACLE_H
#define MEIKEY_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- seeds ---------------------------------------------------------------- */
void orc_chacha20_block(const uint32_t key[8], uint64_t counter, uint32_t out[16]);
void orc_chacha20_seed(const uint8_t key[32], uint64_t index, uint8_t seed[32]);
void orc_counter_seed(const uint8_t key[32], uint64_t index, uint8_t seed[32]);
/* source: 0 = ChaCha20 (default, "secure"), 1 = counter (--low-security) */
void orc_seed(const uint8_t key[32], int source, uint64_t index, uint8_t seed[32]);

/* --- SHA-512 + clamp ------------------------------------------------------ */
void orc_sha512(const uint8_t* msg, size_t len, uint8_t out[64]);
void orc_secret_scalar(const uint8_t seed[32], uint8_t s[32]);

/* --- field / curve (test hooks) ------------------------------------------- */
/* op: 0 add, 1 sub, 2 mul, 3 sq, 4 invert. a, b: 32-byte LE values < 2^256 (reduced
 * mod p on load). out: canonical 32-byte LE. */
void orc_fe_op(int op, const uint8_t a[32], const uint8_t b[32], uint8_t out[32]);
/* Signed 2^w digits of the 251-bit m of a clamped scalar; returns the row count. */
int orc_digits(const uint8_t s[32], int w, int32_t* out);
/* s*B by the signed-digit comb of width w (2..16); pk = compressed, with sign. */
int orc_mul_base_comb(const uint8_t s[32], int w, uint8_t pk[32]);
/* s*B by textbook double-and-add over all 256 bits (independent check). */
void orc_mul_base_ladder(const uint8_t s[32], uint8_t pk[32]);
/* Compressed encoding of k * 2^(w*row + 3) * B (+ 2^254 B on the top row) —
 * the comb table's closed form. */
int orc_comb_entry(int w, int row, uint32_t k, uint8_t pk[32]);
/* Affine (y+x, y-x, 2dxy) of the comb entry (row, k), each canonical 32-byte LE. */
int orc_comb_entry_niels(int w, int row, uint32_t k, uint8_t out[96]);

/* --- keys ------------------------------------------------------------------ */
void orc_public_key(const uint8_t seed[32], uint8_t pk[32]);
/* 68-char base64 of the ssh-ed25519 public-key blob (no NUL). */
void orc_public_b64(const uint8_t pk[32], char out[68]);
/* n public keys for indices first..first+n-1 of one master key (threads: 0 = auto). */
void orc_derive(const uint8_t key[32], int source, uint64_t first, uint64_t n, uint8_t* pks, int threads);

/* --- matching ------------------------------------------------------------- */
enum { ORC_CONTAINS = 0, ORC_SUFFIX = 1, ORC_REGEX = 2 };
/* Exact match of the 43-char hay b64[25..68]. For ORC_REGEX, needle is the DFA table
 * (states x 65 x u16 LE, col 64 = end-of-input flag) and len the state count. */
int orc_match(const char b64[68], const uint8_t* needle, uint32_t len, int case_sensitive, int mode);
/* Sign-unknown form: true if the key matches with pk[31] bit 7 either 0 or 1. */
int orc_match_unsigned(const uint8_t pk[32], const uint8_t* needle, uint32_t len, int case_sensitive, int mode);

/* Exact hit set of [first, first+n): every index whose key matches. Returns the hit
 * count (all of them); writes at most cap indices, ascending. */
uint64_t orc_search(const uint8_t key[32], int source, uint64_t first, uint64_t n, const uint8_t* needle,
                    uint32_t len, int case_sensitive, int mode, uint64_t* hits, uint64_t cap, int threads);

#ifdef __cplusplus
}
#endif
#endif
