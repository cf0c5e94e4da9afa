"""Independent pure-Python references used to pin the oracle (test helpers only).

* Ed25519 key generation as written in RFC 8032 s5.1.5 / s6 (big integers, the
  textbook double-and-add, recover_x for the base point).
* ChaCha20 block function of RFC 8439 s2.3, with the state's last four words given
  explicitly so both the IETF nonce layout and the 64-bit-counter layout (words 12-13
  counter, 14-15 zero) can be expressed.
* base64 of the ssh-ed25519 public-key blob via the standard library.
"""
from __future__ import annotations

import base64
import hashlib
import struct

P = 2 ** 255 - 19
D = -121665 * pow(121666, P - 2, P) % P
SQRT_M1 = pow(2, (P - 1) // 4, P)


def inv(x: int) -> int:
    return pow(x, P - 2, P)


def recover_x(y: int, sign: int) -> int:
    x2 = (y * y - 1) * inv(D * y * y + 1) % P
    x = pow(x2, (P + 3) // 8, P)
    if (x * x - x2) % P != 0:
        x = x * SQRT_M1 % P
    assert (x * x - x2) % P == 0
    if (x & 1) != sign:
        x = P - x
    return x


G_Y = 4 * inv(5) % P
G_X = recover_x(G_Y, 0)
G = (G_X, G_Y, 1, G_X * G_Y % P)
IDENTITY = (0, 1, 1, 0)


def point_add(p1, p2):
    x1, y1, z1, t1 = p1
    x2, y2, z2, t2 = p2
    a = (y1 - x1) * (y2 - x2) % P
    b = (y1 + x1) * (y2 + x2) % P
    c = t1 * 2 * D * t2 % P
    d = z1 * 2 * z2 % P
    e, f, g, h = b - a, d - c, d + c, b + a
    return (e * f % P, g * h % P, f * g % P, e * h % P)


def point_mul(s: int, pt=G):
    q = IDENTITY
    while s > 0:
        if s & 1:
            q = point_add(q, pt)
        pt = point_add(pt, pt)
        s >>= 1
    return q


def compress(pt) -> bytes:
    x, y, z, _ = pt
    zi = inv(z)
    x, y = x * zi % P, y * zi % P
    return int.to_bytes(y | ((x & 1) << 255), 32, "little")


def secret_scalar(seed: bytes) -> int:
    h = hashlib.sha512(seed).digest()
    a = int.from_bytes(h[:32], "little")
    a &= (1 << 254) - 8
    a |= 1 << 254
    return a


def public_key(seed: bytes) -> bytes:
    return compress(point_mul(secret_scalar(seed)))


def _rotl(v: int, n: int) -> int:
    return ((v << n) | (v >> (32 - n))) & 0xFFFFFFFF


def chacha20_block_words(key: bytes, w12: int, w13: int, w14: int, w15: int) -> list:
    const = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    state = const + list(struct.unpack("<8I", key)) + [w12, w13, w14, w15]
    x = list(state)

    def qr(a, b, c, d):
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[d] = _rotl(x[d] ^ x[a], 16)
        x[c] = (x[c] + x[d]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[c], 12)
        x[a] = (x[a] + x[b]) & 0xFFFFFFFF; x[d] = _rotl(x[d] ^ x[a], 8)
        x[c] = (x[c] + x[d]) & 0xFFFFFFFF; x[b] = _rotl(x[b] ^ x[c], 7)

    for _ in range(10):
        qr(0, 4, 8, 12); qr(1, 5, 9, 13); qr(2, 6, 10, 14); qr(3, 7, 11, 15)
        qr(0, 5, 10, 15); qr(1, 6, 11, 12); qr(2, 7, 8, 13); qr(3, 4, 9, 14)
    return [(x[i] + state[i]) & 0xFFFFFFFF for i in range(16)]


def chacha20_seed(key: bytes, index: int) -> bytes:
    """Seed of `index`: half (index & 1) of the block index >> 1 with a 64-bit counter."""
    ctr = index >> 1
    words = chacha20_block_words(key, ctr & 0xFFFFFFFF, ctr >> 32, 0, 0)
    block = struct.pack("<16I", *words)
    return block[32 * (index & 1): 32 * (index & 1) + 32]


def counter_seed(key: bytes, index: int) -> bytes:
    tail = (int.from_bytes(key[24:], "little") + index) % (1 << 64)
    return key[:24] + tail.to_bytes(8, "little")


def public_b64(pk: bytes) -> str:
    blob = struct.pack(">I", 11) + b"ssh-ed25519" + struct.pack(">I", 32) + pk
    return base64.b64encode(blob).decode()


def hay(pk: bytes) -> str:
    return public_b64(pk)[25:]
