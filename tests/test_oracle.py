"""Pins the CPU oracle before anything is compared with it.

Each check uses an independent source: RFC 8032 / RFC 8439 vectors, hashlib, the
pure-Python textbook references in tests/textbook.py, closed forms of the comb
table and the digit split, and the fixed match cases of the reference's tests.
"""
import hashlib
import os
import random
import re
import struct

import pytest

from tests import textbook as T

HERE = os.path.dirname(os.path.abspath(__file__))


def rfc8032_vectors():
    out = []
    for line in open(os.path.join(HERE, "golden", "rfc8032-ed25519.txt"), encoding="utf-8"):
        if line.startswith("#") or not line.strip():
            continue
        f = line.strip().split(":")
        out.append((f[0], bytes.fromhex(f[1]), bytes.fromhex(f[2])))
    return out


def test_rfc8032_vectors(oracle):
    vecs = rfc8032_vectors()
    assert len(vecs) == 5
    for name, sk, pk in vecs:
        assert oracle.public_key(sk) == pk, name
        assert T.public_key(sk) == pk, name
        assert oracle.mul_base_ladder(oracle.secret_scalar(sk)) == pk, name


def test_sha512_matches_hashlib(oracle):
    rng = random.Random(7)
    for n in list(range(0, 260, 7)) + [111, 112, 127, 128, 129, 239, 240, 256]:
        m = rng.randbytes(n)
        assert oracle.sha512(m) == hashlib.sha512(m).digest(), n


def test_secret_scalar_is_clamped_sha512(oracle):
    rng = random.Random(8)
    for _ in range(200):
        seed = rng.randbytes(32)
        s = oracle.secret_scalar(seed)
        assert int.from_bytes(s, "little") == T.secret_scalar(seed)
        assert s[0] & 7 == 0 and s[31] & 0xC0 == 0x40


# RFC 8439 Appendix A.1: zero key, zero nonce, block counters 0 and 1.
A1_0 = bytes.fromhex(
    "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
    "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
A1_1 = bytes.fromhex(
    "9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
    "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")


def test_chacha20_rfc8439_vectors(oracle):
    zero = [0] * 8
    assert struct.pack("<16I", *oracle.chacha20_block(zero, 0)) == A1_0
    assert struct.pack("<16I", *oracle.chacha20_block(zero, 1)) == A1_1
    assert struct.pack("<16I", *T.chacha20_block_words(bytes(32), 0, 0, 0, 0)) == A1_0
    # RFC 8439 s2.3.2 (IETF nonce layout), first state words of the output
    key = bytes(range(32))
    w = T.chacha20_block_words(key, 1, 0x09000000, 0x4A000000, 0)
    assert w[:4] == [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3]


def test_seed_mapping_matches_textbook(oracle):
    rng = random.Random(9)
    idx = [0, 1, 2, 3, 255, 256, 2**32 - 1, 2**32, 2**32 + 1, 2**63, 2**64 - 2, 2**64 - 1]
    idx += [rng.getrandbits(64) for _ in range(40)]
    for _ in range(4):
        key = rng.randbytes(32)
        for i in idx:
            assert oracle.seed(key, 0, i) == T.chacha20_seed(key, i), i


def test_seeds_are_consecutive_chacha20_output(oracle):
    # gpu.rs:1685-1692: seeds 2b and 2b+1 are the 64 bytes of block b
    key = random.Random(10).randbytes(32)
    kw = struct.unpack("<8I", key)
    for b in (0, 1, 77, 2**40 + 5):
        block = struct.pack("<16I", *oracle.chacha20_block(list(kw), b))
        assert oracle.seed(key, 0, 2 * b) + oracle.seed(key, 0, 2 * b + 1) == block


def test_counter_seed(oracle):
    rng = random.Random(11)
    key = rng.randbytes(24) + (2**64 - 3).to_bytes(8, "little")
    for i in (0, 1, 2, 3, 4, 1000, 2**64 - 1):
        s = oracle.seed(key, 1, i)
        assert s == T.counter_seed(key, i), i
        assert s[:24] == key[:24]
    assert oracle.seed(key, 1, 3) == key[:24] + (0).to_bytes(8, "little")


def test_public_keys_match_textbook(oracle):
    rng = random.Random(12)
    for _ in range(25):
        seed = rng.randbytes(32)
        assert oracle.public_key(seed) == T.public_key(seed)


def test_comb_equals_ladder_every_width(oracle):
    rng = random.Random(13)
    scalars = [oracle.secret_scalar(rng.randbytes(32)) for _ in range(6)]
    # extreme clamped scalars: m = 0 and m = 2^251 - 1
    scalars.append(bytes([0] * 31 + [0x40]))
    scalars.append(bytes([0xF8] + [0xFF] * 30 + [0x7F]))
    for w in range(2, 17):
        for s in scalars:
            assert oracle.mul_base_comb(s, w) == oracle.mul_base_ladder(s), (w, s.hex())


def test_comb_entries_closed_form(oracle):
    rng = random.Random(14)
    for w in (2, 3, 5, 8, 12, 16):
        rows = -(-251 // w)
        cols = (1 << (w - 1)) + 1
        for _ in range(4):
            row, k = rng.randrange(rows), rng.randrange(cols)
            want = T.compress(T.point_mul(k << (w * row + 3)) if row < rows - 1 else
                              T.point_add(T.point_mul(k << (w * row + 3)), T.point_mul(1 << 254)))
            assert oracle.comb_entry(w, row, k) == want, (w, row, k)
            # the table's Niels form describes the same point
            n = oracle.comb_entry_niels(w, row, k)
            ypx, ymx, xy2d = (int.from_bytes(n[i:i + 32], "little") for i in (0, 32, 64))
            y = (ypx + ymx) * T.inv(2) % T.P
            x = (ypx - ymx) * T.inv(2) % T.P
            assert xy2d == 2 * T.D * x * y % T.P
            assert int.to_bytes(y | ((x & 1) << 255), 32, "little") == want


def test_digits_reconstruct(oracle):
    rng = random.Random(15)
    for w in list(range(2, 25)):
        for _ in range(6):
            s = oracle.secret_scalar(rng.randbytes(32))
            m = (int.from_bytes(s, "little") >> 3) & ((1 << 251) - 1)
            d = oracle.digits(s, w)
            assert len(d) == -(-251 // w)
            assert sum(di << (w * i) for i, di in enumerate(d)) == m
            half = 1 << (w - 1)
            assert all(-half <= di < half for di in d[:-1])
            assert 0 <= d[-1] <= half


EDGE = [0, 1, 38, T.P - 1, T.P, T.P + 1, 2 * T.P - 1, 2**256 - 38, 2**256 - 1, 2**255, 2**255 - 1, 2**204 - 1]


def test_field_edge_values(oracle):
    rng = random.Random(16)
    vals = EDGE + [rng.getrandbits(256) for _ in range(20)]
    for a in vals:
        for b in vals[:8] + [rng.getrandbits(256)]:
            assert oracle.fe_op(0, a, b) == (a + b) % T.P
            assert oracle.fe_op(1, a, b) == (a - b) % T.P
            assert oracle.fe_op(2, a, b) == a * b % T.P
        assert oracle.fe_op(3, a, 0) == a * a % T.P
        assert oracle.fe_op(4, a, 0) == pow(a, T.P - 2, T.P)


def test_base64_known_line(oracle):
    sk = rfc8032_vectors()[0][1]
    b64 = oracle.public_b64(oracle.public_key(sk))
    # src/model/key.rs:161-169
    assert b64 == "AAAAC3NzaC1lZDI1NTE5AAAAINdamAGCsQq31Uv+08lkBzoO4XLz2qYjJa8CGmj3B1Ea"
    rng = random.Random(17)
    for _ in range(100):
        pk = rng.randbytes(32)
        assert oracle.public_b64(pk) == T.public_b64(pk)
        assert T.public_b64(pk)[25] in "ABCDEFGHIJKLMNOP"


# A real key and a key whose variable part is all '0' (src/model/pattern.rs:302-304, 349-350)
SAMPLE = "AAAAC3NzaC1lZDI1NTE5AAAAIIQK2/0eXLDku+cWkIDsqHQOwFFTEDDoWaKofXuV/Gy9"
PREFIX_ONLY = "AAAAC3NzaC1lZDI1NTE5AAAAI0000000000000000000000000000000000000000000"

# (b64, needle, case_sensitive, mode, expected) from pattern.rs:352-399
FIXED_CASES = [
    (SAMPLE, "wako", False, 0, True), (SAMPLE, "WAKO", False, 0, True), (SAMPLE, "WaKo", True, 0, True),
    (SAMPLE, "wako", True, 0, False), (SAMPLE, "Gy9", True, 0, True), (SAMPLE, "IQK2", True, 0, True),
    (SAMPLE, "zzzz", False, 0, False), (PREFIX_ONLY, "AAAAC3", True, 0, False),
    (PREFIX_ONLY, "NzaC1lZDI1", True, 0, False), (PREFIX_ONLY, "NTE5", True, 0, False),
    (PREFIX_ONLY, "0000", True, 0, True), (SAMPLE, "Gy9", True, 1, True), (SAMPLE, "gY9", False, 1, True),
    (SAMPLE, "WaKo", True, 1, False),
]


@pytest.mark.parametrize("b64,needle,cs,mode,want", FIXED_CASES)
def test_match_fixed_cases(oracle, b64, needle, cs, mode, want):
    assert oracle.match(b64, needle.encode(), cs, mode) == want


def py_match(b64: str, needle: str, cs: bool, mode: int) -> bool:
    hay = b64[25:]
    if not cs:
        hay, needle = hay.lower(), needle.lower()
    return hay.endswith(needle) if mode == 1 else needle in hay


def test_match_random_vs_python(oracle):
    rng = random.Random(18)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    for _ in range(3000):
        pk = rng.randbytes(32)
        b64 = T.public_b64(pk)
        L = rng.choice([1, 1, 2, 2, 3, 4, 42, 43])
        if rng.random() < 0.5 and L <= 43:
            st = rng.randrange(25, 69 - L)
            needle = b64[st:st + L]
        else:
            needle = "".join(rng.choice(alpha) for _ in range(L))
        if rng.random() < 0.5:
            needle = needle.swapcase()
        cs, mode = rng.random() < 0.5, rng.choice([0, 1])
        assert oracle.match(b64, needle.encode(), cs, mode) == py_match(b64, needle, cs, mode), (b64, needle, cs, mode)
        pk0 = pk[:31] + bytes([pk[31] & 0x7F])
        either = py_match(T.public_b64(pk0), needle, cs, mode) or py_match(
            T.public_b64(pk[:31] + bytes([pk[31] | 0x80])), needle, cs, mode)
        assert oracle.match_unsigned(pk, needle.encode(), cs, mode) == either


def test_derive_and_search_agree(oracle):
    key = random.Random(1).randbytes(32)
    n = 4096
    pks = oracle.derive(key, 0, 1000, n, threads=4)
    for i in (0, 1, 2, 3, n - 1):
        assert pks[32 * i:32 * i + 32] == T.public_key(T.chacha20_seed(key, 1000 + i))
    needle = "ab"
    want = [1000 + i for i in range(n) if py_match(T.public_b64(pks[32 * i:32 * i + 32]), needle, False, 0)]
    got = oracle.search(key, 0, 1000, n, needle.encode(), False, 0, threads=3)
    assert got == want and len(want) > 10


def test_golden_fixture(oracle):
    """The committed fixture (made by tests/golden/make_golden.py from the textbook
    references) agrees with the oracle."""
    import json
    g = json.load(open(os.path.join(HERE, "golden", "golden.json")))
    for case in g["derive"]:
        key = bytes.fromhex(case["key"])
        for idx, pk in case["pks"]:
            assert oracle.derive(key, case["source"], idx, 1) == bytes.fromhex(pk), (case["source"], idx)
    for case in g["search"]:
        key = bytes.fromhex(case["key"])
        got = oracle.search(key, case["source"], case["first"], case["count"], case["needle"].encode(),
                            case["case_sensitive"], case["mode"])
        assert got == case["hits"]
