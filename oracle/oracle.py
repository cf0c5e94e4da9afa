"""ctypes view of the CPU oracle (liboracle.so) — test infrastructure only.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this
module. It restates the reference's per-key algorithm on the CPU; the functions
and their reference citations are in meikey_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libmeikey_ref.so")

CONTAINS, SUFFIX, REGEX = 0, 1, 2


def build(ref: bool = True) -> None:
    """Compile liboracle.so and the SIMD CPU baseline (libmeikey_simd.so), and the
    reference host build where its sources exist."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref and os.path.exists("/root/reference/src/service/gpu/host.cpp"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build(ref=False)
        L = ctypes.CDLL(LIB)
        u8p, u64 = ctypes.c_char_p, ctypes.c_uint64
        L.orc_seed.argtypes = [u8p, ctypes.c_int, u64, ctypes.c_void_p]
        L.orc_chacha20_block.argtypes = [ctypes.c_void_p, u64, ctypes.c_void_p]
        L.orc_sha512.argtypes = [u8p, ctypes.c_size_t, ctypes.c_void_p]
        L.orc_secret_scalar.argtypes = [u8p, ctypes.c_void_p]
        L.orc_fe_op.argtypes = [ctypes.c_int, u8p, u8p, ctypes.c_void_p]
        L.orc_digits.argtypes = [u8p, ctypes.c_int, ctypes.c_void_p]
        L.orc_mul_base_comb.argtypes = [u8p, ctypes.c_int, ctypes.c_void_p]
        L.orc_mul_base_ladder.argtypes = [u8p, ctypes.c_void_p]
        L.orc_comb_entry.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
        L.orc_comb_entry_niels.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_void_p]
        L.orc_public_key.argtypes = [u8p, ctypes.c_void_p]
        L.orc_public_b64.argtypes = [u8p, ctypes.c_void_p]
        L.orc_derive.argtypes = [u8p, ctypes.c_int, u64, u64, ctypes.c_void_p, ctypes.c_int]
        L.orc_match.argtypes = [u8p, u8p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        L.orc_match_unsigned.argtypes = [u8p, u8p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int]
        L.orc_search.argtypes = [u8p, ctypes.c_int, u64, u64, u8p, ctypes.c_uint32, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, u64, ctypes.c_int]
        L.orc_search.restype = u64
        _lib = L
    return _lib


_ref = None


def ref():
    """The reference's own kernel compiled as host C++ (oracle/_ref, built by
    `make ref` from the meikey sources), or None where it was not built. Its functions
    (kernel.cu:2083-2208 test ABI, prefix meikey_test_: 10 limbs, 32 keys per thread,
    packed 24-word table entries) pin the oracle against the reference itself."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF):
            return None
        L = ctypes.CDLL(REF)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.meikey_test_seed.argtypes = [vp, u64, vp]
        L.meikey_test_counter_seed.argtypes = [vp, u64, vp]
        L.meikey_test_secret.argtypes = [vp, vp]
        L.meikey_test_secret_mid.argtypes = [vp, u64, vp]
        L.meikey_test_digits.argtypes = [vp, u32, vp]
        L.meikey_test_derive_batch.argtypes = [vp, u32, vp, u32, u64, vp]
        L.meikey_test_contains.argtypes = [vp, vp, u32, u32, u32]
        L.meikey_test_contains_unsigned.argtypes = [vp, vp, u32, u32, u32]
        L.meikey_test_tail_b64.argtypes = [vp, vp]
        L.meikey_test_search.argtypes = [u64, vp, u32, vp, u32, vp, u32, u32, u32, u64, u32, u32, vp]
        L.meikey_test_batch.restype = u32
        L.meikey_test_entry_words.restype = u32
        _ref = L
    return _ref


# Limb offsets of the ten-limb radix-2^25.5 form and the word of limb i of field f in a
# 32-word entry of this repository's device table (limbs 0-7 of the fields in words
# 0-7 / 8-15 / 16-23, limbs 8-9 in words 24-25 / 26-27 / 28-29).
_LIMB_OFFSETS = [0, 26, 51, 77, 102, 128, 153, 179, 204, 230]


def _entry_word(f: int, i: int) -> int:
    return 8 * f + i if i < 8 else 24 + 2 * f + (i - 8)


_P_WORDS = [0xFFFFFFED] + [0xFFFFFFFF] * 6 + [0x7FFFFFFF]


def _double_mod_p(v):
    """2v mod p for canonical 256-bit values held as 8 little-endian 32-bit words in the
    columns of a uint64 array (one value per row), vectorized over the rows."""
    import numpy as np
    m32 = np.uint64(0xFFFFFFFF)
    d = np.zeros_like(v)
    carry = np.zeros(v.shape[0], dtype=np.uint64)
    for i in range(8):  # 2v < 2p < 2^256: no carry out of word 7
        t = (v[:, i] << np.uint64(1)) | carry
        d[:, i] = t & m32
        carry = t >> np.uint64(32)
    r = np.zeros_like(v)
    borrow = np.zeros(v.shape[0], dtype=np.int64)
    for i in range(8):  # 2v - p, word by word
        t = d[:, i].astype(np.int64) - np.int64(_P_WORDS[i]) - borrow
        borrow = (t < 0).astype(np.int64)
        r[:, i] = (t + (borrow << 32)).astype(np.uint64)
    keep = borrow.astype(bool)  # 2v < p
    r[keep] = d[keep]
    return r


def ref_table_from_device(words):
    """The reference's packed table (24 words per entry: (y+x, y-x, 2dxy) as canonical
    8-word LE values, row 0's third field 2xy, gpu.rs:1455-1480) from the device table's
    entries (numpy uint32, shape (entries, 32), canonical limbs of the halved fields
    ((y+x)/2, (y-x)/2, dxy; row 0 xy): each field is doubled mod p)."""
    import numpy as np
    n = words.shape[0]
    out = np.zeros((n, 24), dtype=np.uint64)
    for f in range(3):
        for i, off in enumerate(_LIMB_OFFSETS):
            limb = words[:, _entry_word(f, i)].astype(np.uint64) << np.uint64(off % 32)
            w = off // 32
            out[:, 8 * f + w] |= limb & np.uint64(0xFFFFFFFF)
            if w + 1 < 8:
                out[:, 8 * f + w + 1] |= limb >> np.uint64(32)
        out[:, 8 * f:8 * f + 8] = _double_mod_p(out[:, 8 * f:8 * f + 8])
    flat = out.astype(np.uint32).reshape(-1)
    return (ctypes.c_uint32 * flat.size).from_buffer_copy(flat.tobytes())


def _buf(n: int):
    return ctypes.create_string_buffer(n)


def seed(key: bytes, source: int, index: int) -> bytes:
    out = _buf(32)
    lib().orc_seed(key, source, index, out)
    return out.raw


def chacha20_block(key_words, counter: int) -> list:
    k = (ctypes.c_uint32 * 8)(*key_words)
    out = (ctypes.c_uint32 * 16)()
    lib().orc_chacha20_block(k, counter, out)
    return list(out)


def sha512(msg: bytes) -> bytes:
    out = _buf(64)
    lib().orc_sha512(msg, len(msg), out)
    return out.raw


def secret_scalar(seed_: bytes) -> bytes:
    out = _buf(32)
    lib().orc_secret_scalar(seed_, out)
    return out.raw


def fe_op(op: int, a: int, b: int) -> int:
    out = _buf(32)
    lib().orc_fe_op(op, a.to_bytes(32, "little"), b.to_bytes(32, "little"), out)
    return int.from_bytes(out.raw, "little")


def digits(s: bytes, w: int) -> list:
    out = (ctypes.c_int32 * 130)()
    rows = lib().orc_digits(s, w, out)
    return list(out)[:rows]


def mul_base_comb(s: bytes, w: int) -> bytes:
    out = _buf(32)
    assert lib().orc_mul_base_comb(s, w, out) == 0
    return out.raw


def mul_base_ladder(s: bytes) -> bytes:
    out = _buf(32)
    lib().orc_mul_base_ladder(s, out)
    return out.raw


def comb_entry(w: int, row: int, k: int) -> bytes:
    out = _buf(32)
    assert lib().orc_comb_entry(w, row, k, out) == 0
    return out.raw


def comb_entry_niels(w: int, row: int, k: int) -> bytes:
    out = _buf(96)
    assert lib().orc_comb_entry_niels(w, row, k, out) == 0
    return out.raw


def public_key(seed_: bytes) -> bytes:
    out = _buf(32)
    lib().orc_public_key(seed_, out)
    return out.raw


def public_b64(pk: bytes) -> str:
    out = _buf(68)
    lib().orc_public_b64(pk, out)
    return out.raw.decode()


def derive(key: bytes, source: int, first: int, n: int, threads: int = 0) -> bytes:
    out = _buf(32 * n) if n else _buf(1)
    lib().orc_derive(key, source, first, n, out, threads)
    return out.raw[: 32 * n]


def match(b64: str, needle: bytes, case_sensitive: bool, mode: int) -> bool:
    n = len(needle) // 130 if mode == REGEX else len(needle)
    return bool(lib().orc_match(b64.encode(), needle, n, int(case_sensitive), mode))


def match_unsigned(pk: bytes, needle: bytes, case_sensitive: bool, mode: int) -> bool:
    n = len(needle) // 130 if mode == REGEX else len(needle)
    return bool(lib().orc_match_unsigned(pk, needle, n, int(case_sensitive), mode))


def search(key: bytes, source: int, first: int, n: int, needle: bytes, case_sensitive: bool, mode: int,
           cap: int = 1 << 20, threads: int = 0) -> list:
    """Exact hit indices of [first, first + n), ascending."""
    length = len(needle) // 130 if mode == REGEX else len(needle)
    hits = (ctypes.c_uint64 * max(cap, 1))()
    total = lib().orc_search(key, source, first, n, needle, length, int(case_sensitive), mode, hits, cap, threads)
    if total > cap:
        raise OverflowError(f"{total} hits exceed cap {cap}")
    return list(hits)[:total]
