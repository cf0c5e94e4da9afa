"""ctypes view of libmeikey_hostcheck.so: the kernel arithmetic headers compiled for the
host (g++), so CPU tests can exercise the exact code the gfx950 kernels run."""
from __future__ import annotations

import ctypes
import os
import struct

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "meikey_amd", "libmeikey_hostcheck.so")

_lib = None
_layouts = {}


def load(layout: str = ""):
    """The CPU build of one layout: "" (ten limbs, 128-byte entries of limbs), "_p10" (ten
    limbs, packed 96-byte entries) or "_r32" (eight 32-bit words, packed entries)."""
    if layout not in _layouts:
        path = os.path.join(ROOT, "meikey_amd", f"libmeikey_hostcheck{layout}.so")
        if not os.path.exists(path):
            from meikey_amd import build
            build.build_hostcheck()
        _layouts[layout] = _bind(ctypes.CDLL(path))
    return _layouts[layout]


def lib():
    global _lib
    if _lib is None:
        _lib = load("")
    return _lib


def _bind(L):
    P, U32, U64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    L.hc_seed.argtypes = [P, U32, U64, P]
    L.hc_secret_scalar.argtypes = [P, P]
    L.hc_secret_scalar_counter.argtypes = [P, U64, P]
    L.hc_fe_mul_limbs.argtypes = [P, P, P]
    L.hc_fe_sq_limbs.argtypes = [P, P]
    L.hc_fe_to_words.argtypes = [P, P]
    L.hc_fe_from_words.argtypes = [P, P]
    L.hc_fe_op.argtypes = [U32, P, P, P]
    L.hc_digits.argtypes = [P, U32, P]
    L.hc_table_words.argtypes = [U32]
    L.hc_table_words.restype = U64
    L.hc_table_build.argtypes = [U32, P]
    L.hc_table_entry.argtypes = [U32, U32, U32, P]
    L.hc_full_pk.argtypes = [P, U32, P, P]
    L.hc_lane.argtypes = [P, U32, P, U32, U64, U32, ctypes.c_char_p, U32, U32, U32, P, P]
    L.hc_build_mask.argtypes = [ctypes.c_char_p, U32, U32, U32, P]
    L.hc_match.argtypes = [P, ctypes.c_char_p, U32, U32, U32, U32]
    L.hc_confirm.argtypes = [P, U32, U64, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, U32, U32, U32]
    L.hc_pk_by_ladder.argtypes = [P, P]
    L.hc_sort_hits.argtypes = [P, P, U32, P]
    L.hc_fe_carried_to_words.argtypes = [P, P]
    L.hc_fe_op_limbs.argtypes = [U32, P, P, P]
    L.hc_layout.restype = ctypes.c_int
    L.hc_claims.argtypes = [U32, U32, U64, U64, P, P, U64]
    L.hc_claims.restype = ctypes.c_long
    return L


def words(b: bytes):
    return (ctypes.c_uint32 * 8)(*struct.unpack("<8I", b))


def wbytes(w) -> bytes:
    return struct.pack("<8I", *list(w)[:8])


def seed(key: bytes, source: int, index: int) -> bytes:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_seed(words(key), source, index, out)
    return wbytes(out)


def secret_scalar(seed_: bytes) -> bytes:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_secret_scalar(words(seed_), out)
    return wbytes(out)


def secret_scalar_counter(key: bytes, index: int) -> bytes:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_secret_scalar_counter(words(key), index, out)
    return wbytes(out)


def pk_by_ladder(s: bytes) -> bytes:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_pk_by_ladder(words(s), out)
    return wbytes(out)


def confirm(key: bytes, source: int, index: int, seed_: bytes, pk: bytes, needle: bytes, fold: bool,
            mode: int) -> int:
    n = len(needle) // 130 if mode == 2 else len(needle)
    return lib().hc_confirm(words(key), source, index, seed_, pk, needle, n, int(fold), mode)


def limbs_of(x: int):
    """Integer to ten limbs at offsets 0,26,51,... (no reduction, x < 2^255)."""
    offs = [0, 26, 51, 77, 102, 128, 153, 179, 204, 230, 255]
    return [(x >> offs[i]) & ((1 << (offs[i + 1] - offs[i])) - 1) for i in range(10)]


# Word index of limb i of field f (0 = y+x, 1 = y-x, 2 = third) in a 32-word table entry:
# limbs 0-7 of the fields in words 0-7, 8-15, 16-23; limbs 8-9 in words 24-25, 26-27, 28-29;
# words 30-31 are padding.
def entry_word(f: int, i: int) -> int:
    return 8 * f + i if i < 8 else 24 + 2 * f + (i - 8)


def entry_limbs(words, f: int):
    return [words[entry_word(f, i)] for i in range(10)]


def value_of(limbs) -> int:
    offs = [0, 26, 51, 77, 102, 128, 153, 179, 204, 230]
    return sum(int(l) << o for l, o in zip(limbs, offs))


def fe_mul_limbs(f, g):
    a = (ctypes.c_uint32 * 10)(*f)
    b = (ctypes.c_uint32 * 10)(*g)
    out = (ctypes.c_uint32 * 10)()
    lib().hc_fe_mul_limbs(a, b, out)
    return list(out)


def fe_sq_limbs(f):
    a = (ctypes.c_uint32 * 10)(*f)
    out = (ctypes.c_uint32 * 10)()
    lib().hc_fe_sq_limbs(a, out)
    return list(out)


def fe_to_int(limbs) -> int:
    a = (ctypes.c_uint32 * 10)(*limbs)
    out = (ctypes.c_uint32 * 8)()
    lib().hc_fe_to_words(a, out)
    return int.from_bytes(wbytes(out), "little")


def fe_op(op: int, a: int, b: int) -> int:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_fe_op(op, words(a.to_bytes(32, "little")), words(b.to_bytes(32, "little")), out)
    return int.from_bytes(wbytes(out), "little")


def digits(s: bytes, w: int):
    out = (ctypes.c_int32 * 130)()
    rows = lib().hc_digits(words(s), w, out)
    return list(out)[:rows]


def table(w: int):
    n = lib().hc_table_words(w)
    buf = (ctypes.c_uint32 * n)()
    lib().hc_table_build(w, buf)
    return buf


def table_entry(w: int, row: int, k: int):
    out = (ctypes.c_uint32 * 32)()
    lib().hc_table_entry(w, row, k, out)
    return list(out)


def full_pk(tab, w: int, s: bytes) -> bytes:
    out = (ctypes.c_uint32 * 8)()
    lib().hc_full_pk(tab, w, words(s), out)
    return wbytes(out)


def lane(tab, w: int, key: bytes, source: int, first: int, K: int, needle: bytes, fold: bool, mode: int):
    n = len(needle) // 130 if mode == 2 else len(needle)
    ys = (ctypes.c_uint32 * (8 * K))()
    cand = (ctypes.c_uint8 * K)()
    rc = lib().hc_lane(tab, w, words(key), source, first, K, needle, n, int(fold), mode, ys, cand)
    assert rc == 0
    return [struct.pack("<8I", *ys[8 * j: 8 * j + 8]) for j in range(K)], list(cand)


def build_mask(needle: bytes, fold: bool, mode: int = 0):
    out = (ctypes.c_uint64 * 64)()
    rc = lib().hc_build_mask(needle, len(needle), int(fold), mode, out)
    assert rc == 0
    return list(out)


def match(pk: bytes, needle: bytes, fold: bool, mode: int, sign_unknown: bool) -> bool:
    n = len(needle) // 130 if mode == 2 else len(needle)
    rc = lib().hc_match(words(pk), needle, n, int(fold), mode, int(sign_unknown))
    assert rc >= 0
    return bool(rc)


def sort_hits(search, index):
    """The library's (search, index) order of the hits: a list of positions."""
    import numpy as np
    s = np.ascontiguousarray(search, dtype=np.uint32)
    x = np.ascontiguousarray(index, dtype=np.uint64)
    out = np.zeros(len(s), dtype=np.uint32)
    lib().hc_sort_hits(s.ctypes.data, x.ctypes.data, len(s), out.ctypes.data)
    return out


def fe_carried_to_int(limbs) -> int:
    """fe_carried_to_words of carried limbs, as an integer."""
    a = (ctypes.c_uint32 * 10)(*limbs)
    out = (ctypes.c_uint32 * 8)()
    lib().hc_fe_carried_to_words(a, out)
    return int.from_bytes(struct.pack("<8I", *out), "little")
