"""meikey_amd — MI355X (gfx950) backend of meikey's Ed25519 vanity-key search.

Python view of the C ABI in include/meikey_hip.h (libmeikey_hip.so, built in-tree by
meikey_amd/build.py). The library holds the HIP kernels; this module only marshals
arguments. Importing it never falls back to CPU code: if the library is missing or
the device cannot be opened, calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import List, Optional, Sequence

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("MEIKEY_HIP_LIB") or os.path.join(HERE, "libmeikey_hip.so")

SEED_CHACHA20 = 0
SEED_COUNTER = 1
MODE_CONTAINS = 0
MODE_SUFFIX = 1
MODE_REGEX = 2
LIMBS_R25 = 10

STATUS = {
    0: "ok",
    1: "driver not found",
    2: "no device",
    3: "out of range",
    4: "self-check failed",
    5: "driver error",
    6: "invalid argument",
    7: "overflow",
}

# Every entry point declared in include/meikey_hip.h.
EXPORTS = (
    "mk_last_error", "mk_device_count", "mk_default_options", "mk_open", "mk_info_get", "mk_close",
    "mk_keys_per_iter", "mk_prepare", "mk_launch_run", "mk_free_launch", "mk_derive", "mk_derive_seeds",
    "mk_table_read", "mk_batch_run", "mk_tile_keys", "mk_last_timing",
)


class MeikeyError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.message = message


class Options(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("blocks", ctypes.c_uint32)]


class Info(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("compute_units", ctypes.c_uint32), ("grid_waves", ctypes.c_uint32),
                ("hbm_bytes", ctypes.c_uint64), ("table_bytes", ctypes.c_uint64), ("scratch_bytes", ctypes.c_uint64),
                ("options", Options), ("rows", ctypes.c_uint32), ("cols", ctypes.c_uint32)]


class HitRec(ctypes.Structure):
    _fields_ = [("index", ctypes.c_uint64), ("search", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("seed", ctypes.c_uint8 * 32), ("pk", ctypes.c_uint8 * 32)]


class SearchSpec(ctypes.Structure):
    _fields_ = [("key", ctypes.c_uint8 * 32), ("seed_source", ctypes.c_int32), ("mode", ctypes.c_uint32),
                ("fold", ctypes.c_uint32), ("needle_len", ctypes.c_uint32), ("needle", ctypes.c_char_p),
                ("stop_on_hit", ctypes.c_uint32), ("pad", ctypes.c_uint32), ("first", ctypes.c_uint64),
                ("count", ctypes.c_uint64)]


@dataclass(frozen=True)
class Hit:
    index: int
    search: int
    seed: bytes
    pk: bytes


_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MeikeyError(1, f"{LIB_PATH} is missing; run meikey_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    P, U32, U64, I32 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    L.mk_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.mk_last_error.restype = None
    L.mk_device_count.argtypes = [ctypes.POINTER(I32)]
    L.mk_default_options.argtypes = [ctypes.POINTER(Options)]
    L.mk_default_options.restype = None
    L.mk_open.argtypes = [I32, ctypes.POINTER(Options), ctypes.POINTER(P)]
    L.mk_info_get.argtypes = [P, ctypes.POINTER(Info)]
    L.mk_close.argtypes = [P]
    L.mk_keys_per_iter.argtypes = [P]
    L.mk_keys_per_iter.restype = U64
    L.mk_tile_keys.argtypes = [P]
    L.mk_tile_keys.restype = U64
    L.mk_prepare.argtypes = [P, ctypes.c_char_p, I32, ctypes.c_char_p, U32, U32, U32, U32, ctypes.POINTER(P)]
    L.mk_launch_run.argtypes = [P, U64, U32, ctypes.POINTER(HitRec), U32, ctypes.POINTER(U32)]
    L.mk_free_launch.argtypes = [P]
    L.mk_derive.argtypes = [P, ctypes.c_char_p, I32, U64, U64, P, P]
    L.mk_derive_seeds.argtypes = [P, ctypes.c_char_p, U32, P]
    L.mk_table_read.argtypes = [P, U32, U32, U32, P]
    L.mk_last_timing.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.mk_batch_run.argtypes = [P, ctypes.POINTER(SearchSpec), U32, ctypes.POINTER(HitRec), U32,
                               ctypes.POINTER(U32), ctypes.POINTER(U64)]
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is ctypes.c_int and name not in ("mk_keys_per_iter", "mk_tile_keys"):
            fn.restype = I32
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().mk_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(status: int) -> None:
    if status != 0:
        raise MeikeyError(status, last_error())


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().mk_device_count(ctypes.byref(n)))
    return n.value


def _hits(recs, n: int) -> List[Hit]:
    return [Hit(recs[i].index, recs[i].search, bytes(recs[i].seed), bytes(recs[i].pk)) for i in range(n)]


class Launch:
    """A prepared single search (the reference's Engine::prepare result)."""

    def __init__(self, gpu: "Gpu", handle: ctypes.c_void_p):
        self.gpu = gpu
        self._h = handle

    def run(self, base: int, iters: int, cap: int = 1 << 16) -> List[Hit]:
        """Engine::launch: exact hits of [base, base + keys_per_iter * iters), ascending."""
        recs = (HitRec * max(cap, 1))()
        n = ctypes.c_uint32(0)
        _check(lib().mk_launch_run(self._h, base, iters, recs, cap, ctypes.byref(n)))
        return _hits(recs, min(n.value, cap))

    def close(self) -> None:
        if self._h:
            lib().mk_free_launch(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Gpu:
    """One device with its comb table (the reference's Gpu + Kernel)."""

    def __init__(self, ordinal: int = 0, width: int = 24, batch: int = 64, blocks: int = 2, limbs: int = LIMBS_R25):
        self.options = Options(width, batch, limbs, blocks)
        h = ctypes.c_void_p()
        _check(lib().mk_open(ordinal, ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h

    def info(self) -> Info:
        out = Info()
        _check(lib().mk_info_get(self._h, ctypes.byref(out)))
        return out

    @property
    def keys_per_iter(self) -> int:
        return lib().mk_keys_per_iter(self._h)

    @property
    def tile_keys(self) -> int:
        return lib().mk_tile_keys(self._h)

    def prepare(self, key: bytes, seed_source: int, needle: bytes, fold: bool, mode: int,
                stop_on_hit: bool = False, needle_len: Optional[int] = None) -> Launch:
        if needle_len is None:
            needle_len = len(needle) // 130 if mode == MODE_REGEX else len(needle)
        h = ctypes.c_void_p()
        _check(lib().mk_prepare(self._h, key, seed_source, needle, needle_len, int(fold), mode, int(stop_on_hit),
                                ctypes.byref(h)))
        return Launch(self, h)

    def derive(self, key: bytes, seed_source: int, first: int, n: int, seeds: bool = False):
        """Compressed public keys (n x 32 bytes) of indices first .. first+n-1."""
        pks = ctypes.create_string_buffer(max(32 * n, 1))
        sd = ctypes.create_string_buffer(max(32 * n, 1)) if seeds else None
        _check(lib().mk_derive(self._h, key, seed_source, first, n, pks, sd))
        if seeds:
            return pks.raw[: 32 * n], sd.raw[: 32 * n]
        return pks.raw[: 32 * n]

    def derive_seeds(self, seeds: bytes) -> bytes:
        n = len(seeds) // 32
        out = ctypes.create_string_buffer(max(32 * n, 1))
        _check(lib().mk_derive_seeds(self._h, seeds, n, out))
        return out.raw[: 32 * n]

    def last_timing(self):
        """(search kernel ms, confirm kernel ms) of the last launch, from HIP events."""
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        _check(lib().mk_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def table_words(self, row: int, k: int, count: int = 1) -> List[int]:
        buf = (ctypes.c_uint32 * (32 * count))()
        _check(lib().mk_table_read(self._h, row, k, count, buf))
        return list(buf)

    def batch_run(self, specs: Sequence[dict], cap: int = 1 << 16):
        """Run many searches in one launch. Each spec: key, seed_source, mode, fold, needle,
        first, count, stop_on_hit (optional). Returns (hits, tries per search)."""
        n = len(specs)
        arr = (SearchSpec * max(n, 1))()
        keep = []
        for i, s in enumerate(specs):
            needle = s["needle"]
            keep.append(needle)
            arr[i].key[:] = list(s["key"])
            arr[i].seed_source = s.get("seed_source", SEED_CHACHA20)
            arr[i].mode = s.get("mode", MODE_CONTAINS)
            arr[i].fold = int(s.get("fold", False))
            arr[i].needle_len = s.get("needle_len", len(needle) // 130 if arr[i].mode == MODE_REGEX else len(needle))
            arr[i].needle = needle
            arr[i].stop_on_hit = int(s.get("stop_on_hit", False))
            arr[i].first = s.get("first", 0)
            arr[i].count = s["count"]
        recs = (HitRec * max(cap, 1))()
        nh = ctypes.c_uint32(0)
        tries = (ctypes.c_uint64 * max(n, 1))()
        _check(lib().mk_batch_run(self._h, arr, n, recs, cap, ctypes.byref(nh), tries))
        return _hits(recs, min(nh.value, cap)), list(tries)[:n]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mk_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
