"""ctypes view of the SIMD CPU baseline (libmeikey_simd.so, from simd/meikey_simd.cpp) —
test and measurement infrastructure only.

The library restates the reference's CPU search path (AVX-512 IFMA or AVX2 lanes, W = 12
comb, 256-key batch inversion, unsigned-first matching). Only tests/ and the benchmark's
CPU-baseline leg import this module; its table comes from the oracle's W = 12 comb
entries.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

from . import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmeikey_simd.so")
W, ROWS, COLS = 12, 21, 2049

_lib = None
_table = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "simd"], check=True)


def lib() -> ctypes.CDLL:
    """The library with its table loaded."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, u32, u64, i = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        L.mks_load_table.argtypes = [vp]
        L.mks_derive.argtypes = [vp, i, u64, u64, vp]
        L.mks_search.argtypes = [vp, i, u64, u64, vp, u32, i, i, vp, u64, i]
        L.mks_search.restype = ctypes.c_longlong
        L.mks_bench.argtypes = [vp, i, u64, vp, u32, i, i, ctypes.c_double, i, ctypes.POINTER(u64),
                                ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)]
        if L.mks_isa() == 0:
            raise RuntimeError("the CPU has neither AVX2 nor AVX-512 IFMA")
        if L.mks_load_table(niels_table()) != 0:
            raise RuntimeError("mks_load_table failed")
        _lib = L
    return _lib


def use(isa_code: int) -> bool:
    """Switch to engine 512 (AVX-512 IFMA) or 2 (AVX2); False where the CPU lacks it."""
    L = lib()
    if L.mks_select(isa_code) != 0:
        return False
    assert L.mks_load_table(niels_table()) == 0
    return True


def niels_table() -> bytes:
    """The oracle's W = 12 comb table, 96 bytes per entry (canonical y+x, y-x, 2dxy)."""
    global _table
    if _table is not None:
        return _table
    L = O.lib()
    buf = ctypes.create_string_buffer(ROWS * COLS * 96)
    out = (ctypes.c_uint8 * 96)()
    for row in range(ROWS):
        for k in range(COLS):
            assert L.orc_comb_entry_niels(W, row, k, out) == 0
            ctypes.memmove(ctypes.addressof(buf) + 96 * (row * COLS + k), out, 96)
    _table = buf.raw
    return _table


def isa() -> int:
    """512 (AVX-512 IFMA), 2 (AVX2) or 0."""
    return lib().mks_isa()


def cpus() -> int:
    return lib().mks_cpu_count()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def derive(key: bytes, source: int, first: int, n: int) -> bytes:
    out = ctypes.create_string_buffer(32 * n)
    assert lib().mks_derive(key, source, first, n, out) == 0
    return out.raw


def search(key: bytes, source: int, first: int, n: int, needle: bytes, cs: bool, mode: int, threads: int = 0,
           cap: int = 1 << 16) -> list:
    """Exact hit indices of [first, first + n), in order (the oracle's search contract)."""
    threads = threads or max(1, min(cpus(), 64))
    hits = (ctypes.c_uint64 * cap)()
    c = lib().mks_search(key, source, first, n, needle, len(needle), int(cs), mode, hits, cap, threads)
    assert c >= 0, "mks_search failed"
    return list(hits[:min(c, cap)])


def bench(key: bytes, needle: bytes, cs: bool, mode: int, seconds: float, threads: int = 0, source: int = 0,
          base: int = 0) -> dict:
    """Timed search on `threads` pinned threads (default: every CPU of the affinity mask)."""
    threads = threads or cpus()
    keys, el, hits = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_uint64()
    assert lib().mks_bench(key, source, base, needle, len(needle), int(cs), mode, seconds, threads,
                           ctypes.byref(keys), ctypes.byref(el), ctypes.byref(hits)) == 0
    return {"keys": keys.value, "seconds": el.value, "hits": hits.value, "threads": threads,
            "keys_per_s": keys.value / el.value}
