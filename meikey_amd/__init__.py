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
LIMBS_R25 = 10   # "10": ten 25.5-bit limbs, packed 96-byte table entries
LIMBS_R25U = 11  # "10u": ten 25.5-bit limbs, 128-byte entries of limbs
LIMBS_R32 = 8    # "8": eight 32-bit words, packed entries
LIMBS_K26 = 26   # "10k": ten 26-bit limbs with Karatsuba products, packed entries
LIMBS_K26U = 27  # "10ku": ten 26-bit limbs with Karatsuba products, 128-byte entries of limbs
ENTRY_WORDS = {LIMBS_R25: 24, LIMBS_R25U: 32, LIMBS_R32: 24, LIMBS_K26: 24, LIMBS_K26U: 32}
# --gpu-limbs codes (src/controller/cli.rs:105 of the reference) to mk_options.limbs
LIMBS_BY_CODE = {"10": LIMBS_R25, "10u": LIMBS_R25U, "8": LIMBS_R32, "10k": LIMBS_K26, "10ku": LIMBS_K26U}

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
    "mk_keys_per_iter", "mk_prepare", "mk_launch_run", "mk_free_launch", "mk_derive", "mk_derive_noalloc",
    "mk_derive_seeds",
    "mk_table_read", "mk_batch_run", "mk_tile_keys", "mk_last_timing",
    "mk_pattern_new", "mk_pattern_last_error", "mk_pattern_free", "mk_pattern_needle", "mk_pattern_needle_bytes",
    "mk_pattern_probability", "mk_pattern_dfa_states", "mk_pattern_is_match",
    "mk_server_default_options", "mk_server_open", "mk_server_submit", "mk_server_step", "mk_server_poll",
    "mk_server_cancel", "mk_server_stats", "mk_server_now", "mk_server_close",
    "mk_search", "mk_openssh_public", "mk_openssh_private", "mk_fastest", "mk_fastest_staged",
)

PATTERN_ERRORS = {1: "Empty", 2: "InvalidChar", 3: "TooLong", 4: "Regex", 5: "TooComplex", 6: "MatchesEmpty",
                  7: "NeverMatches"}


class MeikeyError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.message = message


class Options(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("batch", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("blocks", ctypes.c_uint32), ("lockstep", ctypes.c_uint32), ("pool", ctypes.c_uint32),
                ("lds_table", ctypes.c_uint32)]


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


class ServerOptions(ctypes.Structure):
    _fields_ = [("max_slots", ctypes.c_uint32), ("use_graph", ctypes.c_uint32), ("dfa_budget", ctypes.c_uint64),
                ("step_keys", ctypes.c_uint64), ("active_per_step", ctypes.c_uint32), ("overlap", ctypes.c_uint32),
                ("unit_keys", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("request", ctypes.c_uint64), ("status", ctypes.c_int32), ("pad", ctypes.c_uint32),
                ("tries", ctypes.c_uint64), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("hit", HitRec)]


class TuneResult(ctypes.Structure):
    _fields_ = [("options", Options), ("keys_per_s", ctypes.c_double), ("status", ctypes.c_int32),
                ("pad", ctypes.c_uint32)]


class Stats(ctypes.Structure):
    _fields_ = [("tries", ctypes.c_uint64), ("elapsed_s", ctypes.c_double)]


PROGRESS_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.POINTER(Stats))


class SearchConfig(ctypes.Structure):
    _fields_ = [("pattern", ctypes.c_void_p), ("seed_source", ctypes.c_int32), ("bench", ctypes.c_uint32),
                ("time_limit_s", ctypes.c_double), ("progress_interval_s", ctypes.c_double),
                ("target_launch_s", ctypes.c_double), ("cancel", ctypes.POINTER(ctypes.c_int32)),
                ("key", ctypes.c_char_p)]


class Outcome(ctypes.Structure):
    _fields_ = [("found", ctypes.c_int32), ("pad", ctypes.c_uint32), ("stats", Stats), ("launches", ctypes.c_uint64),
                ("index", ctypes.c_uint64), ("key", ctypes.c_uint8 * 32), ("seed", ctypes.c_uint8 * 32),
                ("pk", ctypes.c_uint8 * 32)]


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
    L.mk_derive_noalloc.argtypes = [P, ctypes.c_char_p, I32, U64, U32, P, P]
    L.mk_derive_seeds.argtypes = [P, ctypes.c_char_p, U32, P]
    L.mk_table_read.argtypes = [P, U32, U32, U32, P]
    L.mk_last_timing.argtypes = [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    L.mk_pattern_new.argtypes = [ctypes.c_char_p, U32, I32, U32, ctypes.POINTER(P), ctypes.POINTER(I32)]
    L.mk_pattern_last_error.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    L.mk_pattern_last_error.restype = None
    L.mk_pattern_free.argtypes = [P]
    L.mk_pattern_needle.argtypes = [P, ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8)), ctypes.POINTER(U32),
                                    ctypes.POINTER(U32), ctypes.POINTER(U32)]
    L.mk_pattern_needle_bytes.argtypes = [P]
    L.mk_pattern_needle_bytes.restype = ctypes.c_size_t
    L.mk_pattern_probability.argtypes = [P]
    L.mk_pattern_probability.restype = ctypes.c_double
    L.mk_pattern_dfa_states.argtypes = [P]
    L.mk_pattern_dfa_states.restype = U32
    L.mk_pattern_is_match.argtypes = [P, ctypes.c_char_p]
    L.mk_server_default_options.argtypes = [ctypes.POINTER(ServerOptions)]
    L.mk_server_default_options.restype = None
    L.mk_server_open.argtypes = [P, ctypes.POINTER(ServerOptions), ctypes.POINTER(P)]
    L.mk_server_submit.argtypes = [P, ctypes.POINTER(SearchSpec), ctypes.POINTER(U64)]
    L.mk_server_step.argtypes = [P, ctypes.POINTER(U32), ctypes.POINTER(U32)]
    L.mk_server_poll.argtypes = [P, ctypes.POINTER(ServerResult), U32, ctypes.POINTER(U32)]
    L.mk_server_cancel.argtypes = [P, U64]
    L.mk_server_stats.argtypes = [P, ctypes.POINTER(U64), ctypes.POINTER(U64)]
    L.mk_server_now.argtypes = [P]
    L.mk_server_now.restype = ctypes.c_double
    L.mk_server_close.argtypes = [P]
    L.mk_search.argtypes = [P, ctypes.POINTER(SearchConfig), PROGRESS_FN, P, ctypes.POINTER(Outcome)]
    L.mk_openssh_public.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t,
                                    ctypes.POINTER(ctypes.c_size_t)]
    L.mk_openssh_private.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, U32, ctypes.c_char_p,
                                     ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    L.mk_fastest.argtypes = [I32, ctypes.POINTER(Options), U32, ctypes.POINTER(TuneResult), ctypes.POINTER(Options)]
    L.mk_fastest_staged.argtypes = [I32, ctypes.POINTER(Options), ctypes.POINTER(TuneResult), U32, ctypes.POINTER(U32)]
    L.mk_batch_run.argtypes = [P, ctypes.POINTER(SearchSpec), U32, ctypes.POINTER(HitRec), U32,
                               ctypes.POINTER(U32), ctypes.POINTER(U64)]
    _lib = L
    return L


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().mk_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def _check(status: int) -> None:
    if status != 0:
        raise MeikeyError(status, last_error())


def openssh_public(pk: bytes, comment: str = "") -> str:
    """authorized_keys line of a public key (VanityKey::public_openssh)."""
    n = ctypes.c_size_t(0)
    c = comment.encode()
    _check(lib().mk_openssh_public(pk, c, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(lib().mk_openssh_public(pk, c, buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def openssh_private(seed: bytes, pk: bytes, comment: str = "", checkint: Optional[int] = None) -> str:
    """Unencrypted openssh-key-v1 private key text (private_openssh)."""
    if checkint is None:
        checkint = int.from_bytes(os.urandom(4), "little")
    n = ctypes.c_size_t(0)
    c = comment.encode()
    _check(lib().mk_openssh_private(seed, pk, c, checkint, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _check(lib().mk_openssh_private(seed, pk, c, checkint, buf, len(buf), ctypes.byref(n)))
    return buf.value.decode()


def fastest(ordinal: int, candidates):
    """Measure each (width, batch, blocks[, limbs[, lockstep[, pool]]]) candidate: (best
    Options, rows as fastest_staged). Limbs default to the library's, lockstep and pool
    to 0."""
    lim = default_options().limbs
    opts = []
    for c in candidates:
        w, b, k = c[:3]
        opts.append(Options(w, b, c[3] if len(c) > 3 else lim, k, c[4] if len(c) > 4 else 0,
                            c[5] if len(c) > 5 else 0))
    cands = (Options * len(opts))(*opts)
    res = (TuneResult * len(candidates))()
    best = Options()
    _check(lib().mk_fastest(ordinal, cands, len(candidates), res, ctypes.byref(best)))
    rows = [(r.options.width, r.options.batch, r.options.blocks, r.keys_per_s, r.status) for r in res]
    return best, rows


def fastest_staged(ordinal: int = 0):
    """The staged autotuner: (best Options, [(width, batch, blocks, keys/s, status, limbs,
    lockstep, pool, lds_table), ...])."""
    best = Options()
    res = (TuneResult * 64)()
    n = ctypes.c_uint32(0)
    _check(lib().mk_fastest_staged(ordinal, ctypes.byref(best), res, 64, ctypes.byref(n)))
    rows = [(r.options.width, r.options.batch, r.options.blocks, r.keys_per_s, r.status, r.options.limbs,
             r.options.lockstep, r.options.pool, r.options.lds_table) for r in res[:n.value]]
    return best, rows


def default_options() -> Options:
    """mk_default_options: the library's default setting."""
    o = Options()
    lib().mk_default_options(ctypes.byref(o))
    return o


def device_count() -> int:
    n = ctypes.c_int(0)
    _check(lib().mk_device_count(ctypes.byref(n)))
    return n.value


def _fill_spec(dst: "SearchSpec", s: dict, keep: list) -> None:
    needle = s["needle"]
    keep.append(needle)
    dst.key[:] = list(s["key"])
    dst.seed_source = s.get("seed_source", SEED_CHACHA20)
    dst.mode = s.get("mode", MODE_CONTAINS)
    dst.fold = int(s.get("fold", False))
    dst.needle_len = s.get("needle_len", len(needle) // 130 if dst.mode == MODE_REGEX else len(needle))
    dst.needle = needle
    dst.stop_on_hit = int(s.get("stop_on_hit", False))
    dst.first = s.get("first", 0)
    dst.count = s.get("count", 0)


HIT_DTYPE = None


def hit_dtype():
    """numpy dtype of mk_hit_rec (index u64, search u32, pad u32, seed[32], pk[32])."""
    global HIT_DTYPE
    if HIT_DTYPE is None:
        import numpy as np
        HIT_DTYPE = np.dtype([("index", "<u8"), ("search", "<u4"), ("pad", "<u4"), ("seed", "u1", 32), ("pk", "u1", 32)])
        assert HIT_DTYPE.itemsize == ctypes.sizeof(HitRec)
    return HIT_DTYPE


def _hits_array(recs, n: int):
    import numpy as np
    return np.frombuffer(recs, dtype=hit_dtype(), count=n).copy()


def _hits(recs, n: int) -> List[Hit]:
    return [Hit(recs[i].index, recs[i].search, bytes(recs[i].seed), bytes(recs[i].pk)) for i in range(n)]


class PatternError(ValueError):
    def __init__(self, kind: int, message: str):
        super().__init__(f"{PATTERN_ERRORS.get(kind, kind)}: {message}")
        self.kind = PATTERN_ERRORS.get(kind, str(kind))
        self.message = message


class Pattern:
    """A validated search pattern (the reference's Pattern::with_mode): substring, suffix
    or regular expression, case-sensitive or not. `needle()` gives what the kernels take."""

    def __init__(self, text: str, case_sensitive: bool = False, mode: int = MODE_CONTAINS):
        raw = text.encode("utf-8")
        h = ctypes.c_void_p()
        kind = ctypes.c_int(0)
        st = lib().mk_pattern_new(raw, len(raw), int(case_sensitive), mode, ctypes.byref(h), ctypes.byref(kind))
        if st != 0:
            buf = ctypes.create_string_buffer(1024)
            lib().mk_pattern_last_error(buf, len(buf))
            raise PatternError(kind.value, buf.value.decode(errors="replace"))
        self._h = h
        self.text, self.case_sensitive, self.mode = text, case_sensitive, mode

    def needle(self):
        """(needle bytes, needle_len, fold, mode) for Gpu.prepare / Gpu.batch_run."""
        ptr = ctypes.POINTER(ctypes.c_uint8)()
        n, fold, mode = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        _check(lib().mk_pattern_needle(self._h, ctypes.byref(ptr), ctypes.byref(n), ctypes.byref(fold),
                                       ctypes.byref(mode)))
        size = lib().mk_pattern_needle_bytes(self._h)
        return bytes(ptr[:size]), n.value, bool(fold.value), mode.value

    def spec(self, key: bytes, first: int, count: int, seed_source: int = SEED_CHACHA20, stop_on_hit=False) -> dict:
        needle, n, fold, mode = self.needle()
        return {"key": key, "seed_source": seed_source, "needle": needle, "needle_len": n, "fold": fold,
                "mode": mode, "first": first, "count": count, "stop_on_hit": stop_on_hit}

    @property
    def probability(self) -> float:
        return lib().mk_pattern_probability(self._h)

    @property
    def expected_tries(self) -> float:
        return 1.0 / self.probability

    @property
    def dfa_states(self) -> int:
        return lib().mk_pattern_dfa_states(self._h)

    def is_match(self, b64: str) -> bool:
        r = lib().mk_pattern_is_match(self._h, b64.encode())
        if r < 0:
            raise ValueError("bad base64 text")
        return bool(r)

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().mk_pattern_free(self._h)
                self._h = None
        except Exception:
            pass


class Launch:
    """A prepared single search (the reference's Engine::prepare result)."""

    def __init__(self, gpu: "Gpu", handle: ctypes.c_void_p):
        self.gpu = gpu
        self._h = handle

    def run(self, base: int, iters: int, cap: int = 1 << 20) -> List[Hit]:
        """Engine::launch: exact hits of [base, base + keys_per_iter * iters), ascending."""
        recs = self.gpu._hit_buffer(cap)
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

    def __init__(self, ordinal: int = 0, width: int = 24, batch: int = 256, blocks: int = 2,
                 limbs: Optional[int] = None, lockstep: int = 0, pool: int = 0, lds_table: int = 0):
        if limbs is None:
            limbs = default_options().limbs
        self.options = Options(width, batch, limbs, blocks, lockstep, pool, lds_table)
        h = ctypes.c_void_p()
        _check(lib().mk_open(ordinal, ctypes.byref(self.options), ctypes.byref(h)))
        self._h = h
        self._hits = None

    def _hit_buffer(self, cap: int):
        """A hit-record array of at least `cap` entries, allocated once per device handle
        (ctypes zero-fills new arrays, which would cost milliseconds per launch)."""
        cap = max(cap, 1)
        if self._hits is None or len(self._hits) < cap:
            self._hits = (HitRec * cap)()
        return self._hits

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

    def derive_noalloc(self, key: bytes, seed_source: int, first: int, n: int, seeds: bool = False):
        """mk_derive_noalloc: derive without any device allocation (n <= 65536)."""
        pks = ctypes.create_string_buffer(max(32 * n, 1))
        sd = ctypes.create_string_buffer(max(32 * n, 1)) if seeds else None
        _check(lib().mk_derive_noalloc(self._h, key, seed_source, first, n, pks, sd))
        if seeds:
            return pks.raw[: 32 * n], sd.raw[: 32 * n]
        return pks.raw[: 32 * n]

    def derive_seeds(self, seeds: bytes) -> bytes:
        n = len(seeds) // 32
        out = ctypes.create_string_buffer(max(32 * n, 1))
        _check(lib().mk_derive_seeds(self._h, seeds, n, out))
        return out.raw[: 32 * n]

    def search(self, pattern: Optional["Pattern"], seed_source: int = SEED_CHACHA20, bench: bool = False,
               time_limit: float = 0.0, progress=None, progress_interval: float = 5.0, target_launch: float = 0.2,
               cancel=None, key: Optional[bytes] = None) -> Outcome:
        """Searcher::run: search until an exact hit (or, with bench, until time_limit).
        progress(tries, elapsed) is called every progress_interval seconds; cancel is an
        optional ctypes.c_int32 that stops the search when set."""
        cfg = SearchConfig(pattern._h if pattern else None, seed_source, int(bench), time_limit,
                           progress_interval, target_launch,
                           ctypes.pointer(cancel) if cancel is not None else None, key)
        cb = PROGRESS_FN((lambda user, st: progress(st.contents.tries, st.contents.elapsed_s)) if progress
                         else (lambda user, st: None))
        out = Outcome()
        _check(lib().mk_search(self._h, ctypes.byref(cfg), cb, None, ctypes.byref(out)))
        return out

    def last_timing(self):
        """(search kernel ms, confirm kernel ms) of the last launch, from HIP events."""
        a, b = ctypes.c_float(0), ctypes.c_float(0)
        _check(lib().mk_last_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def table_words(self, row: int, k: int, count: int = 1, as_array: bool = False):
        """Raw table entries from (row, k) on, 32 words each for LIMBS_R25U and LIMBS_K26U, 24 for
        the packed layouts (a list, or with as_array a numpy array of shape (count, words))."""
        n = ENTRY_WORDS[self.options.limbs]
        buf = (ctypes.c_uint32 * (n * count))()
        _check(lib().mk_table_read(self._h, row, k, count, buf))
        if as_array:
            import numpy as np
            return np.frombuffer(buf, dtype=np.uint32).reshape(count, n).copy()
        return list(buf)

    def batch_run(self, specs: Sequence[dict], cap: int = 1 << 20, as_array: bool = False, out=None):
        """Run many searches in one launch. Each spec: key, seed_source, mode, fold, needle,
        first, count, stop_on_hit (optional). Returns (hits, tries per search); with
        as_array the hits are one numpy structured array (fields of hit_dtype()). With out, a
        C-contiguous numpy array of hit_dtype() that the caller keeps between calls, the
        records are written straight into it and the hits returned are the view out[:n]: no
        copy, valid until out is reused. More hits than len(out) is MK_ERR_OVERFLOW, as for
        cap."""
        n = len(specs)
        arr = (SearchSpec * max(n, 1))()
        keep = []
        for i, s in enumerate(specs):
            _fill_spec(arr[i], s, keep)
        if out is not None:
            if (out.ndim != 1 or out.dtype != hit_dtype() or not out.flags["C_CONTIGUOUS"]
                    or not out.flags["WRITEABLE"]):
                raise ValueError("out must be a writeable, C-contiguous, one-dimensional array of hit_dtype()")
            cap = len(out)
            recs = ctypes.cast(out.ctypes.data, ctypes.POINTER(HitRec)) if cap else None
        else:
            recs = self._hit_buffer(cap)
        nh = ctypes.c_uint32(0)
        tries = (ctypes.c_uint64 * max(n, 1))()
        _check(lib().mk_batch_run(self._h, arr, n, recs, cap, ctypes.byref(nh), tries))
        got = min(nh.value, cap)
        if out is not None:
            return out[:got], list(tries)[:n]
        return (_hits_array(recs, got) if as_array else _hits(recs, got)), list(tries)[:n]

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


@dataclass(frozen=True)
class ServedResult:
    request: int
    status: int  # 1 found, 0 budget spent, -1 invalid or cancelled
    tries: int
    submit_s: float
    admit_s: float
    done_s: float
    index: int
    seed: bytes
    pk: bytes


class Server:
    """Continuous batching of search requests on one Gpu (mk_server_* of the C ABI).

    unit_keys is the keys per lane of a step's claim unit (64, 128, 192 or 256): 128 serves
    the most keys/s; 64 with half the step_keys halves a step, and with it the median time
    to a hit, at about 1% of the served keys/s."""

    def __init__(self, gpu: Gpu, max_slots: int = 1024, use_graph: bool = True, dfa_budget: int = 64 << 20,
                 step_keys: int = 1 << 26, active_per_step: int = 16, overlap: bool = True, unit_keys: int = 128):
        self.gpu = gpu
        opt = ServerOptions(max_slots, int(use_graph), dfa_budget, step_keys, active_per_step, int(overlap),
                            unit_keys, 0)
        h = ctypes.c_void_p()
        _check(lib().mk_server_open(gpu._h, ctypes.byref(opt), ctypes.byref(h)))
        self._h = h

    def submit(self, spec: dict) -> int:
        arr = SearchSpec()
        keep = []
        _fill_spec(arr, spec, keep)
        rid = ctypes.c_uint64(0)
        _check(lib().mk_server_submit(self._h, ctypes.byref(arr), ctypes.byref(rid)))
        return rid.value

    def step(self):
        """One device step; returns (active, queued) before the step."""
        a, q = ctypes.c_uint32(0), ctypes.c_uint32(0)
        _check(lib().mk_server_step(self._h, ctypes.byref(a), ctypes.byref(q)))
        return a.value, q.value

    def poll(self, cap: int = 4096) -> List[ServedResult]:
        out = (ServerResult * cap)()
        n = ctypes.c_uint32(0)
        _check(lib().mk_server_poll(self._h, out, cap, ctypes.byref(n)))
        return [ServedResult(r.request, r.status, r.tries, r.submit_s, r.admit_s, r.done_s, r.hit.index,
                             bytes(r.hit.seed), bytes(r.hit.pk)) for r in out[:n.value]]

    def cancel(self, request: int) -> None:
        _check(lib().mk_server_cancel(self._h, request))

    def now(self) -> float:
        return lib().mk_server_now(self._h)

    def stats(self):
        steps, keys = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(lib().mk_server_stats(self._h, ctypes.byref(steps), ctypes.byref(keys)))
        return steps.value, keys.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mk_server_close(self._h)
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
