"""Parity of the production setting (W = 24 comb, 256 keys per lane, 2 waves per SIMD):
the headline search itself, bulk derivation and the comb table, against the CPU oracle.

The headline workload is a secure-mode (ChaCha20 seed) search with a 6-character
case-sensitive suffix. A random suffix has about 0.06 expected hits in 2^32 candidates,
so the test plants one: it derives the key at a chosen index, takes the last 6
characters of its text as the needle and searches a 2^24-candidate window around it.
The reference's own bulk check of derived keys is src/service/gpu.rs:2077-2107; its
search kernel is src/service/gpu/kernel.cu:1996-2032.
"""
import os
import random
import sys

import pytest

import meikey_amd as M
from meikey_amd import workloads as W
from tests import textbook as T

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu_default():
    say("open gpu at the defaults (W=24, 256 keys per lane, 2 waves per SIMD)")
    g = M.Gpu(0)
    opt = g.info().options
    assert (opt.width, opt.batch, opt.blocks) == (24, 256, 2)
    yield g
    g.close()


def test_headline_planted_suffix_hit_set(gpu_default, oracle):
    """Config 2's search at the defaults: the master key of random.Random(2), a 6-char
    case-sensitive suffix planted at index i = 3 * 2^30 + 777, a 2^24-candidate window
    around i. The full hit set (index, seed, public key) equals the oracle's and holds i;
    the seeded headline suffix over the same window agrees with the oracle too."""
    g = gpu_default
    key, seeded = W.single()
    target = 3 * (1 << 30) + 777
    pk = g.derive(key, M.SEED_CHACHA20, target, 1)
    assert pk == oracle.derive(key, 0, target, 1)
    needle = T.hay(pk)[-6:].encode()
    first, count = target - (1 << 23) + 4321, 1 << 24
    say(f"headline planted suffix {needle!r} window [{first}, +2^24)")
    hits, tries = g.batch_run([
        {"key": key, "seed_source": M.SEED_CHACHA20, "needle": needle, "fold": False, "mode": M.MODE_SUFFIX,
         "first": first, "count": count},
        {"key": key, "seed_source": M.SEED_CHACHA20, "needle": seeded.encode(), "fold": False,
         "mode": M.MODE_SUFFIX, "first": first, "count": count}])
    assert tries == [count, count]
    planted = [h for h in hits if h.search == 0]
    want = oracle.search(key, 0, first, count, needle, True, M.MODE_SUFFIX, threads=THREADS)
    assert target in want
    assert [h.index for h in planted] == want
    for h in planted:
        assert h.pk == oracle.derive(key, 0, h.index, 1)
        assert h.seed == oracle.seed(key, 0, h.index)
        assert T.hay(h.pk).endswith(needle.decode())
    say(f"headline planted hits={len(want)}")
    # the seeded suffix (about 2^-36 per key): compared over the first 2^22 of the window
    sub = 1 << 22
    want2 = oracle.search(key, 0, first, sub, seeded.encode(), True, M.MODE_SUFFIX, threads=THREADS)
    assert [h.index for h in hits if h.search == 1 and h.index < first + sub] == want2


@pytest.mark.parametrize("source", [0, 1])
def test_bulk_derivation_at_the_defaults(gpu_default, oracle, source):
    """2^18 consecutive public keys (x sign included) from the W = 24 table equal the
    oracle's (its own W = 12 comb), across the 2^32 index boundary for ChaCha20 seeds
    and across a wrapping 64-bit key tail for counter seeds."""
    rng = random.Random(140 + source)
    key = rng.randbytes(32)
    n = 1 << 18
    if source == 0:
        first = (1 << 32) - (n // 2)
    else:
        key = key[:24] + ((1 << 64) - n // 3).to_bytes(8, "little")
        first = 0
    say(f"derive {n} keys at the defaults source={source}")
    got, seeds = gpu_default.derive(key, source, first, n, seeds=True)
    assert got == oracle.derive(key, source, first, n, threads=THREADS)
    for i in (0, 1, n // 3, n // 2, n - 1):
        assert seeds[32 * i:32 * i + 32] == oracle.seed(key, source, first + i)


def _entry_point(words):
    """(x, y, third field) of a device table entry: the fields are (y+x)/2, (y-x)/2 and
    d*x*y (row 0: x*y), ten limbs each."""
    from tests.hostcheck import entry_limbs, value_of
    ypx, ymx, third = (value_of(entry_limbs(words, f)) for f in range(3))
    for v in (ypx, ymx, third):
        assert v < T.P  # every stored field is canonical
    return (ypx - ymx) % T.P, (ypx + ymx) % T.P, third


def test_w24_table_entries_closed_form(gpu_default, oracle):
    """Sampled entries of the production W = 24 table against the closed form
    k * 2^(24 * row + 3) * B (+ 2^254 * B on the top row), computed by the oracle's
    separate double-and-add (orc_comb_entry): the compressed point, the curve equation
    and the third field d*x*y (x*y on row 0)."""
    g = gpu_default
    info = g.info()
    rows, cols = info.rows, info.cols
    assert (rows, cols) == (11, (1 << 23) + 1)
    rng = random.Random(141)
    picks = [(0, 0), (0, 1), (0, cols - 1), (1, 1), (rows - 1, 0), (rows - 1, 1), (rows - 1, 2048),
             (rows - 1, cols - 1), (rows - 2, cols - 1), (5, 1 << 22)]
    picks += [(rng.randrange(rows), rng.randrange(cols)) for _ in range(30)]
    for row, k in picks:
        x, y, third = _entry_point(g.table_words(row, k))
        assert (y * y - x * x - 1 - T.D * x * x % T.P * y * y) % T.P == 0, (row, k)
        assert int.to_bytes(y | ((x & 1) << 255), 32, "little") == oracle.comb_entry(24, row, k), (row, k)
        assert third == (x * y % T.P if row == 0 else T.D * x % T.P * y % T.P), (row, k)


@pytest.mark.parametrize("source,limbs", [(M.SEED_CHACHA20, M.LIMBS_R32), (M.SEED_COUNTER, M.LIMBS_K26U)],
                         ids=["secure-8", "low-security-10ku"])
def test_full_range_hit_sets_of_two_field_layouts(gpu_default, oracle, source, limbs):
    """The headline range at its full size: all 2^32 candidates of the headline key
    (ChaCha20 seeds, as the headline search; counter seeds, as the low-security one),
    searched for a 3-char case-sensitive suffix and a 4-char case-insensitive substring,
    by the default ten-limb kernel ("10u") and by a layout whose field arithmetic shares
    no code with it (the eight-word "8"; the Karatsuba "10ku"). The two complete hit sets
    (index, seed, public key) are equal; every 97th hit is re-derived by the oracle and
    matched as text; the suffix count lies within 6 sigma of 2^32 / 64^3 = 16384
    (sigma 128)."""
    key, _ = W.single()
    specs = [{"key": key, "seed_source": source, "needle": b"Zq9", "fold": False, "mode": M.MODE_SUFFIX,
              "first": 0, "count": 1 << 32},
             {"key": key, "seed_source": source, "needle": b"meek", "fold": True, "mode": M.MODE_CONTAINS,
              "first": 0, "count": 1 << 32}]
    say(f"full 2^32 range, seed source {source}, on the default layout")
    hits, tries = gpu_default.batch_run(specs)
    assert tries == [1 << 32, 1 << 32]
    say(f"open gpu W=24 batch=256 limbs={limbs}")
    other = M.Gpu(0, width=24, batch=256, limbs=limbs)
    try:
        say(f"full 2^32 range, seed source {source}, on layout {limbs}")
        hits8, tries8 = other.batch_run(specs)
    finally:
        other.close()
    assert tries8 == tries
    suffix = [h for h in hits if h.search == 0]
    substring = [h for h in hits if h.search == 1]
    say(f"hits: suffix {len(suffix)}, substring {len(substring)}")
    assert abs(len(suffix) - 16384) <= 6 * 128
    assert len(substring) > 100000
    assert [(h.search, h.index, h.seed, h.pk) for h in hits] == [(h.search, h.index, h.seed, h.pk) for h in hits8]
    for h in suffix[::97] + substring[::97]:
        assert h.pk == oracle.derive(key, source, h.index, 1)
        hay = T.hay(h.pk)
        assert hay.endswith("Zq9") if h.search == 0 else "meek" in hay.lower()


def test_batched_workload_on_two_field_layouts(gpu_default, oracle):
    """The 64-search batched workload at its full size (16 each of case-insensitive and
    case-sensitive substrings, suffixes and regex DFAs, 2^26 candidates each, about 600 k
    hits) on the default ten-limb kernel and on the eight-word layout, whose field
    arithmetic shares no code with it: equal complete hit sets for every search. Every
    499th hit is re-derived by the oracle and matched against its pattern with Python's
    own string and re semantics."""
    import re
    specs, pats, keys = W.batched(1 << 26)
    say("batched workload on the default layout")
    hits, tries = gpu_default.batch_run(specs, as_array=True)
    assert tries == [1 << 26] * 64
    say("open gpu W=24 batch=256 limbs=8")
    other = M.Gpu(0, width=24, batch=256, limbs=M.LIMBS_R32)
    try:
        say("batched workload on the eight-word layout")
        hits8, tries8 = other.batch_run(specs, as_array=True)
    finally:
        other.close()
    assert tries8 == tries
    assert len(hits) > 500000 and hits.tobytes() == hits8.tobytes()
    counts = [int((hits["search"] == i).sum()) for i in range(64)]
    assert sum(c > 0 for c in counts) >= 48, counts  # 51: the long case-sensitive patterns find none
    for h in hits[::499]:
        i, index, pk = int(h["search"]), int(h["index"]), bytes(h["pk"])
        assert pk == oracle.derive(keys[i], 0, index, 1)
        hay, p = T.hay(pk), pats[i]
        if p.mode == M.MODE_REGEX:
            ok = re.search(p.text, hay, re.ASCII | (0 if p.case_sensitive else re.IGNORECASE)) is not None
        else:
            h2, n2 = (hay, p.text) if p.case_sensitive else (hay.lower(), p.text.lower())
            ok = h2.endswith(n2) if p.mode == M.MODE_SUFFIX else n2 in h2
        assert ok, (i, p.text, hay)
