"""Parity of the gfx950 kernels with the CPU oracle, through the C ABI.

Every comparison is bit-exact: public keys (x sign included), table entries and full
hit sets. Sizes are chosen so the oracle finishes in seconds; the 2^24 config-1 range
is compared in full.
"""
import json
import math
import os
import random
import sys

import pytest

import meikey_amd as M
from tests import textbook as T

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "golden.json")))


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    say("open gpu W=16 batch=32")
    g = M.Gpu(0, width=16, batch=32)
    yield g
    g.close()


@pytest.fixture(scope="module")
def gpu_small():
    say("open gpu W=12 batch=4 blocks=1")
    g = M.Gpu(0, width=12, batch=4, blocks=1)
    yield g
    g.close()


def test_rfc8032_known_answers(gpu):
    from tests.test_oracle import rfc8032_vectors
    vecs = rfc8032_vectors()
    seeds = b"".join(v[1] for v in vecs)
    pks = gpu.derive_seeds(seeds)
    assert [pks[32 * i:32 * i + 32] for i in range(len(vecs))] == [v[2] for v in vecs]


def test_derive_golden(gpu):
    for case in GOLDEN["derive"]:
        key = bytes.fromhex(case["key"])
        for idx, pk in case["pks"]:
            assert gpu.derive(key, case["source"], idx, 1) == bytes.fromhex(pk), (case["source"], idx)


@pytest.mark.parametrize("source", [0, 1])
def test_derive_bulk_vs_oracle(gpu, oracle, source):
    key = random.Random(40 + source).randbytes(32)
    n = 1 << 18
    first = (1 << 33) - 1000 if source == 0 else 12345
    say(f"derive {n} keys source={source}")
    got, seeds = gpu.derive(key, source, first, n, seeds=True)
    want = oracle.derive(key, source, first, n)
    assert got == want
    for i in (0, 1, n - 1):
        assert seeds[32 * i:32 * i + 32] == oracle.seed(key, source, first + i)


def test_table_entries_vs_oracle(gpu, oracle):
    from tests.hostcheck import entry_limbs, value_of
    w = 16
    rows, cols = 16, (1 << 15) + 1
    rng = random.Random(41)
    picks = [(0, 0), (0, 1), (rows - 1, 0), (rows - 1, cols - 1), (7, cols - 1)]
    picks += [(rng.randrange(rows), rng.randrange(cols)) for _ in range(20)]
    for row, k in picks:
        words = gpu.table_words(row, k)
        n = oracle.comb_entry_niels(w, row, k)
        half = T.inv(2)  # the device table holds every field halved
        assert value_of(entry_limbs(words, 0)) == int.from_bytes(n[0:32], "little") * half % T.P, (row, k)
        assert value_of(entry_limbs(words, 1)) == int.from_bytes(n[32:64], "little") * half % T.P, (row, k)
        third = value_of(entry_limbs(words, 2))
        xy2d = int.from_bytes(n[64:96], "little")
        assert third == (xy2d * T.inv(2 * T.D) % T.P if row == 0 else xy2d * half % T.P), (row, k)


def test_device_keys_equal_the_reference_kernel(gpu):
    """The reference's own kernel (kernel.cu compiled for the host into oracle/_ref)
    derives the same public keys, x sign included, as the device: over the device's
    W = 16 table repacked into the reference's layout, both seed sources, at index
    ranges that start at 0, straddle 2^32 and sit past 2^40."""
    import ctypes
    import struct
    import numpy as np
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        pytest.skip("the reference kernel's host build (oracle/_ref) is absent")
    info = gpu.info()
    assert info.options.width == 16
    rows = [gpu.table_words(r, 0, info.cols, as_array=True) for r in range(info.rows)]
    table = O.ref_table_from_device(np.concatenate(rows))
    batch = R.meikey_test_batch()
    rng = random.Random(47)
    for src in (0, 1):
        key = rng.randbytes(32)
        kw = (ctypes.c_uint32 * 8)(*struct.unpack("<8I", key))
        for first in (0, 12345, (1 << 32) - batch // 2, (1 << 40) + 7):
            say(f"reference derive src={src} first={first}")
            out = ctypes.create_string_buffer(32 * batch)
            R.meikey_test_derive_batch(table, 16, kw, src, first, out)
            assert out.raw == gpu.derive(key, src, first, batch), (src, first)


def test_search_golden(gpu):
    for case in GOLDEN["search"]:
        key = bytes.fromhex(case["key"])
        hits, tries = gpu.batch_run([{
            "key": key, "seed_source": case["source"], "needle": case["needle"].encode(),
            "fold": not case["case_sensitive"], "mode": case["mode"], "first": case["first"],
            "count": case["count"]}])
        end = case["first"] + case["count"]
        got = [h.index for h in hits if h.index < end]
        assert got == case["hits"]


def _spec(key, source, needle, cs, mode, first, count, **kw):
    d = {"key": key, "seed_source": source, "needle": needle, "fold": not cs, "mode": mode, "first": first,
         "count": count}
    d.update(kw)
    return d


@pytest.mark.parametrize("source,needle,cs,mode", [
    (0, b"ab", False, M.MODE_CONTAINS),
    (0, b"Q9", True, M.MODE_SUFFIX),
    (1, b"k+", False, M.MODE_CONTAINS),
    (0, b"z", True, M.MODE_SUFFIX),
])
def test_search_vs_oracle(gpu, oracle, source, needle, cs, mode):
    key = random.Random(int.from_bytes(needle, "little")).randbytes(32)
    first = 1 << 20
    count = 1 << 20
    say(f"search {needle} source={source} mode={mode}")
    hits, tries = gpu.batch_run([_spec(key, source, needle, cs, mode, first, count)])
    assert tries[0] >= count
    want = oracle.search(key, source, first, tries[0], needle, cs, mode)
    assert [h.index for h in hits] == want
    pks = {h.index: h.pk for h in hits}
    for i in want[:32]:
        assert pks[i] == oracle.derive(key, source, i, 1)
        assert [h.seed for h in hits if h.index == i][0] == oracle.seed(key, source, i)


@pytest.mark.parametrize("batch", [256, 192, 128, 200])
def test_tail_units_match_whole_tiles(gpu, oracle, batch):
    """The last two tiles per wave of a launch are cut into 64-key units claimed from a
    second counter (4, 3 and 2 units at 256, 192 and 128 keys per lane; a batch that is no
    multiple of 64 keeps whole tiles), and the final round of units is claimed as halves.
    The launch spans more than two tiles per wave of the grid, so it runs through the
    whole-tile claims, the units and the halves. Hit sets over the same ranges equal those
    of whole 32-key tiles, and the oracle's on the smaller searches."""
    say(f"open gpu W=16 batch={batch} blocks=1")
    big = M.Gpu(0, width=16, batch=batch, blocks=1)
    try:
        tile = 64 * batch
        step = tile * 2048 // math.gcd(tile, 2048)  # whole tiles of both kernels

        def span(n):
            return -(-n // step) * step

        waves = big.info().grid_waves
        rng = random.Random(77 + batch)
        specs = []
        for count, needle, cs, mode in [(span((2 * waves + 3) * tile), b"Kq7", True, M.MODE_CONTAINS),
                                        (span(1 << 20), b"zz", False, M.MODE_SUFFIX),
                                        (span(5 * tile), b"mk", False, M.MODE_CONTAINS)]:
            specs.append(_spec(rng.randbytes(32), rng.choice([0, 1]), needle, cs, mode, rng.getrandbits(40), count))
        hits, tries = big.batch_run(specs)
        assert tries == [s["count"] for s in specs]
        ref, _ = gpu.batch_run(specs)
        assert [(h.search, h.index, h.pk) for h in hits] == [(h.search, h.index, h.pk) for h in ref]
        for i in (1, 2):
            s = specs[i]
            want = oracle.search(s["key"], s["seed_source"], s["first"], s["count"], s["needle"], not s["fold"],
                                 s["mode"])
            assert [h.index for h in hits if h.search == i] == want and want
    finally:
        big.close()


def test_low_security_counter_mode(gpu, oracle):
    """Low-security mode (counter seeds through the per-search SHA-512 mid-state): exact
    hit sets across the 2^32 index boundary and across a key tail that wraps past 2^64,
    and every 2^24-th key of the 2^32 range re-derived against the oracle."""
    key = random.Random(3).randbytes(32)
    wrap = key[:24] + (2**64 - 3000).to_bytes(8, "little")
    for k, first in [(key, (1 << 32) - (1 << 19)), (wrap, 0)]:
        say(f"counter search first={first}")
        hits, tries = gpu.batch_run([_spec(k, M.SEED_COUNTER, b"mk7", False, M.MODE_CONTAINS, first, 1 << 20)])
        want = oracle.search(k, M.SEED_COUNTER, first, tries[0], b"mk7", False, M.MODE_CONTAINS)
        assert [h.index for h in hits] == want and len(want) > 0
        for h in hits[:16]:
            assert h.seed == oracle.seed(k, M.SEED_COUNTER, h.index)
            assert h.pk == oracle.derive(k, M.SEED_COUNTER, h.index, 1)
    for i in range(0, 1 << 32, 1 << 24):
        assert gpu.derive(key, M.SEED_COUNTER, i, 1) == oracle.derive(key, M.SEED_COUNTER, i, 1), i


@pytest.mark.parametrize("length", [32, 33, 40, 43])
def test_long_planted_needles(gpu, oracle, length):
    """Needles of 32 characters and more (the matcher's 64-bit shift-and states above 32;
    43 is the whole hay, sign character included), cut from the text of a planted key:
    substring and suffix, both case rules. Each window's full hit set is the oracle's and
    holds the planted index."""
    rng = random.Random(300 + length)
    specs, planted = [], []
    for mode, cs in [(M.MODE_CONTAINS, True), (M.MODE_CONTAINS, False), (M.MODE_SUFFIX, True),
                     (M.MODE_SUFFIX, False)]:
        key = rng.randbytes(32)
        source = rng.choice([0, 1])
        target = rng.getrandbits(40)
        hay = T.hay(oracle.derive(key, source, target, 1))
        if mode == M.MODE_SUFFIX:
            text = hay[-length:]
        else:
            o = rng.randrange(43 - length + 1)
            text = hay[o:o + length]
        if not cs:
            text = "".join(c.swapcase() if rng.random() < 0.5 else c for c in text)
        needle = text.encode() if cs else text.lower().encode()
        first = target - (1 << 15) - 7
        specs.append(_spec(key, source, needle, cs, mode, first, 1 << 16))
        planted.append(target)
    hits, tries = gpu.batch_run(specs)
    for i, s in enumerate(specs):
        want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"], s["mode"])
        assert planted[i] in want and [h.index for h in hits if h.search == i] == want, (length, i)


def test_sign_dependent_suffix(gpu, oracle):
    """Suffixes of 2-3 chars cover tail position 42, whose value holds the x sign; the
    exact hit set must follow the real sign."""
    key = random.Random(43).randbytes(32)
    for needle in (b"A", b"AB", b"gQ", b"Yw"):
        hits, tries = gpu.batch_run([_spec(key, 0, needle, True, M.MODE_SUFFIX, 0, 1 << 19)])
        want = oracle.search(key, 0, 0, tries[0], needle, True, 1)
        assert [h.index for h in hits] == want, needle


def test_engine_launch_ranges(gpu_small, oracle):
    """Engine::launch semantics: a launch covers [base, base + keys_per_iter * iters)."""
    g = gpu_small
    key = random.Random(44).randbytes(32)
    kpi = g.keys_per_iter
    l = g.prepare(key, M.SEED_CHACHA20, b"xy", True, M.MODE_CONTAINS)
    base = 3 * kpi + 17
    got = [h.index for h in l.run(base, 2)]
    want = oracle.search(key, 0, base, 2 * kpi, b"xy", False, 0)
    assert got == want and len(want) > 0
    # consecutive launches tile the index space
    a = [h.index for h in l.run(0, 1)] + [h.index for h in l.run(kpi, 1)]
    assert a == oracle.search(key, 0, 0, 2 * kpi, b"xy", False, 0)
    l.close()


def test_stop_on_hit_reports_a_true_hit(gpu_small, oracle):
    g = gpu_small
    key = random.Random(45).randbytes(32)
    l = g.prepare(key, M.SEED_CHACHA20, b"q", True, M.MODE_CONTAINS, stop_on_hit=True)
    hits = l.run(0, 4)
    assert len(hits) >= 1
    for h in hits:
        assert oracle.derive(key, 0, h.index, 1) == h.pk
        assert "q" in T.hay(h.pk).lower()  # fold = case-insensitive
    l.close()


def test_batched_equals_single(gpu, oracle):
    """Searches batched in one launch return the same hit sets as one launch each."""
    rng = random.Random(46)
    specs = []
    for i in range(8):
        key = rng.randbytes(32)
        needle = rng.choice([b"ab", b"Zz", b"9", b"k/"])
        mode = rng.choice([M.MODE_CONTAINS, M.MODE_SUFFIX])
        specs.append(_spec(key, i % 2, needle, i % 3 == 0, mode, rng.getrandbits(40), 64 * 32 * rng.randint(1, 40)))
    batched, tries = gpu.batch_run(specs)
    for i, s in enumerate(specs):
        single, t1 = gpu.batch_run([s])
        assert t1[0] == tries[i]
        assert [h.index for h in batched if h.search == i] == [h.index for h in single]
        want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"],
                             s["mode"])
        assert [h.index for h in single] == want


def test_batch_run_into_caller_buffer(gpu):
    """batch_run(out=buf) writes the records straight into the caller's array and returns
    the view buf[:n]: the same records as the copying as_array form. A buffer shorter than
    the hit list is an overflow, as a short cap is for the C entry point."""
    import numpy as np
    rng = random.Random(47)
    specs = [_spec(rng.randbytes(32), i % 2, b"ab", False, M.MODE_CONTAINS, rng.getrandbits(40), 64 * 32 * 20)
             for i in range(3)]
    want, tries = gpu.batch_run(specs, as_array=True)
    assert len(want) > 40
    buf = np.zeros(len(want) + 7, dtype=M.hit_dtype())
    got, t2 = gpu.batch_run(specs, out=buf)
    assert t2 == tries
    assert np.shares_memory(got, buf) and len(got) == len(want)
    assert got.tobytes() == want.tobytes()
    with pytest.raises(M.MeikeyError) as e:
        gpu.batch_run(specs, out=np.zeros(40, dtype=M.hit_dtype()))
    assert e.value.status == 7  # MK_ERR_OVERFLOW
    for bad in (np.zeros(8, dtype=np.uint64), np.zeros((4, 100), dtype=M.hit_dtype()),
                np.zeros(200, dtype=M.hit_dtype())[::2]):
        with pytest.raises(ValueError):
            gpu.batch_run(specs, out=bad)


def test_config1_full_hit_set(gpu, oracle):
    """Config 1: 2^24 ChaCha20 candidates of the master key from PRNG seed 1, a seeded
    4-char case-insensitive substring; every hit index and key equal to the oracle's."""
    rng = random.Random(1)
    key = rng.randbytes(32)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    needle = "".join(rng.choice(alpha) for _ in range(4)).lower().encode()
    n = 1 << 24
    say(f"config1 needle={needle}")
    hits, tries = gpu.batch_run([_spec(key, 0, needle, False, M.MODE_CONTAINS, 0, n)])
    assert tries[0] == n
    want = oracle.search(key, 0, 0, n, needle, False, 0, threads=16)
    assert [h.index for h in hits] == want
    for h in hits:
        assert "".join(T.hay(h.pk)).lower().find(needle.decode()) >= 0
    say(f"config1 hits={len(want)}")


@pytest.mark.parametrize("text,cs", [
    ("ab", True), ("^[A-P]a|z$", False), ("(ab){2,}", True), ("\\b9", True), ("[0-9][A-F]x", False),
    ("Q.$", True), ("[a-z]{2}[0-9]$", True), ("^[A-D][0-9]", True), ("a.b", False),
    ("(?i)abc", True), ("\\d{3}$", True), ("(ab|c(d|e))+f", True), ("[^A-Za-z]{4}", True), ("x{1,3}y$", False),
])
def test_regex_search_vs_oracle(gpu, oracle, text, cs):
    """DFA matching in the search kernel (LDS-resident table, sign fork at tail position
    42) against the oracle's DFA walk, and every hit against Python re on the real key."""
    import re
    p = M.Pattern(text, cs, M.MODE_REGEX)
    import zlib
    key = random.Random(zlib.crc32(text.encode()) * 2 + int(cs)).randbytes(32)
    first = 12345
    hits, tries = gpu.batch_run([p.spec(key, first, 1 << 19)])
    table = p.needle()[0]
    want = oracle.search(key, 0, first, tries[0], table, cs, M.MODE_REGEX)
    assert [h.index for h in hits] == want
    rx = re.compile(text, re.ASCII | (0 if cs else re.IGNORECASE))
    for h in hits[:50]:
        assert rx.search(T.hay(h.pk)) is not None


def test_full_scale_hits_are_exact(gpu, oracle):
    """2^30 candidates in one launch at the default 256 keys per lane: every reported hit is
    re-derived and re-matched on the CPU, and four random 2^16-key windows of the range
    have exactly the oracle's hit sets (so nothing is missed there either)."""
    say("open gpu defaults for 2^30 candidates")
    big = M.Gpu(0)
    try:
        key = random.Random(95).randbytes(32)
        needle = b"mk7q"
        hits, tries = big.batch_run([_spec(key, 0, needle, False, M.MODE_CONTAINS, 1 << 33, 1 << 30)],
                                    as_array=True)
        assert tries[0] == 1 << 30
        n = len(hits)
        per_position = 1.0
        for c in needle.decode():
            per_position *= 2 / 64 if c.isalpha() else 1 / 64  # case-insensitive letters: 2 of 64
        expect = (1 << 30) * (43 - len(needle) + 1) * per_position
        assert 0.7 * expect < n < 1.3 * expect, (n, expect)
        idx = hits["index"].tolist()
        assert idx == sorted(idx) and len(set(idx)) == n
        for i in range(0, n, max(1, n // 2000)):
            pk = oracle.derive(key, 0, idx[i], 1)
            assert pk == hits["pk"][i].tobytes()
            assert oracle.match(oracle.public_b64(pk), needle, False, M.MODE_CONTAINS)
        rng = random.Random(96)
        for _ in range(4):
            lo = (1 << 33) + rng.randrange((1 << 30) - (1 << 16))
            want = oracle.search(key, 0, lo, 1 << 16, needle, False, M.MODE_CONTAINS)
            assert [i for i in idx if lo <= i < lo + (1 << 16)] == want
    finally:
        big.close()


def test_overflow_is_an_error_not_a_short_list(gpu):
    """More candidates than the device lists hold (2^20) in a search that does not stop
    on its first hit is MK_ERR_OVERFLOW, never a silently truncated hit set; the same
    search with stop_on_hit succeeds."""
    key = random.Random(97).randbytes(32)
    spec = _spec(key, 0, b"A", False, M.MODE_CONTAINS, 0, 1 << 23)  # ~2^23 * 43/32 candidates
    with pytest.raises(M.MeikeyError) as e:
        gpu.batch_run([spec])
    assert e.value.status == 7
    hits, _ = gpu.batch_run([dict(spec, stop_on_hit=True)])
    assert len(hits) >= 1


def test_searcher_counter_mode_and_served_large_dfa(gpu, oracle):
    """The Searcher in low-security mode returns a true hit; a served request whose DFA
    (> 126 states) is walked from HBM returns the same key alone as the oracle finds."""
    key = random.Random(98).randbytes(32)
    out = gpu.search(M.Pattern("Qz", True, M.MODE_SUFFIX), M.SEED_COUNTER, key=key)
    assert out.found == 1 and bytes(out.pk) == oracle.derive(key, 1, out.index, 1)
    assert bytes(out.seed) == oracle.seed(key, 1, out.index)
    p = M.Pattern("(ab|cd|ef|gh).{0,10}(12|34|56)", True, M.MODE_REGEX)
    with M.Server(gpu, max_slots=4, use_graph=True, step_keys=1 << 22) as srv:
        rid = srv.submit(p.spec(key, 0, 0, M.SEED_CHACHA20, stop_on_hit=True))
        done = []
        for _ in range(400):
            srv.step()
            done += srv.poll()
            if done:
                break
        assert done and done[0].request == rid and done[0].status == 1
        r = done[0]
        first = oracle.search(key, 0, 0, r.index + 1, p.needle()[0], True, M.MODE_REGEX)
        assert first and first[0] == r.index and r.pk == oracle.derive(key, 0, r.index, 1)


def test_large_dfa_from_global_memory(gpu, oracle):
    """A DFA of more than 126 states does not fit the LDS staging buffers and is walked
    from HBM; hit sets still equal the oracle's."""
    p = M.Pattern("(ab|cd|ef|gh).{0,10}(12|34|56)", True, M.MODE_REGEX)
    assert p.dfa_states > 126
    key = random.Random(91).randbytes(32)
    hits, tries = gpu.batch_run([p.spec(key, 777, 1 << 20)])
    want = oracle.search(key, 0, 777, tries[0], p.needle()[0], True, M.MODE_REGEX)
    assert [h.index for h in hits] == want and len(want) > 50


def test_edge_inputs(gpu, oracle):
    """Empty batch, zero-count and empty-needle searches next to a normal one, and a
    43-character needle (the whole hay) that matches exactly one key."""
    assert gpu.batch_run([]) == ([], [])
    key = random.Random(92).randbytes(32)
    specs = [_spec(key, 0, b"Ab", True, M.MODE_CONTAINS, 0, 0),
             _spec(key, 0, b"", True, M.MODE_CONTAINS, 0, 1 << 16),
             _spec(key, 0, b"Ab", True, M.MODE_CONTAINS, 5, 1 << 16)]
    hits, tries = gpu.batch_run(specs)
    assert tries[0] == 0 and tries[1] >= 1 << 16
    assert {h.search for h in hits} <= {2}
    assert [h.index for h in hits] == oracle.search(key, 0, 5, tries[2], b"Ab", True, M.MODE_CONTAINS)
    target = 4242
    hay = T.hay(oracle.derive(key, 0, target, 1)).encode()
    assert len(hay) == 43
    hits, _ = gpu.batch_run([_spec(key, 0, hay, True, M.MODE_CONTAINS, target - 1000, 4096)])
    assert [h.index for h in hits] == [target]
    hits, _ = gpu.batch_run([_spec(key, 0, hay.lower(), False, M.MODE_SUFFIX, target - 1000, 4096)])
    assert [h.index for h in hits] == [target]


@pytest.mark.parametrize("source", [0, 1])
def test_index_range_across_2_64(gpu, oracle, source):
    """A range that starts just below 2^64 continues at index 0 (u64 indices, as the
    reference's index arithmetic wraps): the hit set equals the oracle's, hits on both
    sides of the wrap, listed in ascending index order."""
    key = random.Random(93 + source).randbytes(32)
    first = (1 << 64) - 3000
    hits, tries = gpu.batch_run([_spec(key, source, b"ab", False, M.MODE_CONTAINS, first, 6144)])
    want = oracle.search(key, source, first, tries[0], b"ab", False, M.MODE_CONTAINS)
    got = [h.index for h in hits]
    assert got == sorted(got)
    assert got == sorted(want)
    assert any(i >= first for i in got) and any(i < 4096 for i in got)


def test_one_key_per_lane_and_eight_waves(oracle):
    """The extremes of the launch shape: 1 key per lane (an inversion per key) and a grid
    sized for 8 waves per SIMD (more waves than fit, queued by the hardware)."""
    for batch, blocks in [(1, 1), (8, 8)]:
        say(f"open gpu W=10 batch={batch} blocks={blocks}")
        g = M.Gpu(0, width=10, batch=batch, blocks=blocks)
        try:
            key = random.Random(93 + batch).randbytes(32)
            assert g.derive(key, 1, 99, 512) == oracle.derive(key, 1, 99, 512)
            hits, tries = g.batch_run([_spec(key, 0, b"q", True, M.MODE_SUFFIX, 3, 1 << 17)])
            assert [h.index for h in hits] == oracle.search(key, 0, 3, tries[0], b"q", True, M.MODE_SUFFIX)
        finally:
            g.close()


def config4_specs(count):
    from meikey_amd import workloads as W
    return W.batched(count)


def _planted_twin(pat, key, index, oracle, rng):
    """A pattern of the same kind, case rule and length as `pat` that the key at `index`
    matches: a slice of that key's text (for a regex, the slice as an escaped literal,
    anchored where `pat` is)."""
    hay = T.hay(oracle.derive(key, 0, index, 1))
    if pat.mode == M.MODE_REGEX:
        n = 4
        if pat.text.startswith("^"):
            body, pre, post = hay[:n], "^", ""
        elif pat.text.endswith("$"):
            body, pre, post = hay[-n:], "", "$"
        else:
            o = rng.randrange(43 - n + 1)
            body, pre, post = hay[o:o + n], "", ""
        return M.Pattern(pre + body.replace("+", "\\+") + post, pat.case_sensitive, M.MODE_REGEX)
    n = len(pat.text)
    if pat.mode == M.MODE_SUFFIX:
        return M.Pattern(hay[-n:], pat.case_sensitive, M.MODE_SUFFIX)
    o = rng.randrange(43 - n + 1)
    return M.Pattern(hay[o:o + n], pat.case_sensitive, M.MODE_CONTAINS)


def test_config4_batched_equals_single_and_oracle(gpu, oracle):
    """64 concurrent searches of mixed kinds in one launch (config 4), 2^22 candidates
    each, then checked three ways:
    - per-search hit sets equal one search per launch over the full range;
    - every hit of every search is re-derived (seed and public key) and re-matched by the
      oracle;
    - around up to two hits of each search, a 2^14-candidate window has exactly the
      oracle's hit set (nothing missed there).
    Long patterns have no hit in 2^22 candidates (a 7-char case-sensitive substring needs
    about 10^11), so each search without one gets a planted twin in the same launch: the
    same key, kind, case rule and length, with the text of a chosen key as its pattern;
    its window around the planted key is checked the same way. Every one of the 64
    search configurations is thereby checked on at least one oracle-confirmed hit
    (src/service/gpu.rs:2155-2229, match agreement)."""
    n = 1 << 22
    specs, pats, keys = config4_specs(n)
    say("config4 batched")
    batched, tries = gpu.batch_run(specs)
    assert all(t == n for t in tries)
    by_search = {i: [] for i in range(64)}
    for h in batched:
        by_search[h.search].append(h)
    for i, s in enumerate(specs):
        single, _ = gpu.batch_run([s])
        assert [h.index for h in by_search[i]] == [h.index for h in single], i
    for i, s in enumerate(specs):
        cs = not s["fold"]
        for h in by_search[i]:
            assert h.seed == oracle.seed(keys[i], 0, h.index), (i, h.index)
            assert h.pk == oracle.derive(keys[i], 0, h.index, 1), (i, h.index)
            assert oracle.match(oracle.public_b64(h.pk), s["needle"], cs, s["mode"]), (i, h.index)
    rng = random.Random(48)
    win = 1 << 14

    def window_check(i, spec, hits, around):
        lo = max(0, min(around - win // 2, n - win))
        want = oracle.search(spec["key"], 0, lo, win, spec["needle"], not spec["fold"], spec["mode"])
        assert around in want, (i, around)
        assert [h.index for h in hits if lo <= h.index < lo + win] == want, (i, lo)

    twins = {}
    for i, s in enumerate(specs):
        hs = by_search[i]
        if hs:
            for h in rng.sample(hs, min(2, len(hs))):
                window_check(i, s, hs, h.index)
        else:
            planted = rng.randrange(n)
            twins[i] = (planted, _planted_twin(pats[i], keys[i], planted, oracle, rng).spec(keys[i], 0, n))
    say(f"config4 hits={len(batched)}; planted twins for {len(twins)} searches")
    if twins:
        order = sorted(twins)
        thits, ttries = gpu.batch_run(specs + [twins[i][1] for i in order])
        assert all(t == n for t in ttries)
        for j, i in enumerate(order):
            planted, spec = twins[i]
            hs = [h for h in thits if h.search == 64 + j]
            for h in hs:
                assert h.pk == oracle.derive(keys[i], 0, h.index, 1), (i, h.index)
            window_check(i, spec, hs, planted)
        # the original 64 searches are unchanged by the extra ones in the launch
        assert [(h.search, h.index) for h in thits if h.search < 64] == [(h.search, h.index) for h in batched]


def _served_checks(oracle, spec, r):
    """A served result is the first exact hit at or after `first`: the oracle finds the
    same lowest hit in the range the request covered, and the key and seed re-derive."""
    assert r.status == 1
    assert r.tries > 0 and r.index >= spec["first"]
    want = oracle.search(spec["key"], spec["seed_source"], spec["first"], r.index - spec["first"] + 1,
                         spec["needle"], not spec["fold"], spec["mode"])
    assert want and want[0] == r.index
    assert oracle.derive(spec["key"], spec["seed_source"], r.index, 1) == r.pk
    assert oracle.seed(spec["key"], spec["seed_source"], r.index) == r.seed


def _serve_specs(n, seed):
    rng = random.Random(seed)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    out = []
    for i in range(n):
        kind = i % 4
        L = rng.randint(2, 3)
        text = "".join(rng.choice(alpha) for _ in range(L))
        if kind == 0:
            p = M.Pattern(text, False)
        elif kind == 1:
            p = M.Pattern(text, True)
        elif kind == 2:
            p = M.Pattern(text[:2], True, M.MODE_SUFFIX)
        else:
            p = M.Pattern(rng.choice(["[0-9]", "[A-F]", "a"]) + rng.choice(alpha[:62]), True, M.MODE_REGEX)
        out.append(p.spec(random.Random(seed * 1000 + i).randbytes(32), rng.getrandbits(48), 0,
                          seed_source=i % 2))
    return out


@pytest.mark.parametrize("use_graph,overlap,unit", [(True, True, 128), (False, True, 128), (True, False, 128),
                                                   (True, True, 64), (False, False, 64)])
def test_server_results_alone_and_in_the_mix(gpu, oracle, use_graph, overlap, unit):
    """Continuous batching: requests submitted while others run; each result equals the
    result of the same request served alone, and the oracle confirms it. With overlap, two
    steps serving disjoint requests are in flight on two streams. Steps claimed in 64-key
    units (half the default) give the same results."""
    specs = _serve_specs(24, 7)
    with M.Server(gpu, max_slots=8, use_graph=use_graph, step_keys=1 << 22, overlap=overlap, unit_keys=unit) as srv:
        ids = {}
        results = {}
        i = 0
        steps = 0
        while len(results) < len(specs):
            # two arrivals per step; at most 8 slots, so the rest queue
            for _ in range(2):
                if i < len(specs):
                    ids[srv.submit(specs[i])] = i
                    i += 1
            srv.step()
            steps += 1
            for r in srv.poll():
                results[ids[r.request]] = r
            assert steps < 2000
    for k, s in enumerate(specs):
        _served_checks(oracle, s, results[k])
    # alone
    for k in (0, 5, 11, 23):
        with M.Server(gpu, max_slots=4, use_graph=use_graph, step_keys=1 << 21, overlap=overlap,
                      unit_keys=unit) as srv:
            rid = srv.submit(specs[k])
            got = []
            while not got:
                srv.step()
                got = srv.poll()
            assert got[0].request == rid and got[0].index == results[k].index and got[0].pk == results[k].pk


def test_server_graph_replays_match_eager(gpu):
    """The captured step graph, replayed many times with host work in between, gives the
    same per-request results as enqueuing each step."""
    specs = _serve_specs(16, 9)
    out = {}
    for use_graph in (True, False):
        with M.Server(gpu, max_slots=16, use_graph=use_graph, step_keys=1 << 21) as srv:
            ids = {srv.submit(s): n for n, s in enumerate(specs)}
            res = {}
            while len(res) < len(specs):
                srv.step()
                for r in srv.poll():
                    res[ids[r.request]] = (r.status, r.index, r.pk)
            out[use_graph] = res
    assert out[True] == out[False]


def test_server_units_at_192_keys_per_lane(oracle):
    """At 192 keys per lane a serving step is claimed in three 64-key units per tile
    (unit_keys = 64), or in whole tiles where the unit does not divide the tile (128) or
    is the tile (192). Every request's result is its first exact hit, the same in all
    three, and the oracle's."""
    specs = _serve_specs(12, 31)
    say("open gpu W=16 batch=192 blocks=1")
    g = M.Gpu(0, width=16, batch=192, blocks=1)
    try:
        out = {}
        for unit in (64, 128, 192):
            say(f"serving in {unit}-key units")
            with M.Server(g, max_slots=6, step_keys=1 << 21, unit_keys=unit) as srv:
                ids = {srv.submit(s): k for k, s in enumerate(specs)}
                res = {}
                steps = 0
                while len(res) < len(specs):
                    srv.step()
                    steps += 1
                    for r in srv.poll():
                        res[ids[r.request]] = r
                    assert steps < 4000
            out[unit] = {k: (r.status, r.index, r.pk) for k, r in res.items()}
            if unit == 64:
                for k, s in enumerate(specs):
                    _served_checks(oracle, s, res[k])
        assert out[64] == out[128] == out[192]
    finally:
        g.close()


def test_server_unit_keys_checked(gpu):
    """unit_keys is 64, 128, 192 or 256 (0 means the default 128); other values are refused
    before anything is allocated."""
    for bad in (32, 96, 320):
        with pytest.raises(M.MeikeyError) as e:
            M.Server(gpu, max_slots=2, unit_keys=bad)
        assert e.value.status == 3  # MK_ERR_OUT_OF_RANGE
    for ok in (0, 192, 256):
        with M.Server(gpu, max_slots=2, step_keys=1 << 20, unit_keys=ok) as srv:
            assert srv.stats()[1] >= 1 << 20


@pytest.mark.parametrize("overlap,use_graph,per_step", [(True, True, 16), (True, False, 16), (False, True, 16),
                                                        (True, True, 2)])
def test_server_churn(gpu, oracle, overlap, use_graph, per_step):
    """Steps under churn (two in flight with overlap; replayed graphs or eager steps; 16 or
    2 requests per step, so most residents wait while the least-served run): 6 slots and a DFA pool that holds three of the
    larger DFAs, so slots and DFA memory are reused while the other step still runs;
    arrivals every step; candidate budgets that run out; unbounded requests and random
    others cancelled while queued or running. Every request gets exactly one result: a
    found one is the oracle's first exact hit at or after its first index, a spent one
    has searched at least its budget, a cancelled one was cancelled."""
    rng = random.Random(2024)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    big = M.Pattern("[A-Z]{2}[0-9]zz$", True, M.MODE_REGEX)
    assert big.dfa_states > 1
    specs = []
    for i in range(120):
        kind = i % 5
        key, src, first = rng.randbytes(32), rng.choice([0, 1]), rng.getrandbits(40)
        count = 0
        if kind == 0:
            p = M.Pattern("".join(rng.choice(alpha) for _ in range(2)), False)
        elif kind == 1:
            p = M.Pattern("".join(rng.choice(alpha) for _ in range(2)), True, M.MODE_SUFFIX)
        elif kind == 2:
            p = big if rng.random() < 0.3 else M.Pattern(rng.choice(["[0-9][A-F]", "^[A-P]", "Q.$", "a[a-f]"]),
                                                            True, M.MODE_REGEX)
        elif kind == 3:  # about 2^36 tries to a hit: the budget runs out first
            p, count = M.Pattern("".join(rng.choice(alpha) for _ in range(6)), True, M.MODE_SUFFIX), 1 << 21
        else:  # unbounded and as rare: cancelled
            p = M.Pattern("".join(rng.choice(alpha) for _ in range(6)), True, M.MODE_SUFFIX)
        specs.append((p.spec(key, first, count, seed_source=src), kind))
    with M.Server(gpu, max_slots=6, dfa_budget=3 * len(big.needle()[0]), step_keys=1 << 20, overlap=overlap,
                  use_graph=use_graph, active_per_step=per_step) as srv:
        ids, results, cancelled = {}, {}, set()
        waiting = list(range(len(specs)))
        steps = 0
        while len(results) < len(specs):
            for _ in range(rng.randint(0, 3)):
                if waiting:
                    k = waiting.pop(0)
                    ids[srv.submit(specs[k][0])] = k
            for rid, k in list(ids.items()):
                if k in results or rid in cancelled:
                    continue
                if (specs[k][1] == 4 and rng.random() < 0.2) or rng.random() < 0.002:
                    try:
                        srv.cancel(rid)
                        cancelled.add(rid)
                    except M.MeikeyError:
                        pass  # it finished in the last step; its result is waiting to be polled
            srv.step()
            steps += 1
            for r in srv.poll():
                k = ids[r.request]
                assert k not in results
                results[k] = r
            assert steps < 20000
    for rid, k in ids.items():
        r, (spec, kind) = results[k], specs[k]
        if r.status == 1:
            assert rid not in cancelled
            _served_checks(oracle, spec, r)
        elif r.status == 0:
            assert spec["count"] and r.tries >= spec["count"], (k, r.tries)
        else:
            assert r.status == -1 and rid in cancelled, (k, r.status)
    statuses = [r.status for r in results.values()]
    assert statuses.count(1) >= 60 and statuses.count(0) >= 20 and statuses.count(-1) >= 20, statuses


def test_six_hundred_searches_in_one_launch(gpu, oracle):
    """600 searches in one launch with ragged counts (0, 1, a tile less one, a tile, a
    tile and one, a few tiles), both seed sources and every match kind: each search covers
    its count rounded up to whole tiles, and its full hit set is the oracle's over that
    range."""
    rng = random.Random(600)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    tile = gpu.tile_keys
    counts = [0, 1, tile - 1, tile, tile + 1, 3 * tile + 17]
    specs = []
    for i in range(600):
        kind = i % 4
        text = "".join(rng.choice(alpha) for _ in range(2))
        if kind == 0:
            p = M.Pattern(text, False)
        elif kind == 1:
            p = M.Pattern(text, True)
        elif kind == 2:
            p = M.Pattern(text, rng.random() < 0.5, M.MODE_SUFFIX)
        else:
            p = M.Pattern(rng.choice(["[0-9][a-f]", "^[A-D]Q", "Zq?.$", "a[a-f]b"]), True, M.MODE_REGEX)
        specs.append(p.spec(rng.randbytes(32), rng.getrandbits(44), counts[i % len(counts)], seed_source=i % 2))
    hits, tries = gpu.batch_run(specs)
    assert tries == [-(-s["count"] // tile) * tile for s in specs]
    by = {}
    for h in hits:
        by.setdefault(h.search, []).append(h.index)
    assert sum(len(v) for v in by.values()) > 10000
    for i, s in enumerate(specs):
        want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"], s["mode"],
                             threads=16) if tries[i] else []
        assert by.get(i, []) == want, i


def test_batched_stop_on_hit_dense_and_sparse(gpu, oracle):
    """Stop-on-hit searches in one launch, three whose patterns match most keys (every
    key, most keys, one key in 64) beside three sparse ones, in both seed sources: each
    search's lowest returned hit is its first exact hit (the oracle's), and every returned
    hit re-derives and matches."""
    rng = random.Random(515)
    pats = [M.Pattern("^[A-P]", True, M.MODE_REGEX), M.Pattern("[0-9][A-F]", True, M.MODE_REGEX),
            M.Pattern("Q.$", True, M.MODE_REGEX), M.Pattern("k9w", False), M.Pattern("Zq", True, M.MODE_SUFFIX),
            M.Pattern("[a-f]{2}[0-9]zz", True, M.MODE_REGEX)]
    specs = [p.spec(rng.randbytes(32), rng.getrandbits(40), 1 << 22, seed_source=i % 2, stop_on_hit=True)
             for i, p in enumerate(pats)]
    hits, tries = gpu.batch_run(specs)
    for i, s in enumerate(specs):
        mine = [h for h in hits if h.search == i]
        assert mine, i
        low = min(h.index for h in mine)
        want = oracle.search(s["key"], s["seed_source"], s["first"], low - s["first"] + 1, s["needle"], not s["fold"],
                             s["mode"])
        assert want and want[0] == low, (i, pats[i].text)
        for h in mine[:8]:
            assert h.pk == oracle.derive(s["key"], s["seed_source"], h.index, 1)
            assert pats[i].is_match(T.public_b64(h.pk)), (i, h.index)


def test_server_budget_and_cancel(gpu):
    """Admission waits for DFA-pool memory; cancelled requests end with status -1; a
    request with a candidate budget ends with status 0 when it is spent."""
    big = M.Pattern("[A-Z]{3}[0-9]zz$", True, M.MODE_REGEX)
    assert big.dfa_states > 1
    table_bytes = len(big.needle()[0])
    with M.Server(gpu, max_slots=8, use_graph=True, dfa_budget=table_bytes + 16, step_keys=1 << 20) as srv:
        with pytest.raises(M.MeikeyError):
            srv.submit(big.spec(b"\x03" * 32, 0, 1 << 21, seed_source=7))
        a = srv.submit(big.spec(b"\x01" * 32, 0, 1 << 21))
        b = srv.submit(big.spec(b"\x02" * 32, 0, 1 << 21))
        active, queued = srv.step()
        assert (active, queued) == (1, 1)  # only one DFA fits the budget
        srv.cancel(b)
        done = {}
        while a not in done:
            srv.step()
            for r in srv.poll():
                done[r.request] = r
        assert done[b].status == -1
        assert done[a].status in (0, 1)
        if done[a].status == 0:
            assert done[a].tries >= 1 << 21


def test_searcher_exact_hit_and_openssh(gpu, oracle, tmp_path):
    """Searcher::run (search.rs:34-40, gpu.rs:999-1012): the reported key is a real hit of
    the pattern, derived from the master key at the reported index; its OpenSSH files
    carry the same key."""
    from tests.test_openssh import parse_private
    key = random.Random(60).randbytes(32)
    for text, cs, mode in [("Zq", True, M.MODE_SUFFIX), ("k9w", False, M.MODE_CONTAINS), ("^[A-D][0-9]", True, M.MODE_REGEX)]:
        say(f"searcher {text!r}")
        pat = M.Pattern(text, cs, mode)
        seen = []
        out = gpu.search(pat, M.SEED_CHACHA20, key=key, progress=lambda t, e: seen.append(t), progress_interval=1e-3)
        assert out.found == 1 and bytes(out.key) == key
        seed = oracle.seed(key, M.SEED_CHACHA20, out.index)
        assert bytes(out.seed) == seed
        pk = oracle.derive(key, M.SEED_CHACHA20, out.index, 1)
        assert bytes(out.pk) == pk
        assert pat.is_match(T.public_b64(pk))
        assert out.stats.tries >= out.index + 1 and out.launches >= 1
        assert seen == sorted(seen)
        import base64
        line = M.openssh_public(pk, "vanity")
        assert line.split() == ["ssh-ed25519", base64.b64encode(b"\0\0\0\x0bssh-ed25519\0\0\0\x20" + pk).decode(),
                                "vanity"]
        _, _, _, _, pub, priv = parse_private(M.openssh_private(seed, pk, "vanity", 7))
        assert pub.endswith(pk) and seed + pk in priv


def test_c_client_of_the_abi(oracle):
    """meikey_amd/mk_example (csrc/mk_example.c, plain C against include/meikey_hip.h):
    its Searcher run returns the oracle's first hit and the matching OpenSSH line."""
    import subprocess
    exe = os.path.join(os.path.dirname(HERE), "meikey_amd", "mk_example")
    if not os.path.exists(exe):
        pytest.skip("C client not built")
    key = random.Random(99).randbytes(32)
    say("run mk_example")
    res = subprocess.run([exe, "Zq", key.hex()], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.splitlines()
    index = int(lines[0].split()[1])
    pk = bytes.fromhex(lines[1].split()[1])
    assert pk == oracle.derive(key, 0, index, 1)
    first = oracle.search(key, 0, 0, index + 1, b"zq", False, M.MODE_SUFFIX)
    assert first and first[0] == index
    assert lines[3] == M.openssh_public(pk, "example")


def test_searcher_bench_runs_to_the_time_limit(gpu_small):
    pat = M.Pattern("A", True, M.MODE_CONTAINS)
    out = gpu_small.search(pat, M.SEED_COUNTER, bench=True, time_limit=0.5, target_launch=0.05)
    assert out.found == 0
    assert out.launches >= 3 and out.stats.elapsed_s >= 0.5
    assert out.stats.tries % gpu_small.keys_per_iter == 0


def test_searcher_cancel(gpu_small):
    import ctypes
    flag = ctypes.c_int32(1)
    out = gpu_small.search(M.Pattern("AAAAAAAAAA", True, M.MODE_CONTAINS), M.SEED_COUNTER, cancel=flag)
    assert out.found == 0 and out.launches == 1


def test_autotuner_measures_and_picks_the_fastest():
    """fastest (gpu.rs:1014-1060): every setting opened, self-checked and measured; an
    invalid one is rejected with its status; the best is the measured maximum."""
    say("autotune W=12/14")
    best, rows = M.fastest(0, [(12, 8, 1), (14, 16, 2), (40, 8, 1)])
    assert rows[2][4] == 3 and rows[2][3] == 0  # out of range
    ok = [r for r in rows if r[4] == 0]
    assert len(ok) == 2 and all(r[3] > 1e6 for r in ok)
    top = max(ok, key=lambda r: r[3])
    assert (best.width, best.batch, best.blocks) == top[:3]
