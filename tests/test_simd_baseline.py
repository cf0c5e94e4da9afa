"""The SIMD CPU baseline (oracle/libmeikey_simd.so): the restatement of the reference's
CPU search path that bench.py times. Each engine (AVX-512 IFMA, and AVX2, where the CPU
has them) must give exactly the oracle's public keys and hit sets: the committed golden
fixtures, both seed sources, ragged batches and unaligned starts, every match mode, and
the unsigned-first path's sign handling."""
import json
import os
import random

import pytest

import meikey_amd as M
from tests import textbook as T

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def simd():
    from oracle import simd as S
    try:
        S.lib()
    except RuntimeError as e:
        pytest.skip(str(e))
    return S


def _engines(S):
    out = []
    for isa in (512, 2):
        if S.use(isa):
            out.append(isa)
    return out


@pytest.fixture(scope="module", params=[512, 2], ids=["ifma", "avx2"])
def engine(request, simd):
    if not simd.use(request.param):
        pytest.skip(f"this CPU has no {request.param} engine")
    yield simd
    simd.use(512) or simd.use(2)


def test_golden_fixture(engine):
    """The committed fixture (tests/golden/golden.json, from the textbook references)."""
    g = json.load(open(os.path.join(HERE, "golden", "golden.json")))
    for case in g["derive"]:
        key = bytes.fromhex(case["key"])
        for idx, pk in case["pks"]:
            assert engine.derive(key, case["source"], idx, 1) == bytes.fromhex(pk), (case["source"], idx)
    for case in g["search"]:
        key = bytes.fromhex(case["key"])
        got = engine.search(key, case["source"], case["first"], case["count"], case["needle"].encode(),
                            case["case_sensitive"], case["mode"], threads=3)
        assert got == case["hits"]


def test_textbook_keys(engine):
    """A few keys against the pure-Python RFC 8032 derivation."""
    key = random.Random(300).randbytes(32)
    pks = engine.derive(key, 0, 41, 9)
    for i in range(9):
        assert pks[32 * i:32 * i + 32] == T.public_key(T.chacha20_seed(key, 41 + i))


@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 700])
def test_derive_ragged(engine, oracle, n):
    rng = random.Random(301 + n)
    key = rng.randbytes(32)
    for source in (0, 1):
        first = rng.choice([0, 1, 12345, (1 << 32) - 3, rng.getrandbits(40)])
        assert engine.derive(key, source, first, n) == oracle.derive(key, source, first, n), (source, first)


def test_counter_seed_wrap(engine, oracle):
    """Counter seeds wrap at 2^64 (src/model/search.rs:45-50)."""
    key = bytes(24) + (2 ** 64 - 100).to_bytes(8, "little")
    assert engine.derive(key, 1, 0, 300) == oracle.derive(key, 1, 0, 300)


@pytest.mark.parametrize("needle,cs,mode", [
    (b"ab", True, M.MODE_CONTAINS), (b"Zz", False, M.MODE_CONTAINS), (b"9k", True, M.MODE_SUFFIX),
    (b"k/", False, M.MODE_SUFFIX), (b"+", True, M.MODE_CONTAINS), (b"abc", False, M.MODE_CONTAINS)])
def test_search_equals_oracle(engine, oracle, needle, cs, mode):
    rng = random.Random(hash((needle, cs, mode)) & 0xFFFF)
    key = rng.randbytes(32)
    for source in (0, 1):
        first, n = rng.getrandbits(40), (1 << 15) + 77
        want = oracle.search(key, source, first, n, needle, cs, mode, threads=8)
        assert engine.search(key, source, first, n, needle, cs, mode, threads=5) == want


@pytest.mark.parametrize("text,cs", [("^[A-D][0-9]", True), ("k.$", False), ("[a-f]{2}Q", True)])
def test_search_regex_equals_oracle(engine, oracle, text, cs):
    p = M.Pattern(text, cs, M.MODE_REGEX)
    needle, n_len, _, mode = p.needle()
    key = random.Random(302).randbytes(32)
    first, n = 999, 1 << 14
    want = oracle.search(key, 0, first, n, needle, cs, mode, threads=8)
    got = engine.search(key, 0, first, n, needle, cs, mode, threads=4)
    assert got == want and len(want) > 0


def test_sign_character_hits(engine, oracle):
    """Suffixes whose last-but-one character holds the x sign: the unsigned-first check
    must pass a key whose signed text matches and reject one whose opposite sign would."""
    rng = random.Random(303)
    key = rng.randbytes(32)
    first, n = 5000, 4096
    pks = oracle.derive(key, 0, first, n)
    for i in (3, 100, 2000):
        pk = pks[32 * i:32 * i + 32]
        signed = T.hay(pk)[-2:].encode()
        flipped = T.hay(pk[:31] + bytes([pk[31] ^ 0x80]))[-2:].encode()
        for needle in (signed, flipped):
            want = oracle.search(key, 0, first, n, needle, True, M.MODE_SUFFIX, threads=8)
            assert engine.search(key, 0, first, n, needle, True, M.MODE_SUFFIX, threads=2) == want
        assert first + i in oracle.search(key, 0, first, n, signed, True, M.MODE_SUFFIX, threads=8)


def test_bench_counts(simd):
    """The timed run reports whole batches on every thread it was given."""
    key = random.Random(304).randbytes(32)
    r = simd.bench(key, b"AbCdEf", True, M.MODE_SUFFIX, 0.3, threads=2)
    assert r["threads"] == 2 and r["keys"] > 0 and r["keys"] % 256 == 0 and r["seconds"] >= 0.3
