"""Patterns and the regex -> DFA compiler (host side, no GPU).

The DFA is pinned against Python's `re` (ASCII mode) on the 43-character hay, which
has the same semantics as the reference's regex configuration for the syntax used
here; against the fixed cases of the reference's own tests (src/model/pattern.rs,
src/model/pattern/dfa.rs); and through the closed-form probabilities those tests use.
"""
import math
import random
import zlib
import re
import struct

import pytest

import meikey_amd as M
from tests import textbook as T

ALPHA = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
SAMPLE = "AAAAC3NzaC1lZDI1NTE5AAAAIIQK2/0eXLDku+cWkIDsqHQOwFFTEDDoWaKofXuV/Gy9"
PREFIX_ONLY = "AAAAC3NzaC1lZDI1NTE5AAAAI0000000000000000000000000000000000000000000"


def dfa_walk(table: bytes, values, start=1):
    s = start
    for v in values:
        s = struct.unpack_from("<H", table, (s * 65 + v) * 2)[0]
        if s == 0:
            return True
    return struct.unpack_from("<H", table, (s * 65 + 64) * 2)[0] != 0


def hay_values(b64: str):
    return [ALPHA.index(c) for c in b64[25:]]


@pytest.mark.parametrize("text,error", [
    ("wakokara", None), ("WaKoKaRa", None), ("gack951", None), ("a+b/c", None), ("a", None),
    ("a" * 41, None), ("a" * 42, None), ("", "Empty"), ("wako-kara", "InvalidChar"), ("wako kara", "InvalidChar"),
    ("wako=", "InvalidChar"), ("x" * 43, "TooLong"),
])
@pytest.mark.parametrize("cs", [False, True])
def test_new_validates(text, error, cs):
    # src/model/pattern.rs:311-342
    if error is None:
        p = M.Pattern(text, cs)
        assert p.text == text and p.case_sensitive == cs
    else:
        with pytest.raises(M.PatternError) as e:
            M.Pattern(text, cs)
        assert e.value.kind == error


@pytest.mark.parametrize("b64,text,cs,mode,want", [
    # pattern.rs:352-399
    (SAMPLE, "wako", False, 0, True), (SAMPLE, "WAKO", False, 0, True), (SAMPLE, "WaKo", True, 0, True),
    (SAMPLE, "wako", True, 0, False), (SAMPLE, "Gy9", True, 0, True), (SAMPLE, "IQK2", True, 0, True),
    (SAMPLE, "zzzz", False, 0, False), (PREFIX_ONLY, "AAAAC3", True, 0, False),
    (PREFIX_ONLY, "NzaC1lZDI1", True, 0, False), (PREFIX_ONLY, "NTE5", True, 0, False),
    (PREFIX_ONLY, "0000", True, 0, True),
    (SAMPLE, "Gy9", True, 1, True), (SAMPLE, "gY9", False, 1, True), (SAMPLE, "WaKo", True, 1, False),
    (SAMPLE, "^IQK2", True, 2, True), (SAMPLE, "^AQK2", True, 2, False), (SAMPLE, "Wa.o", True, 2, True),
    (SAMPLE, "wa.o", False, 2, True), (SAMPLE, "[0-9]$", True, 2, True), (SAMPLE, "[0-9]", True, 2, True),
    (SAMPLE, "NzaC1l", True, 2, False),
    # dfa.rs:225-239
    (SAMPLE, "WaKo", True, 2, True), (SAMPLE, "wako", True, 2, False), (SAMPLE, "wako", False, 2, True),
    (SAMPLE, "^QK2", True, 2, False), (SAMPLE, "Gy9$", True, 2, True), (SAMPLE, "Gy$", True, 2, False),
    (SAMPLE, "[0-9]{2}[a-z]", True, 2, False), (SAMPLE, "zzzz|ofX", True, 2, True), (SAMPLE, "NTE5", True, 2, False),
])
def test_fixed_matches(b64, text, cs, mode, want):
    try:
        p = M.Pattern(text, cs, mode)
    except M.PatternError as e:
        # dfa.rs builds "^QK2" with Dfa::new alone; as a Pattern it never matches any key
        assert e.kind == "NeverMatches" and not want
        return
    assert p.is_match(b64) == want
    if mode == M.MODE_REGEX:
        needle, n, fold, m = p.needle()
        assert (n, fold, m) == (p.dfa_states, False, M.MODE_REGEX)
        assert dfa_walk(needle, hay_values(b64)) == want


@pytest.mark.parametrize("text,kind", [
    # pattern.rs:400-446
    ("a(", "Regex"), ("a*|b", "MatchesEmpty"), ("^", "MatchesEmpty"), ("a=b", "NeverMatches"),
    ("^Q", "NeverMatches"), ("[a-z]{44}", "NeverMatches"), ("a[A-Za-z0-9+/]{20}", "TooComplex"), ("", "Empty"),
    ("(", "Regex"), ("a)", "Regex"), ("[a", "Regex"), ("a{2,1}", "Regex"), ("*a", "Regex"), ("\\q", "Regex"),
])
def test_regex_errors(text, kind):
    with pytest.raises(M.PatternError) as e:
        M.Pattern(text, True, M.MODE_REGEX)
    assert e.value.kind == kind


def test_gpu_table_layout():
    # dfa.rs:297-306
    p = M.Pattern("ab", True, M.MODE_REGEX)
    t, n, _, _ = p.needle()
    assert len(t) == n * 65 * 2
    word = lambda s, i: struct.unpack_from("<H", t, (s * 65 + i) * 2)[0]
    assert all(word(0, i) == 0 for i in range(64))
    assert word(0, 64) == 1 and word(1, 64) == 0
    assert all(word(s, i) < n for s in range(n) for i in range(64))


@pytest.mark.parametrize("text,want", [
    # dfa.rs:283-295
    ("z", 1.0 - (63 / 64) ** 42), ("z$", 1 / 64), ("^A", 1 / 16), ("ab$", 1 / 4096), ("(?i)ab$", 4 / 4096),
])
def test_regex_probability(text, want):
    assert abs(M.Pattern(text, True, M.MODE_REGEX).probability - want) < 1e-12


@pytest.mark.parametrize("mode,text,cs,want", [
    # pattern.rs:458-477
    (M.MODE_SUFFIX, "ab", True, 4096.0), (M.MODE_SUFFIX, "ab", False, 1024.0), (M.MODE_SUFFIX, "12", False, 4096.0),
    (M.MODE_REGEX, "ab$", True, 4096.0), (M.MODE_REGEX, "[ab]$", True, 32.0),
])
def test_expected_tries_with_mode(mode, text, cs, want):
    assert abs(M.Pattern(text, cs, mode).expected_tries / want - 1) < 1e-9


@pytest.mark.parametrize("text,insensitive,sensitive", [
    # pattern.rs:503-530
    ("abcdefg", 1.0 / (36.0 * (2 / 64) ** 7 + (2 / 64) ** 6 / 16), 64 ** 7 / 36),
    ("gack951", 1.0 / (36.0 * 16 / 64 ** 7 + 8 / 64 ** 6 / 16), 64 ** 7 / 36),
    ("wakokara", 64 ** 8 / (2 ** 8 * 35), 64 ** 8 / 35),
    ("W", 64 / (2 * 42), 64 / 42),
    ("A", 1 / (2 * 42 / 64 + 1 / 16), 1 / (42 / 64 + 1 / 16)),
    ("pZ", 1 / (41 * 4 / 4096 + 2 / 64 / 16), 4096 / 41),
    ("Pz", 1 / (41 * 4 / 4096 + 2 / 64 / 16), 1 / (41 / 4096 + 1 / 64 / 16)),
    ("Qz", 4096 / (41 * 4), 4096 / 41),
    ("2026", 64 ** 4 / 39, 64 ** 4 / 39),
    ("+/+/", 64 ** 4 / 39, 64 ** 4 / 39),
])
def test_expected_tries(text, insensitive, sensitive):
    assert abs(M.Pattern(text, False).expected_tries / insensitive - 1) < 1e-12
    assert abs(M.Pattern(text, True).expected_tries / sensitive - 1) < 1e-12


def test_regex_probability_close_to_contains():
    rng = random.Random(5)
    for _ in range(64):
        n = rng.randint(4, 6)
        text = "".join(rng.choice(ALPHA[:62]) for _ in range(n))
        cs = rng.random() < 0.5
        c = M.Pattern(text, cs).probability
        r = M.Pattern(text, cs, M.MODE_REGEX).probability
        assert abs(r / c - 1) < 1e-3


def random_hay(rng):
    return rng.choice("ABCDEFGHIJKLMNOP") + "".join(rng.choice(ALPHA) for _ in range(42))


def rand_regex(rng, depth=0):
    """A random pattern in the syntax shared by Python re and the reference's regex."""
    def atom():
        r = rng.random()
        if r < 0.45:
            c = rng.choice(ALPHA)
            return "\\" + c if c in "+/" else c
        if r < 0.6:
            a, b = sorted(rng.sample(range(62), 2))
            return f"[{ALPHA[a]}-{ALPHA[b]}]" if (a < 26) == (b < 26) and (a < 52) == (b < 52) else f"[{ALPHA[a]}{ALPHA[b]}]"
        if r < 0.68:
            return "[^" + "".join(rng.sample(ALPHA[:62], 3)) + "]"
        if r < 0.74:
            return "."
        if r < 0.8:
            return rng.choice(["\\d", "\\w", "\\W", "\\D"])
        if r < 0.85 and depth < 2:
            return "(" + rand_regex(rng, depth + 1) + ")"
        if r < 0.9:
            return rng.choice(["\\b", "\\B"])
        return rng.choice(ALPHA[:10])

    parts = []
    for _ in range(rng.randint(1, 4)):
        a = atom()
        q = rng.random()
        if a in ("\\b", "\\B"):
            pass  # Python re rejects quantified look-arounds
        elif q < 0.12:
            a += "?"
        elif q < 0.2:
            a += "+"
        elif q < 0.26:
            a += "*"
        elif q < 0.32:
            n = rng.randint(1, 3)
            a += f"{{{n},{n + rng.randint(0, 2)}}}"
        parts.append(a)
    body = "".join(parts)
    if rng.random() < 0.2:
        body = "^" + body
    if rng.random() < 0.2:
        body = body + "$"
    if depth == 0 and rng.random() < 0.2:
        body = body + "|" + rand_regex(rng, depth + 1)
    return body


def test_dfa_equals_python_re():
    """For random patterns and random hays (first char A-P, as in real keys), the DFA
    and re.search agree, both case-sensitive and not; so does the either-sign walk."""
    rng = random.Random(99)
    tested = hits = 0
    for _ in range(400):
        text = rand_regex(rng)
        cs = rng.random() < 0.5
        try:
            p = M.Pattern(text, cs, M.MODE_REGEX)
        except M.PatternError as e:
            if e.kind in ("MatchesEmpty", "NeverMatches", "TooComplex"):
                continue
            raise
        rx = re.compile(text, re.ASCII | (0 if cs else re.IGNORECASE))
        table, n, _, _ = p.needle()
        # hays over a few characters the patterns use, so that many of them match
        small = sorted(set(c for c in text if c in ALPHA) | set("aA0+B"))
        for i in range(60):
            if i % 2:
                hay = random_hay(rng)
            else:
                hay = rng.choice("ABCDEFGHIJKLMNOP") + "".join(rng.choice(small) for _ in range(42))
            want = rx.search(hay) is not None
            hits += want
            vals = [ALPHA.index(c) for c in hay]
            assert dfa_walk(table, vals) == want, (text, cs, hay)
        tested += 1
    assert tested > 250 and hits > 2000


@pytest.mark.parametrize("text,same", [
    ("(a?){100}b", "a{0,100}b"), ("x(y?){45,}z", "xy*z"), ("(ab|c){2,60}", "(?:ab|c){2,60}"),
    ("[A-P]{0,1000}x", "[A-P]*x"), ("(a|b|c){3,500}$", "[abc]{3,}$"), ("^(.){43,50}", "^.{43}"),
    ("(.){43}", ".{43}"), ("((a?b?){50}){50}c", "[ab]*c")])
def test_large_repetition_counts_match_python_re(text, same):
    """Counted repetitions above 44 are built as 44 copies (the hay has 43 characters).
    The DFA agrees with re.search of an equivalent pattern that Python's backtracking
    matcher runs in linear time, on every hay; and it compiles in well under a second."""
    import time
    rng = random.Random(zlib.crc32(text.encode()))
    t0 = time.perf_counter()
    p = M.Pattern(text, True, M.MODE_REGEX)
    assert time.perf_counter() - t0 < 2.0
    rx = re.compile(same, re.ASCII)
    table, n, _, _ = p.needle()
    small = sorted(set(c for c in text if c in ALPHA) | set("aA0+B"))
    hits = 0
    for i in range(200):
        if i % 2:
            hay = random_hay(rng)
        else:
            hay = rng.choice("ABCDEFGHIJKLMNOP") + "".join(rng.choice(small) for _ in range(42))
        want = rx.search(hay) is not None
        hits += want
        assert dfa_walk(table, [ALPHA.index(c) for c in hay]) == want, (text, hay)
    assert hits > 0


@pytest.mark.parametrize("text,kind", [("(.){44}", "NeverMatches"), ("(.){45,1000}", "NeverMatches"),
                                       ("(\\b.){60}", "NeverMatches"),
                                       ("(a{1000}){1000}", "NeverMatches"), ("(a|b|c){100000}", "NeverMatches"),
                                       ("((((a{44}){44}){44}){44})b", "TooComplex"),
                                       # nested counts whose product overflows 64 bits: not "matches empty"
                                       ("((((a{65536}){65536}){65536}){65536})", "TooComplex"),
                                       ("((((a{65536}){65536}){65536}){65536})|b", "TooComplex"),
                                       # small automata whose subset construction explodes: the build limit
                                       ("((a|.{0,43}){0,43}){0,6}b.{12}", "TooComplex"),
                                       ("(a|.{0,43}){0,43}(b|.{0,43}){0,43}c.{12}", "TooComplex"),
                                       ("(a|.{0,43}){0,43}b.{12}", "TooComplex")])
def test_huge_repetitions_fail_fast(text, kind):
    """Patterns that need more than the 43 hay characters never match, whatever their
    counts; nested counts that would build a huge automaton, and subset constructions
    past the 16 MiB build limit (regex-automata's determinize_size_limit,
    src/model/pattern/dfa.rs:23, 64-65), are too complex. Each is rejected within
    seconds, not after building millions of states."""
    import time
    t0 = time.perf_counter()
    with pytest.raises(M.PatternError) as e:
        M.Pattern(text, True, M.MODE_REGEX)
    assert e.value.kind == kind, e.value
    assert time.perf_counter() - t0 < 3.0


@pytest.mark.parametrize("text", ["ab+c", "[0-9][A-F]x", "^[A-P]a|z$", "(ab){2,}", "\\b9", "a.b"])
@pytest.mark.parametrize("cs", [False, True])
def test_matches_regex_crate_cases(text, cs):
    # dfa.rs:241-281, against Python re with the same configuration
    rng = random.Random(zlib.crc32(text.encode()) * 2 + int(cs))
    p = M.Pattern(text, cs, M.MODE_REGEX)
    rx = re.compile(text, re.ASCII | (0 if cs else re.IGNORECASE))
    table = p.needle()[0]
    for _ in range(1024):
        hay = random_hay(rng)
        assert dfa_walk(table, [ALPHA.index(c) for c in hay]) == (rx.search(hay) is not None), hay


def test_needle_for_text_patterns():
    n, ln, fold, mode = M.Pattern("WaKo", False).needle()
    assert (n, ln, fold, mode) == (b"wako", 4, True, M.MODE_CONTAINS)
    n, ln, fold, mode = M.Pattern("WaKo", True, M.MODE_SUFFIX).needle()
    assert (n, ln, fold, mode) == (b"WaKo", 4, False, M.MODE_SUFFIX)


def test_probability_matches_real_keys(oracle):
    """pattern.rs:531-560 style: the mean count of a 1-char pattern over real keys is
    close to Pattern::probability (A appears more often than Z: hay[0] is A-P)."""
    key = random.Random(77).randbytes(32)
    pks = oracle.derive(key, 0, 0, 4000)
    hays = [T.hay(pks[32 * i:32 * i + 32]) for i in range(4000)]
    for ch in "AZ":
        mean = sum(h.count(ch) for h in hays) / 4000
        assert abs(mean - M.Pattern(ch, True).probability) < 0.05
