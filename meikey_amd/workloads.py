"""Seeded synthetic workloads of the benchmark configurations.

Master keys are the bytes of random.Random(seed).randbytes(32); pattern characters are
drawn from the same generator over the 64-character base64 alphabet.

* single(): the headline search -- ChaCha20 candidates of the key of seed 2, a 6-char
  case-sensitive suffix.
* batched(count): 64 searches with keys of seeds 1000..1063, 25% each of
  case-insensitive substring (3-7 chars), case-sensitive substring (3-7), suffix (3-6,
  either case rule) and a regular expression whose DFA has at most 64 states and that is
  rare enough to be a search (at least 2^10 expected tries), anchored or not.
* serving(rate, seconds): Poisson arrivals at `rate` per second for `seconds`, pattern
  length uniform 3-6, the same kind mix, each request stopping at its first hit.
"""
from __future__ import annotations

import random
from dataclasses import dataclass
from typing import List

from . import MODE_CONTAINS, MODE_REGEX, MODE_SUFFIX, SEED_CHACHA20, Pattern, PatternError

ALPHABET = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"


def single():
    rng = random.Random(2)
    key = rng.randbytes(32)
    needle = "".join(rng.choice(ALPHABET) for _ in range(6))
    return key, needle


def _regex(rng: random.Random, atoms: int) -> Pattern:
    while True:
        body = "".join(rng.choice(["[A-Z]", "[0-9]", "[a-f]", rng.choice(ALPHABET[:62]), rng.choice(ALPHABET[:62]),
                                   "."]) for _ in range(atoms))
        text = rng.choice(["^", "", ""]) + body + rng.choice(["$", "", ""])
        try:
            p = Pattern(text, True, MODE_REGEX)
        except PatternError:
            continue
        if p.dfa_states <= 64 and p.probability <= 2.0 ** -10:
            return p


def pattern_of_kind(rng: random.Random, kind: int, length: int) -> Pattern:
    text = "".join(rng.choice(ALPHABET) for _ in range(length))
    if kind == 0:
        return Pattern(text, False)
    if kind == 1:
        return Pattern(text, True)
    if kind == 2:
        return Pattern(text, rng.random() < 0.5, MODE_SUFFIX)
    return _regex(rng, max(2, length - 1))


def batched(count: int, seed: int = 4):
    """(specs, patterns, keys) of the 64-search batch, each over indices 0 .. count-1."""
    rng = random.Random(seed)
    pats = []
    for i in range(64):
        kind = i % 4
        length = rng.randint(3, 6) if kind == 2 else rng.randint(3, 7)
        pats.append(pattern_of_kind(rng, kind, length))
    keys = [random.Random(1000 + i).randbytes(32) for i in range(64)]
    return [p.spec(k, 0, count) for p, k in zip(pats, keys)], pats, keys


@dataclass
class Request:
    arrival_s: float
    pattern: Pattern
    key: bytes

    def spec(self) -> dict:
        return self.pattern.spec(self.key, 0, 0, SEED_CHACHA20, stop_on_hit=True)


def serving(rate: float = 50.0, seconds: float = 60.0, seed: int = 5) -> List[Request]:
    rng = random.Random(seed)
    out, t, i = [], 0.0, 0
    while True:
        t += rng.expovariate(rate)
        if t >= seconds:
            return out
        kind = i % 4
        out.append(Request(t, pattern_of_kind(rng, kind, rng.randint(3, 6)),
                           random.Random(10_000 + i).randbytes(32)))
        i += 1
