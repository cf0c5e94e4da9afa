"""Writes tests/golden/golden.json from the pure-Python textbook references
(tests/textbook.py: RFC 8032 key generation, RFC 8439 ChaCha20, Python string search),
independent of the oracle and of the kernels. Run from the repository root:

    python tests/golden/make_golden.py
"""
import json
import os
import random
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from tests import textbook as T  # noqa: E402


def seed_of(key, source, i):
    return T.chacha20_seed(key, i) if source == 0 else T.counter_seed(key, i)


def main():
    out = {"derive": [], "search": []}
    # config-1 style master key: bytes of random.Random(1); counter key of random.Random(3)
    k1 = random.Random(1).randbytes(32)
    k3 = random.Random(3).randbytes(32)
    # a counter base whose tail wraps around 2^64 inside the range
    kw = k3[:24] + (2**64 - 5).to_bytes(8, "little")
    for key, source, idx in [
        (k1, 0, list(range(0, 48)) + [1000, 1001, 2**20 - 1, 2**24 - 1, 2**32, 2**32 + 1, 2**63 + 7, 2**64 - 1]),
        (k3, 1, list(range(0, 16)) + [2**31, 2**40 + 3]),
        (kw, 1, list(range(0, 10))),
    ]:
        pks = [[i, T.public_key(seed_of(key, source, i)).hex()] for i in idx]
        out["derive"].append({"key": key.hex(), "source": source, "pks": pks})
    # hit sets of short patterns over small ranges
    rng = random.Random(2)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    for source, key, first, count, needle, cs, mode in [
        (0, k1, 0, 1024, "".join(rng.choice(alpha) for _ in range(2)), False, 0),
        (0, k1, 5000, 1024, "A", True, 1),
        (1, k3, 77, 1024, "q/", False, 0),
        (0, k1, 2**32 - 512, 1024, "zz", False, 0),
    ]:
        hays = [T.hay(T.public_key(seed_of(key, source, first + i))) for i in range(count)]
        n = needle if cs else needle.lower()
        hits = []
        for i, h in enumerate(hays):
            h2 = h if cs else h.lower()
            if (h2.endswith(n) if mode == 1 else n in h2):
                hits.append(first + i)
        out["search"].append({"key": key.hex(), "source": source, "first": first, "count": count,
                              "needle": needle, "case_sensitive": cs, "mode": mode, "hits": hits})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, sum(len(c["pks"]) for c in out["derive"]), "keys,",
          [len(c["hits"]) for c in out["search"]], "hits")


if __name__ == "__main__":
    main()
