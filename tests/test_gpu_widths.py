"""Comb widths at the ends of the range mk_open accepts (2..28; the reference's is 2..20,
src/controller/cli.rs:85-94) and past the production W = 24: W = 2 (126 rows of 3
entries), W = 25 and 26 (23.6 and 43 GB tables, where address translation slows the
gathers), and W = 28 (9 rows of 2^27 + 1 entries, a 154.6 GB table, the largest), in the
default ten-limb layout and, at W = 2, 25 and 28, in the Karatsuba layouts. Each must give
the closed-form table entries, the oracle's public keys in both seed modes and the
oracle's hit sets."""
import os
import random
import sys

import pytest

import meikey_amd as M
from tests import textbook as T
from tests.test_gpu_defaults import _entry_point

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _entry_point_any(words, limbs):
    """(x, y, third field) of a raw entry in any layout."""
    from tests.test_gpu_layouts import _entry_values
    if limbs == M.LIMBS_R25U:
        return _entry_point(words)
    ypx, ymx, third = _entry_values(words, limbs)
    assert max(ypx, ymx, third) < T.P
    return (ypx - ymx) % T.P, (ypx + ymx) % T.P, third


@pytest.mark.parametrize("width,limbs", [(2, M.LIMBS_R25U), (25, M.LIMBS_R25U), (26, M.LIMBS_R25U),
                                         (28, M.LIMBS_R25U), (2, M.LIMBS_K26U), (25, M.LIMBS_K26),
                                         (28, M.LIMBS_K26U)])
def test_width_extremes(oracle, width, limbs):
    rows, cols = -(-251 // width), (1 << (width - 1)) + 1
    say(f"open gpu W={width} batch=32 limbs={limbs}: {rows} rows x {cols} entries")
    g = M.Gpu(0, width=width, batch=32, limbs=limbs)
    try:
        info = g.info()
        assert (info.rows, info.cols) == (rows, cols)
        assert info.table_bytes == rows * cols * 4 * M.ENTRY_WORDS[limbs]
        rng = random.Random(600 + width + limbs)
        # entries: the compressed point k * 2^(w * row + 3) * B (+ 2^254 * B on the top row)
        # from the oracle's own double-and-add, the curve equation and the third field
        picks = [(0, 0), (0, 1), (0, cols - 1), (rows - 1, 0), (rows - 1, 1), (rows - 1, cols - 1)]
        picks += [(rng.randrange(rows), rng.randrange(cols)) for _ in range(10)]
        for row, k in picks:
            x, y, third = _entry_point_any(g.table_words(row, k), limbs)
            assert (y * y - x * x - 1 - T.D * x * x % T.P * y * y) % T.P == 0, (row, k)
            assert int.to_bytes(y | ((x & 1) << 255), 32, "little") == oracle.comb_entry(width, row, k), (row, k)
            assert third == (x * y % T.P if row == 0 else T.D * x % T.P * y % T.P), (row, k)
        for source in (0, 1):
            key = rng.randbytes(32)
            first = rng.getrandbits(40)
            n = 1 << 14
            say(f"derive {n} keys W={width} source={source}")
            assert g.derive(key, source, first, n) == oracle.derive(key, source, first, n, threads=THREADS)
        specs = []
        for source, needle, cs, mode in [(0, b"ab", False, M.MODE_CONTAINS), (1, b"Q9", True, M.MODE_SUFFIX),
                                         (0, b"zK", True, M.MODE_CONTAINS)]:
            specs.append({"key": rng.randbytes(32), "seed_source": source, "needle": needle, "fold": not cs,
                          "mode": mode, "first": rng.getrandbits(40), "count": (1 << 18) + 64 * 32 * 3 + 5})
        say(f"batch W={width}")
        hits, tries = g.batch_run(specs)
        for i, s in enumerate(specs):
            want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"],
                                 s["mode"], threads=THREADS)
            got = [h for h in hits if h.search == i]
            assert [h.index for h in got] == want and want, (width, i)
            for h in got[:4]:
                assert h.pk == oracle.derive(s["key"], s["seed_source"], h.index, 1)
    finally:
        g.close()
