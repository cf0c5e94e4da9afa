"""The other field / table layouts on the device (mk_options.limbs): ten 25.5-bit limbs
with packed 96-byte entries ("10"), eight 32-bit words ("8") and ten 26-bit limbs with
Karatsuba products, packed ("10k") or as limbs ("10ku"). Each must give exactly the keys,
table and hit sets of the oracle and of the "10u" build: known answers, bulk derivation,
sampled table entries, mixed search batches, a served request, and the headline search at
the production width."""
import os
import random
import struct
import sys

import pytest

import meikey_amd as M
from tests import textbook as T

pytestmark = pytest.mark.gpu
THREADS = min(os.cpu_count() or 1, 16)
LAYOUTS = [M.LIMBS_R25, M.LIMBS_R32, M.LIMBS_K26, M.LIMBS_K26U]
IDS = ["10", "8", "10k", "10ku"]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(key, source, needle, cs, mode, first, count, **kw):
    d = {"key": key, "seed_source": source, "needle": needle, "fold": not cs, "mode": mode, "first": first,
         "count": count}
    d.update(kw)
    return d


@pytest.fixture(scope="module")
def plain():
    say("open gpu W=16 batch=32 limbs=10u")
    g = M.Gpu(0, width=16, batch=32, limbs=M.LIMBS_R25U)
    yield g
    g.close()


@pytest.fixture(scope="module", params=LAYOUTS, ids=IDS)
def gpu_layout(request):
    say(f"open gpu W=16 batch=32 limbs={request.param}")
    g = M.Gpu(0, width=16, batch=32, limbs=request.param)
    assert g.info().options.limbs == request.param
    yield g
    g.close()


def test_known_answers_and_bulk_derivation(gpu_layout, oracle):
    from tests.test_oracle import rfc8032_vectors
    vecs = rfc8032_vectors()
    pks = gpu_layout.derive_seeds(b"".join(v[1] for v in vecs))
    assert [pks[32 * i:32 * i + 32] for i in range(len(vecs))] == [v[2] for v in vecs]
    rng = random.Random(190)
    for source in (0, 1):
        key = rng.randbytes(32)
        first = (1 << 32) - 5000 if source == 0 else rng.getrandbits(40)
        n = 1 << 16
        say(f"derive {n} keys source={source}")
        assert gpu_layout.derive(key, source, first, n) == oracle.derive(key, source, first, n, threads=THREADS)


def _entry_values(words, limbs):
    """The three canonical fields of a raw entry: 24 words of packed 256-bit values, or ("10ku")
    32 words of ten 26-bit limbs per field (limbs 0-7 in words 8f .. 8f+7, 8-9 in 24+2f, 25+2f)."""
    if limbs != M.LIMBS_K26U:
        assert len(words) == 24
        return [int.from_bytes(struct.pack("<8I", *words[8 * f:8 * f + 8]), "little") for f in range(3)]
    assert len(words) == 32 and words[30] == 0 and words[31] == 0
    out = []
    for f in range(3):
        ls = [words[8 * f + i] if i < 8 else words[24 + 2 * f + i - 8] for i in range(10)]
        assert all(x < (1 << 26) for x in ls) and ls[9] < (1 << 21)
        out.append(sum(x << (26 * i) for i, x in enumerate(ls)))
    return out


def test_table_entries(gpu_layout, oracle):
    info = gpu_layout.info()
    rows, cols = info.rows, info.cols
    rng = random.Random(191)
    half = T.inv(2)
    picks = [(0, 0), (0, 1), (rows - 1, 0), (rows - 1, cols - 1)] + [
        (rng.randrange(rows), rng.randrange(cols)) for _ in range(20)]
    for row, k in picks:
        words = gpu_layout.table_words(row, k)
        ypx, ymx, third = _entry_values(words, info.options.limbs)
        assert max(ypx, ymx, third) < T.P
        n = oracle.comb_entry_niels(16, row, k)
        assert ypx == int.from_bytes(n[0:32], "little") * half % T.P, (row, k)
        assert ymx == int.from_bytes(n[32:64], "little") * half % T.P, (row, k)
        xy2d = int.from_bytes(n[64:96], "little")
        assert third == (xy2d * T.inv(2 * T.D) % T.P if row == 0 else xy2d * half % T.P), (row, k)


def test_search_batches_equal_10u_and_oracle(gpu_layout, plain, oracle):
    rng = random.Random(192)
    specs = []
    for i in range(6):
        needle = rng.choice([b"ab", b"Zz", b"9k", b"k/"])
        mode = rng.choice([M.MODE_CONTAINS, M.MODE_SUFFIX])
        specs.append(_spec(rng.randbytes(32), i % 2, needle, i % 3 == 0, mode, rng.getrandbits(40),
                           64 * 32 * rng.randint(1, 200)))
    p = M.Pattern("^[A-D][0-9]", True, M.MODE_REGEX)
    specs.append(p.spec(rng.randbytes(32), 777, 1 << 19))
    say("layout batch")
    hits, tries = gpu_layout.batch_run(specs)
    ref, rtries = plain.batch_run(specs)
    assert tries == rtries
    assert [(h.search, h.index, h.pk, h.seed) for h in hits] == [(h.search, h.index, h.pk, h.seed) for h in ref]
    for i, s in enumerate(specs):
        want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"], s["mode"])
        assert [h.index for h in hits if h.search == i] == want, i


def test_served_request(gpu_layout, oracle):
    key = random.Random(193).randbytes(32)
    spec = M.Pattern("Qz", True, M.MODE_SUFFIX).spec(key, 123, 0)
    with M.Server(gpu_layout, max_slots=4, step_keys=1 << 22) as srv:
        rid = srv.submit(spec)
        got = []
        for _ in range(2000):
            srv.step()
            got = srv.poll()
            if got:
                break
    assert got and got[0].request == rid and got[0].status == 1
    r = got[0]
    first = oracle.search(key, 0, 123, r.index - 123 + 1, b"Qz", True, M.MODE_SUFFIX)
    assert first and first[0] == r.index and r.pk == oracle.derive(key, 0, r.index, 1)


@pytest.mark.parametrize("limbs", LAYOUTS, ids=IDS)
def test_headline_search_at_the_production_width(oracle, limbs):
    """W = 24, 256 keys per lane: a planted 6-char case-sensitive suffix over a 2^22 window
    of the headline key gives the oracle's full hit set."""
    from meikey_amd import workloads as W
    say(f"open gpu W=24 batch=256 limbs={limbs}")
    g = M.Gpu(0, width=24, batch=256, limbs=limbs)
    try:
        key, _ = W.single()
        target = (1 << 31) + 4321
        pk = g.derive(key, 0, target, 1)
        assert pk == oracle.derive(key, 0, target, 1)
        needle = T.hay(pk)[-6:].encode()
        first = target - (1 << 21)
        hits, tries = g.batch_run([_spec(key, 0, needle, True, M.MODE_SUFFIX, first, 1 << 22)])
        want = oracle.search(key, 0, first, tries[0], needle, True, M.MODE_SUFFIX, threads=THREADS)
        assert target in want and [h.index for h in hits] == want
        assert g.derive(key, 1, 12345, 1 << 14) == oracle.derive(key, 1, 12345, 1 << 14, threads=THREADS)
    finally:
        g.close()


def test_staged_autotuner_chooses_the_layout():
    """fastest_staged (gpu.rs:1085-1113): the first stage measures the five integer layouts
    (gpu.rs:364-370 measures every Limbs first), every later stage keeps the layout it
    chose, the width stage includes an LDS-resident table and the inversion stage the
    pooled settings, and the final setting is the fastest of the last stage."""
    say("staged autotuner")
    best, rows = M.fastest_staged(0)
    for r in rows:
        say(f"  W={r[0]} K={r[1]} waves={r[2]} limbs={r[5]} lockstep={r[6]} pool={r[7]} lds={r[8]}: "
            f"{r[3] / 1e9:.3f} G keys/s status {r[4]}")
    stage1 = rows[:5]
    assert sorted(r[5] for r in stage1) == sorted([M.LIMBS_R25U, M.LIMBS_R25, M.LIMBS_R32, M.LIMBS_K26,
                                                   M.LIMBS_K26U])
    assert all(r[4] == 0 and r[3] > 1e8 for r in stage1)
    chosen = max(stage1, key=lambda r: r[3])[5]
    assert best.limbs == chosen and all(r[5] == chosen for r in rows[5:])
    last = rows[-3:]
    top = max((r for r in last if r[4] == 0), key=lambda r: r[3])
    assert (best.width, best.batch, best.blocks, best.lockstep, best.pool, best.lds_table) == \
        (top[0], top[1], top[2], top[6], top[7], top[8])
    # the width stage measures an LDS-resident table, the inversion stage pooled settings
    assert any(r[8] == 1 and r[0] == 5 and r[4] == 0 for r in rows)
    assert any(r[7] == 4 and r[1] == 8 and r[4] == 0 for r in rows)
