"""The lockstep search kernel (one workgroup per CU whose waves add each comb row
together, mk_options.lockstep = 1) returns exactly what the one-wave-per-workgroup kernel
and the oracle return: batches of mixed searches (whole tiles and tail units, groups of
claims that run past the end of the queue), stop-on-hit searches (waves that pass their
barriers without work) and served requests."""
import random
import sys

import pytest

import meikey_amd as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(key, source, needle, cs, mode, first, count, **kw):
    d = {"key": key, "seed_source": source, "needle": needle, "fold": not cs, "mode": mode, "first": first,
         "count": count}
    d.update(kw)
    return d


@pytest.fixture(scope="module")
def pair():
    say("open gpu W=16 batch=256 lockstep=1 and lockstep=0")
    lock = M.Gpu(0, width=16, batch=256, blocks=2, lockstep=1)
    plain = M.Gpu(0, width=16, batch=256, blocks=2, lockstep=0)
    assert lock.info().options.lockstep == 1
    yield lock, plain
    lock.close()
    plain.close()


def test_lockstep_batches_equal_plain_and_oracle(pair, oracle):
    lock, plain = pair
    tile = 64 * 256
    rng = random.Random(160)
    specs = []
    # the first search spans whole tiles and tail units; the odd tile counts leave the
    # last group of claims partly empty
    for count, needle, cs, mode in [((1 << 25) + 3 * tile, b"Kq7", True, M.MODE_CONTAINS),
                                    (1 << 20, b"zz", False, M.MODE_SUFFIX),
                                    (5 * tile, b"mk", False, M.MODE_CONTAINS),
                                    (7 * tile + 1, b"Q", True, M.MODE_SUFFIX)]:
        specs.append(_spec(rng.randbytes(32), rng.choice([0, 1]), needle, cs, mode, rng.getrandbits(40), count))
    say("lockstep batch of 4 searches")
    hits, tries = lock.batch_run(specs)
    ref, rtries = plain.batch_run(specs)
    assert tries == rtries
    assert [(h.search, h.index, h.pk, h.seed) for h in hits] == [(h.search, h.index, h.pk, h.seed) for h in ref]
    for i in (1, 2, 3):
        s = specs[i]
        want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"], s["mode"])
        assert [h.index for h in hits if h.search == i] == want and want


def test_lockstep_small_launches(oracle):
    """Launches smaller than one claim per wave of the grid, at 32 keys per lane (no tail
    units) and with one wave per SIMD (4-wave workgroups)."""
    for batch, blocks in [(32, 2), (256, 1)]:
        say(f"open gpu W=14 batch={batch} blocks={blocks} lockstep=1")
        g = M.Gpu(0, width=14, batch=batch, blocks=blocks, lockstep=1)
        try:
            key = random.Random(161 + batch).randbytes(32)
            for count in (64 * batch, 1 << 18):
                hits, tries = g.batch_run([_spec(key, 0, b"ab", False, M.MODE_CONTAINS, 1000, count)])
                want = oracle.search(key, 0, 1000, tries[0], b"ab", False, M.MODE_CONTAINS)
                assert [h.index for h in hits] == want and want, (batch, blocks, count)
        finally:
            g.close()


def test_lockstep_stop_on_hit(pair, oracle):
    """Searches that stop on their first candidate: the waves whose claims are skipped pass
    the same barriers as the ones that search."""
    lock, _ = pair
    rng = random.Random(162)
    specs = [_spec(rng.randbytes(32), 0, n, True, M.MODE_CONTAINS, 0, 1 << 24, stop_on_hit=True)
             for n in (b"ab9", b"Zq7", b"k/x")]
    specs.append(_spec(rng.randbytes(32), 0, b"k7", False, M.MODE_SUFFIX, 0, 1 << 22))
    say("lockstep stop-on-hit batch")
    hits, tries = lock.batch_run(specs)
    for i in range(3):
        mine = [h for h in hits if h.search == i]
        assert mine, i
        for h in mine:
            assert h.pk == oracle.derive(specs[i]["key"], 0, h.index, 1)
            assert specs[i]["needle"].decode() in oracle.public_b64(h.pk)[25:]
    s = specs[3]
    want = oracle.search(s["key"], 0, 0, tries[3], s["needle"], False, M.MODE_SUFFIX)
    assert [h.index for h in hits if h.search == 3] == want


def test_lockstep_serving(pair, oracle):
    """Served requests (one 128-key unit per wave per step, round-robin over requests,
    stopped requests skipped) give each request's first exact hit, as without lockstep,
    also in 64-key units."""
    lock, plain = pair
    rng = random.Random(163)
    specs = []
    for i in range(12):
        text = "".join(rng.choice("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789") for _ in range(2))
        p = M.Pattern(text, i % 2 == 0, M.MODE_SUFFIX if i % 3 == 0 else M.MODE_CONTAINS)
        specs.append(p.spec(rng.randbytes(32), rng.getrandbits(40), 0))
    results = {}
    for name, g, unit in (("lock", lock, 128), ("lock64", lock, 64), ("plain", plain, 128)):
        say(f"serving on {name}")
        with M.Server(g, max_slots=8, step_keys=1 << 22, unit_keys=unit) as srv:
            ids = {srv.submit(s): k for k, s in enumerate(specs)}
            res = {}
            steps = 0
            while len(res) < len(specs):
                srv.step()
                steps += 1
                for r in srv.poll():
                    res[ids[r.request]] = (r.status, r.index, r.pk)
                assert steps < 4000
        results[name] = res
    assert results["lock"] == results["plain"] == results["lock64"]
    for k, s in enumerate(specs):
        status, index, pk = results["lock"][k]
        assert status == 1
        first = oracle.search(s["key"], 0, s["first"], index - s["first"] + 1, s["needle"], not s["fold"], s["mode"])
        assert first and first[0] == index and pk == oracle.derive(s["key"], 0, index, 1)
