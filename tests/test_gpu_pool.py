"""The pooled inversion (mk_options.pool: the waves of a workgroup share one inversion,
the reference's block-pooled inversion) returns exactly what the per-lane inversion and
the oracle return: batches of mixed searches at 8, 32 and 256 keys per lane (whole tiles,
tail units, groups of claims that run past the end of the queue), workgroups of 8 and 4
waves, stop-on-hit searches (waves that pass the pool's barriers with nothing to search),
served requests, and the headline search at the production width with 8 keys per lane."""
import random
import sys

import pytest

import meikey_amd as M
from tests import textbook as T

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spec(key, source, needle, cs, mode, first, count, **kw):
    d = {"key": key, "seed_source": source, "needle": needle, "fold": not cs, "mode": mode, "first": first,
         "count": count}
    d.update(kw)
    return d


def _mixed_specs(rng, tile):
    specs = []
    for count, needle, cs, mode in [((1 << 23) + 3 * tile, b"Kq7", True, M.MODE_CONTAINS),
                                    (1 << 20, b"zz", False, M.MODE_SUFFIX),
                                    (5 * tile, b"mk", False, M.MODE_CONTAINS),
                                    (7 * tile + 1, b"Q", True, M.MODE_SUFFIX)]:
        specs.append(_spec(rng.randbytes(32), rng.choice([0, 1]), needle, cs, mode, rng.getrandbits(40), count))
    p = M.Pattern("^[A-D][0-9]", True, M.MODE_REGEX)
    specs.append(p.spec(rng.randbytes(32), 777, 1 << 19))
    return specs


def test_pool_rejects_bad_settings():
    for kw in (dict(pool=8, blocks=1), dict(pool=6), dict(pool=4, lockstep=1), dict(pool=4, blocks=4)):
        with pytest.raises(M.MeikeyError):
            M.Gpu(0, width=12, batch=8, **kw)


@pytest.mark.parametrize("batch,pool,blocks,limbs", [(8, 8, 2, M.LIMBS_R25U), (32, 8, 2, M.LIMBS_R25U),
                                                    (256, 8, 2, M.LIMBS_R25U), (16, 4, 2, M.LIMBS_R25U),
                                                    (16, 4, 1, M.LIMBS_R25U), (8, 4, 2, M.LIMBS_K26U),
                                                    (32, 8, 2, M.LIMBS_K26)])
def test_pool_batches_equal_per_lane_and_oracle(oracle, batch, pool, blocks, limbs):
    say(f"open gpu W=16 batch={batch} pool={pool} blocks={blocks} limbs={limbs} and the per-lane twin")
    g = M.Gpu(0, width=16, batch=batch, blocks=blocks, pool=pool, limbs=limbs)
    ref_gpu = M.Gpu(0, width=16, batch=batch, blocks=blocks, limbs=limbs)
    try:
        info = g.info()
        assert info.options.pool == pool
        assert info.scratch_bytes == ref_gpu.info().scratch_bytes
        specs = _mixed_specs(random.Random(400 + batch + pool + blocks + limbs), 64 * batch)
        say("pooled batch")
        hits, tries = g.batch_run(specs)
        ref, rtries = ref_gpu.batch_run(specs)
        assert tries == rtries
        assert [(h.search, h.index, h.pk, h.seed) for h in hits] == [(h.search, h.index, h.pk, h.seed) for h in ref]
        for i in range(1, len(specs)):
            s = specs[i]
            want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"],
                                 s["mode"])
            assert [h.index for h in hits if h.search == i] == want and want, i
    finally:
        g.close()
        ref_gpu.close()


@pytest.fixture(scope="module")
def pooled():
    say("open gpu W=16 batch=8 pool=8")
    g = M.Gpu(0, width=16, batch=8, blocks=2, pool=8)
    yield g
    g.close()


def test_pool_small_launches(pooled, oracle):
    """Launches smaller than one claim per wave of the grid."""
    key = random.Random(410).randbytes(32)
    for count in (64 * 8, 64 * 8 * 5 + 3, 1 << 16):
        hits, tries = pooled.batch_run([_spec(key, 0, b"ab", False, M.MODE_CONTAINS, 1000, count)])
        want = oracle.search(key, 0, 1000, tries[0], b"ab", False, M.MODE_CONTAINS)
        assert [h.index for h in hits] == want and want, count


def test_pool_stop_on_hit(pooled, oracle):
    """Searches that stop on their first candidate: waves whose claims are skipped hand the
    pool P = 1 and pass the same two barriers."""
    rng = random.Random(411)
    specs = [_spec(rng.randbytes(32), 0, n, True, M.MODE_CONTAINS, 0, 1 << 22, stop_on_hit=True)
             for n in (b"ab9", b"Zq7", b"k/x")]
    specs.append(_spec(rng.randbytes(32), 0, b"k7", False, M.MODE_SUFFIX, 0, 1 << 21))
    say("pooled stop-on-hit batch")
    hits, tries = pooled.batch_run(specs)
    for i in range(3):
        mine = [h for h in hits if h.search == i]
        assert mine, i
        for h in mine:
            assert h.pk == oracle.derive(specs[i]["key"], 0, h.index, 1)
            assert specs[i]["needle"].decode() in oracle.public_b64(h.pk)[25:]
    s = specs[3]
    want = oracle.search(s["key"], 0, 0, tries[3], s["needle"], False, M.MODE_SUFFIX)
    assert [h.index for h in hits if h.search == 3] == want


def test_pool_serving(pooled, oracle):
    """Served requests give each request's first exact hit."""
    rng = random.Random(412)
    specs = []
    for i in range(10):
        text = "".join(rng.choice("ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789") for _ in range(2))
        p = M.Pattern(text, i % 2 == 0, M.MODE_SUFFIX if i % 3 == 0 else M.MODE_CONTAINS)
        specs.append(p.spec(rng.randbytes(32), rng.getrandbits(40), 0))
    say("serving on the pooled kernel")
    with M.Server(pooled, max_slots=8, step_keys=1 << 21) as srv:
        ids = {srv.submit(s): k for k, s in enumerate(specs)}
        res = {}
        steps = 0
        while len(res) < len(specs):
            srv.step()
            steps += 1
            for r in srv.poll():
                res[ids[r.request]] = (r.status, r.index, r.pk)
            assert steps < 4000
    for k, s in enumerate(specs):
        status, index, pk = res[k]
        assert status == 1
        first = oracle.search(s["key"], 0, s["first"], index - s["first"] + 1, s["needle"], not s["fold"], s["mode"])
        assert first and first[0] == index and pk == oracle.derive(s["key"], 0, index, 1)


def test_pool_headline_shape(oracle):
    """W = 24 with 8 keys per lane and one inversion per 8-wave workgroup: a planted
    6-char case-sensitive suffix over a 2^22 window of the headline key gives the oracle's
    full hit set."""
    from meikey_amd import workloads as W
    say("open gpu W=24 batch=8 pool=8")
    g = M.Gpu(0, width=24, batch=8, blocks=2, pool=8)
    try:
        key, _ = W.single()
        target = (1 << 30) + 4321
        pk = g.derive(key, 0, target, 1)
        assert pk == oracle.derive(key, 0, target, 1)
        needle = T.hay(pk)[-6:].encode()
        first = target - (1 << 21)
        hits, tries = g.batch_run([_spec(key, 0, needle, True, M.MODE_SUFFIX, first, 1 << 22)])
        want = oracle.search(key, 0, first, tries[0], needle, True, M.MODE_SUFFIX, threads=16)
        assert target in want and [h.index for h in hits] == want
    finally:
        g.close()
