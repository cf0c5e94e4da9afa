"""The LDS-table search kernel (mk_options.lds_table: each CU copies the whole comb table
into LDS, W <= 5, or W = 6 with 96-byte entries) returns exactly what the same table read
from HBM / L2 and the oracle return: mixed batches over both seed sources and every match
mode, stop-on-hit searches, served requests, the low-security workload's 8-char
case-insensitive substring with a planted hit, and the size checks."""
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


def test_lds_table_rejects_what_does_not_fit():
    for kw in (dict(width=6, limbs=M.LIMBS_R25U), dict(width=7, limbs=M.LIMBS_R25), dict(width=5, blocks=1),
               dict(width=5, pool=8), dict(width=5, lockstep=1)):
        with pytest.raises(M.MeikeyError):
            M.Gpu(0, batch=32, lds_table=1, **{"blocks": 2, **kw})


@pytest.mark.parametrize("width,limbs", [(5, M.LIMBS_R25U), (5, M.LIMBS_R25), (6, M.LIMBS_R25), (6, M.LIMBS_R32),
                                         (3, M.LIMBS_R25U), (5, M.LIMBS_K26U), (6, M.LIMBS_K26)])
def test_lds_table_batches_equal_global_and_oracle(oracle, width, limbs):
    say(f"open gpu W={width} limbs={limbs} lds_table=1 and the global-table twin")
    g = M.Gpu(0, width=width, batch=32, blocks=2, limbs=limbs, lds_table=1)
    ref_gpu = M.Gpu(0, width=width, batch=32, blocks=2, limbs=limbs)
    try:
        assert g.info().options.lds_table == 1
        rng = random.Random(500 + width + limbs)
        specs = []
        for count, needle, cs, mode in [((1 << 21) + 3 * 64 * 32, b"Kq", True, M.MODE_CONTAINS),
                                        (1 << 19, b"zz", False, M.MODE_SUFFIX),
                                        (5 * 64 * 32 + 7, b"mk", False, M.MODE_CONTAINS)]:
            specs.append(_spec(rng.randbytes(32), rng.choice([0, 1]), needle, cs, mode, rng.getrandbits(40), count))
        p = M.Pattern("^[A-D][0-9]", True, M.MODE_REGEX)
        specs.append(p.spec(rng.randbytes(32), 777, 1 << 18))
        say("batch")
        hits, tries = g.batch_run(specs)
        ref, rtries = ref_gpu.batch_run(specs)
        assert tries == rtries
        assert [(h.search, h.index, h.pk, h.seed) for h in hits] == [(h.search, h.index, h.pk, h.seed) for h in ref]
        for i, s in enumerate(specs):
            want = oracle.search(s["key"], s["seed_source"], s["first"], tries[i], s["needle"], not s["fold"],
                                 s["mode"])
            assert [h.index for h in hits if h.search == i] == want and want, i
    finally:
        g.close()
        ref_gpu.close()


@pytest.fixture(scope="module")
def lds5():
    say("open gpu W=5 lds_table=1")
    g = M.Gpu(0, width=5, batch=32, blocks=2, lds_table=1)
    yield g
    g.close()


def test_lds_table_stop_on_hit_and_serving(lds5, oracle):
    rng = random.Random(510)
    specs = [_spec(rng.randbytes(32), 0, n, True, M.MODE_CONTAINS, 0, 1 << 22, stop_on_hit=True)
             for n in (b"ab9", b"Zq7")]
    say("stop-on-hit batch")
    hits, _ = lds5.batch_run(specs)
    for i in range(2):
        mine = [h for h in hits if h.search == i]
        assert mine, i
        for h in mine:
            assert h.pk == oracle.derive(specs[i]["key"], 0, h.index, 1)
            assert specs[i]["needle"].decode() in oracle.public_b64(h.pk)[25:]
    key = rng.randbytes(32)
    spec = M.Pattern("Qz", True, M.MODE_SUFFIX).spec(key, 123, 0)
    say("served request")
    with M.Server(lds5, max_slots=4, step_keys=1 << 20) as srv:
        rid = srv.submit(spec)
        got = []
        for _ in range(4000):
            srv.step()
            got = srv.poll()
            if got:
                break
    assert got and got[0].request == rid and got[0].status == 1
    r = got[0]
    first = oracle.search(key, 0, 123, r.index - 123 + 1, b"Qz", True, M.MODE_SUFFIX)
    assert first and first[0] == r.index and r.pk == oracle.derive(key, 0, r.index, 1)


def test_lds_table_low_security_planted(lds5, oracle):
    """The low-security workload (counter seeds of the key of random.Random(3), an 8-char
    case-insensitive substring) with a hit planted at a known index."""
    import random as R
    key = R.Random(3).randbytes(32)
    target = (1 << 31) + 99
    pk = oracle.derive(key, 1, target, 1)
    needle = T.hay(pk)[10:18].lower().encode()
    first = target - (1 << 19)
    hits, tries = lds5.batch_run([_spec(key, 1, needle, False, M.MODE_CONTAINS, first, 1 << 20)])
    want = oracle.search(key, 1, first, tries[0], needle, False, M.MODE_CONTAINS, threads=16)
    assert target in want and [h.index for h in hits] == want
