"""The allocation-free derivation (mk_derive_noalloc) through the C ABI: the same public
keys and seeds as mk_derive, on two calls in a row (the second smaller, so stale staging
contents past its end would show), for both seed sources; sizes beyond the staging array
are rejected. Runs last: it must pass alone in a fresh process and after every other test."""
import random
import sys

import pytest

import meikey_amd as M

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_derive_noalloc_equals_derive(oracle):
    say("open gpu W=16 batch=32 for mk_derive_noalloc")
    with M.Gpu(0, width=16, batch=32) as g:
        rng = random.Random(150)
        for source in (M.SEED_CHACHA20, M.SEED_COUNTER):
            key = rng.randbytes(32)
            for first, n in [((1 << 32) - 30000, 1 << 16), (rng.getrandbits(48), 777)]:
                say(f"noalloc derive source={source} first={first} n={n}")
                want_pk, want_seed = g.derive(key, source, first, n, seeds=True)
                got_pk, got_seed = g.derive_noalloc(key, source, first, n, seeds=True)
                assert got_pk == want_pk and got_seed == want_seed, (source, first, n)
                assert g.derive_noalloc(key, source, first, n) == want_pk
                for i in (0, n // 2, n - 1):
                    assert got_pk[32 * i:32 * i + 32] == oracle.derive(key, source, first + i, 1)
        assert g.derive_noalloc(b"\0" * 32, 0, 0, 0) == b""
        with pytest.raises(M.MeikeyError) as e:
            g.derive_noalloc(b"\0" * 32, 0, 0, (1 << 16) + 1)
        assert e.value.status == 3  # MK_ERR_OUT_OF_RANGE
