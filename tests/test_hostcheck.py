"""The kernel arithmetic (meikey_amd/csrc/mk_*.h) compiled for the host, against the
oracle and Python integers: the same source the gfx950 kernels run, checked without a
GPU, including the limb bounds the kernels rely on."""
import random
import zlib
import struct

import pytest

from tests import textbook as T

P = T.P
MASK26, MASK25 = (1 << 26) - 1, (1 << 25) - 1


def test_seed_and_scalar(oracle, hostcheck):
    rng = random.Random(21)
    for src in (0, 1):
        key = rng.randbytes(32)
        for i in [0, 1, 2, 3, 2**32 - 1, 2**32, 2**64 - 1] + [rng.getrandbits(64) for _ in range(20)]:
            s = hostcheck.seed(key, src, i)
            assert s == oracle.seed(key, src, i)
            assert hostcheck.secret_scalar(s) == oracle.secret_scalar(s)


def test_counter_scalar_from_mid_state(oracle, hostcheck):
    """The SHA-512 mid-state path (rounds 0-3 and W16/W17 shared by a search) equals the
    full hash of each counter seed, including a key tail that wraps past 2^64."""
    rng = random.Random(22)
    keys = [rng.randbytes(32) for _ in range(3)] + [rng.randbytes(24) + b"\xff" * 8]
    for key in keys:
        for i in [0, 1, 2, 255, 2**32 - 1, 2**64 - 1] + [rng.getrandbits(64) for _ in range(20)]:
            want = oracle.secret_scalar(oracle.seed(key, 1, i))
            assert hostcheck.secret_scalar_counter(key, i) == want, (key.hex(), i)


def _rand_limbs(rng, even_max, odd_max):
    return [rng.randrange(0, (even_max if i % 2 == 0 else odd_max) + 1) for i in range(10)]


@pytest.mark.parametrize("case", ["carried", "sub", "t_side", "extreme"])
def test_fe_mul_within_bounds(hostcheck, case):
    """fe_mul is exact for the operand bounds the point formulas produce: g up to
    3*2^26 / 3*2^25 per limb (a subtraction result), f up to 2^28 (the t operand)."""
    rng = random.Random(zlib.crc32(case.encode()))
    for _ in range(300):
        if case == "carried":
            f = _rand_limbs(rng, MASK26, MASK25)
            g = _rand_limbs(rng, MASK26, MASK25)
        elif case == "sub":
            f = _rand_limbs(rng, 3 << 26, 3 << 25)
            g = _rand_limbs(rng, 3 << 26, 3 << 25)
        elif case == "t_side":
            f = _rand_limbs(rng, 1 << 28, 1 << 27)
            g = _rand_limbs(rng, (3 << 26) + (3 << 13), (3 << 25) + (3 << 13))
        else:
            f = [(1 << 28) if i % 2 == 0 else (1 << 27) for i in range(10)]
            g = [(3 << 26) + (3 << 13) if i % 2 == 0 else (3 << 25) + (3 << 13) for i in range(10)]
        r = hostcheck.fe_mul_limbs(f, g)
        assert all(x < (1 << 26) + (1 << 14) for x in r)
        want = hostcheck.value_of(f) * hostcheck.value_of(g) % P
        assert hostcheck.fe_to_int(r) == want
        assert hostcheck.value_of(r) < 2**255 + 2**142  # the bound fe_carried_to_words relies on
        assert hostcheck.fe_carried_to_int(r) == want
        q = hostcheck.fe_sq_limbs(r)
        assert hostcheck.fe_to_int(q) == hostcheck.value_of(r) ** 2 % P
        assert hostcheck.fe_carried_to_int(q) == hostcheck.value_of(r) ** 2 % P


def test_fe_ops_edge_values(oracle, hostcheck):
    from tests.test_oracle import EDGE
    rng = random.Random(22)
    vals = EDGE + [rng.getrandbits(256) for _ in range(12)]
    for a in vals:
        for b in vals[:6]:
            for op in (0, 1, 2):
                assert hostcheck.fe_op(op, a, b) == oracle.fe_op(op, a, b), (op, a, b)
        assert hostcheck.fe_op(3, a, 0) == a * a % P
        assert hostcheck.fe_op(4, a, 0) == pow(a, P - 2, P)
        assert hostcheck.fe_op(5, a, 0) == (-a) % P


def test_to_words_canonical(hostcheck):
    # values just below, at and above p in carried limb form
    for v in [P - 1, P, P + 1, P + 18, 2**255 - 1, 0, 1]:
        limbs = hostcheck.limbs_of(v % (1 << 255))
        assert hostcheck.fe_to_int(limbs) == v % P
        assert hostcheck.fe_carried_to_int(limbs) == v % P


def test_carried_to_words_at_the_carried_bound(hostcheck):
    """fe_carried_to_words takes what fe_mul returns without a carry pass: limb 1 up to
    2^25 + 2^18.3 and limb 5 up to 2^25 + 2^14 (values up to 2^255 + 2^142, above p)
    reduce to the canonical value."""
    top = [(1 << 26) - 1 if i % 2 == 0 else (1 << 25) - 1 for i in range(10)]
    cases = [top]
    for l1 in (1 << 25, (1 << 25) + (1 << 18), (1 << 25) + 321000):
        for l5 in ((1 << 25) - 1, (1 << 25) + (1 << 14) - 1):
            cases.append(top[:1] + [l1] + top[2:5] + [l5] + top[6:])
    rng = random.Random(25)
    for _ in range(300):
        c = _rand_limbs(rng, MASK26, MASK25)
        c[1] = rng.randrange(0, (1 << 25) + 321000)
        c[5] = rng.randrange(0, (1 << 25) + (1 << 14))
        cases.append(c)
    for c in cases:
        v = hostcheck.value_of(c)
        assert v < 2 * P
        assert hostcheck.fe_carried_to_int(c) == v % P, c
    assert any(hostcheck.value_of(c) >= P for c in cases)


def test_digits(oracle, hostcheck):
    rng = random.Random(23)
    for w in range(2, 25):
        for _ in range(5):
            s = oracle.secret_scalar(rng.randbytes(32))
            assert hostcheck.digits(s, w) == oracle.digits(s, w), w


def _entry_ints(words):
    from tests.hostcheck import entry_limbs, value_of
    return [value_of(entry_limbs(words, f)) for f in range(3)]


def test_entry_layout_is_a_permutation():
    """Every limb of the three fields has its own word, words 30-31 are free, and y+x /
    y-x occupy mirrored 16-byte chunks (0-1 vs 2-3, halves of chunk 6)."""
    from tests.hostcheck import entry_word
    used = sorted(entry_word(f, i) for f in range(3) for i in range(10))
    assert used == list(range(30))
    for i in range(8):
        assert entry_word(1, i) == entry_word(0, i) + 8
    assert [entry_word(0, 8) // 4, entry_word(1, 8) // 4] == [6, 6]


@pytest.mark.parametrize("w", [2, 3, 5, 8])
def test_table_matches_oracle(oracle, hostcheck, w):
    tab = hostcheck.table(w)
    rows, cols = -(-251 // w), (1 << (w - 1)) + 1
    rng = random.Random(w)
    picks = [(0, 0), (0, 1), (rows - 1, 0), (rows - 1, cols - 1)] + [
        (rng.randrange(rows), rng.randrange(cols)) for _ in range(12)]
    for row, k in picks:
        words = list(tab[(row * cols + k) * 32:(row * cols + k + 1) * 32])
        ypx, ymx, third = _entry_ints(words)
        n = oracle.comb_entry_niels(w, row, k)
        half = T.inv(2)  # every field halved: (y+x)/2, (y-x)/2, dxy
        assert ypx == int.from_bytes(n[0:32], "little") * half % P
        assert ymx == int.from_bytes(n[32:64], "little") * half % P
        xy2d = int.from_bytes(n[64:96], "little")
        if row == 0:
            assert third == xy2d * T.inv(2 * T.D) % P  # row 0 holds xy
        else:
            assert third == xy2d * half % P
        assert words[30] == 0 and words[31] == 0
        # canonical limbs
        from tests.hostcheck import entry_limbs
        assert all(l <= (MASK26 if i % 2 == 0 else MASK25) for f in range(3)
                   for i, l in enumerate(entry_limbs(words, f)))


@pytest.mark.parametrize("w", [3, 7, 8])
def test_full_pk_and_lane(oracle, hostcheck, w):
    tab = hostcheck.table(w)
    rng = random.Random(100 + w)
    key = rng.randbytes(32)
    for src in (0, 1):
        first = rng.getrandbits(40)
        K = 6
        pks = oracle.derive(key, src, first, K)
        for j in range(K):
            s = oracle.secret_scalar(oracle.seed(key, src, first + j))
            assert hostcheck.full_pk(tab, w, s) == pks[32 * j:32 * j + 32]
        needle = T.hay(pks[0:32])[10:13].encode()
        ys, cand = hostcheck.lane(tab, w, key, src, first, K, needle, False, 0)
        for j in range(K):
            pk = pks[32 * j:32 * j + 32]
            assert ys[j] == pk[:31] + bytes([pk[31] & 0x7F])
            assert cand[j] == int(oracle.match_unsigned(pk, needle, False, 0))
        assert cand[0] == 1


@pytest.mark.parametrize("K", [1, 2, 3, 64])
def test_lane_batched_inversion_edges(oracle, hostcheck, K):
    """The lane tile's batched inversion (Q_j = yn_j * P_(j-1) spilled with yd_j, one
    inversion, the back-sweep y_j = Q_j / P_j) gives every key's y for a tile of one, two
    and three keys, where the prefetch and the last step have no predecessor, and for 64."""
    w = 5
    tab = hostcheck.table(w)
    rng = random.Random(300 + K)
    key = rng.randbytes(32)
    for src in (0, 1):
        first = rng.getrandbits(40)
        pks = oracle.derive(key, src, first, K)
        needle = T.hay(pks[0:32])[10:13].encode()
        ys, _ = hostcheck.lane(tab, w, key, src, first, K, needle, False, 0)
        for j in range(K):
            pk = pks[32 * j:32 * j + 32]
            assert ys[j] == pk[:31] + bytes([pk[31] & 0x7F]), (K, src, j)


def test_bitap_matcher_vs_oracle(oracle, hostcheck):
    rng = random.Random(24)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    for _ in range(6000):
        pk = rng.randbytes(32)
        b64 = T.public_b64(pk)
        # lengths on both sides of the suffix automaton's start points (8 / 16) and of the
        # switch from 32-bit to 64-bit automaton states (32)
        L = rng.choice([1, 2, 3, 5, 7, 8, 9, 12, 16, 17, 30, 31, 32, 33, 42, 43])
        r = rng.random()
        if r < 0.4:
            needle = b64[68 - L:]  # a true suffix of the key
        elif r < 0.7:
            st = rng.randrange(25, 69 - L)
            needle = b64[st:st + L]
        else:
            needle = "".join(rng.choice(alpha) for _ in range(L))
        if rng.random() < 0.3:
            needle = needle.swapcase()
        cs, mode = rng.random() < 0.5, rng.choice([0, 1])
        nb = needle.encode()
        assert hostcheck.match(pk, nb, not cs, mode, False) == oracle.match(b64, nb, cs, mode)
        # sign-unknown matching takes y with the sign bit 0, as the search kernel has it
        y = pk[:31] + bytes([pk[31] & 0x7F])
        assert hostcheck.match(y, nb, not cs, mode, True) == oracle.match_unsigned(pk, nb, cs, mode)


def test_bitap_mask_folds_case(hostcheck):
    m = hostcheck.build_mask(b"aB", True)
    alpha = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    assert m[alpha.index("a")] == 1 and m[alpha.index("A")] == 1
    assert m[alpha.index("b")] == 2 and m[alpha.index("B")] == 2
    assert sum(1 for x in m if x) == 4
    m = hostcheck.build_mask(b"aB", False)
    assert m[alpha.index("a")] == 1 and m[alpha.index("A")] == 0 and m[alpha.index("B")] == 2


def test_ladder_public_key(oracle, hostcheck):
    """The table-free double-and-add used to re-derive emitted keys on the host."""
    rng = random.Random(26)
    for _ in range(40):
        seed = rng.randbytes(32)
        assert hostcheck.pk_by_ladder(oracle.secret_scalar(seed)) == oracle.public_key(seed)


def test_host_confirm_accepts_true_keys_and_refuses_others(oracle, hostcheck):
    """The confirmation before a key is handed out (the reference's confirm,
    gpu.rs:973-996, whose tests require a wrong hit to be refused): the true key passes;
    a flipped x sign, another index's seed or key, or a pattern it does not match fail."""
    import meikey_amd as M
    rng = random.Random(27)
    for source in (0, 1):
        key = rng.randbytes(32)
        for _ in range(6):
            i = rng.getrandbits(48)
            seed, pk = oracle.seed(key, source, i), oracle.derive(key, source, i, 1)
            b64 = oracle.public_b64(pk)
            needle = b64[60:66].encode()
            assert hostcheck.confirm(key, source, i, seed, pk, needle, False, 0) == 1
            assert hostcheck.confirm(key, source, i, seed, pk, b64[62:].encode(), False, 1) == 1
            dfa = M.Pattern(b64[40:44], True, M.MODE_REGEX).needle()[0]
            assert hostcheck.confirm(key, source, i, seed, pk, dfa, False, 2) == 1
            flipped = pk[:31] + bytes([pk[31] ^ 0x80])
            assert hostcheck.confirm(key, source, i, seed, flipped, needle, False, 0) == 0
            other = oracle.derive(key, source, i + 1, 1)
            assert hostcheck.confirm(key, source, i, seed, other, needle, False, 0) == 0
            assert hostcheck.confirm(key, source, i, oracle.seed(key, source, i + 1), pk, needle, False, 0) == 0
            assert hostcheck.confirm(key, source, i, seed, pk, b"+/+/+/", False, 0) == 0


@pytest.mark.parametrize("case", ["empty", "one", "dense", "wide", "sparse_searches", "duplicates"])
def test_hit_order(hostcheck, case):
    """The library's hit order (counting sort by search, radix sort by index offset) is a
    permutation that puts the hits in (search, index) order, as numpy's lexsort does,
    for index ranges that need 0 to 6 radix passes."""
    import numpy as np
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    if case == "empty":
        s, x = np.zeros(0, np.uint32), np.zeros(0, np.uint64)
    elif case == "one":
        s, x = np.array([7], np.uint32), np.array([2**63 + 5], np.uint64)
    elif case == "dense":  # 64 searches of 2^26 keys, as in a batched launch
        s = rng.integers(0, 64, 200000).astype(np.uint32)
        x = (rng.choice(1 << 26, 200000, replace=False).astype(np.uint64) + (s.astype(np.uint64) << np.uint64(40)))
    elif case == "wide":  # ranges of 2^12 .. 2^64 within a search
        s = rng.integers(0, 6, 30000).astype(np.uint32)
        span = np.array([1 << 12, 1 << 22, 1 << 33, 1 << 44, 1 << 55, 2**64 - 1], dtype=object)
        x = np.array([int(rng.integers(0, 2**63)) * 2 % int(span[k]) if span[k] > 1 else 0 for k in s], np.uint64)
        x[:6] = [0, 2**64 - 1, 1, 2**64 - 2, 2**32, 2**63]
        s[:6] = 5
    elif case == "sparse_searches":  # empty and one-hit groups between large ones
        s = rng.choice([0, 1, 3, 900, 4095], 5000).astype(np.uint32)
        s[0] = 2
        x = rng.integers(0, 2**40, 5000).astype(np.uint64)
    else:  # equal keys: any order among them, the same sorted keys
        s = rng.integers(0, 3, 3000).astype(np.uint32)
        x = rng.integers(0, 50, 3000).astype(np.uint64)
    order = hostcheck.sort_hits(s, x)
    assert sorted(order.tolist()) == list(range(len(s)))
    want = np.lexsort((x, s))
    assert s[order].tolist() == s[want].tolist()
    assert x[order].tolist() == x[want].tolist()
    # a second call on a shorter, reordered input reuses the buffers without carrying
    # anything over
    s2, x2 = s[::-2].copy(), x[::-2].copy()
    order2 = hostcheck.sort_hits(s2, x2)
    assert sorted(order2.tolist()) == list(range(len(s2)))
    want2 = np.lexsort((x2, s2))
    assert s2[order2].tolist() == s2[want2].tolist() and x2[order2].tolist() == x2[want2].tolist()


@pytest.mark.parametrize("K", [1, 8, 64, 128, 192, 200, 256])
def test_claims_cover_every_key_once(hostcheck, K):
    """The search kernel's claim layout (mk_claims.h): whole tiles, tail units, the halved
    final round. For launches of every size around the tail's boundaries, the claims the
    counters hand out cover the launch's 64 * K * tiles keys exactly once, halves
    (K / 2U keys per lane) appear only in the last round of the tail, one per wave, and
    serving steps are cut into units of the server's size (128 or 64 keys per lane) and
    never halved."""
    import numpy as np
    waves = 64
    for serving in (0, 128, 64):
        if serving:
            U = K // serving if K >= serving and K % serving == 0 else 1
        else:
            U = K // 64 if K >= 128 and K % 64 == 0 else 1
        for tiles in [0, 1, 2, 5, waves - 1, waves, waves + 1, 2 * waves - 1, 2 * waves, 2 * waves + 1, 3 * waves + 7,
                      10 * waves]:
            cap = tiles * 2 * max(U, 1) + 8
            starts = np.zeros(cap, np.uint64)
            keffs = np.zeros(cap, np.uint32)
            n = hostcheck.lib().hc_claims(K, serving, tiles, waves, starts.ctypes.data, keffs.ctypes.data, cap)
            assert n >= 0, (K, serving, tiles)
            s, k = starts[:n].astype(np.int64), keffs[:n].astype(np.int64)
            order = np.argsort(s, kind="stable")
            ends = s[order] + 64 * k[order]
            if n:
                assert s[order][0] == 0 and ends[-1] == 64 * K * tiles, (K, serving, tiles)
                assert (s[order][1:] == ends[:-1]).all(), (K, serving, tiles)  # no gap, no overlap
            else:
                assert tiles == 0
            halves = int((k < K // max(U, 1)).sum())
            if serving or U == 1:
                assert halves == 0
                assert set(k.tolist()) <= {K, K // max(U, 1)}
            elif tiles * U >= 2 * waves:
                # the final round: the last `waves` units of the tail, each as two halves
                assert halves == 2 * waves and (k[-2 * waves:] == K // (2 * U)).all(), (K, tiles)
            else:
                assert halves == 0
