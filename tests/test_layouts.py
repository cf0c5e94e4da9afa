"""The packed ten-limb ("10"), eight-word ("8") and Karatsuba ("10k", "10ku") builds of
the kernel arithmetic, compiled for the host (libmeikey_hostcheck_p10.so / _r32.so / _k26.so /
_k26u.so from the same headers as the gfx950 kernels), against Python integers, the oracle
and, where it was built, the reference's own eight-word and Karatsuba fields (kernel.cu:650-848
and 1102-1316 host builds, prefixes meikey_b32_l8_test_ and meikey_b32_l26_test_)."""
import ctypes
import random
import struct

import pytest

from tests import hostcheck as H
from tests import textbook as T

P = T.P
LAYOUTS = {"_p10": 101, "_r32": 81, "_k26": 261, "_k26u": 260}
PACKED = ["_k26", "_p10", "_r32"]
K26 = ["_k26", "_k26u"]
W8 = ctypes.c_uint32 * 8
L10 = ctypes.c_uint32 * 10

EDGES = [0, 1, 2, 19, 38, P - 1, P, P + 1, 2 * P - 1, 2 * P, 2 ** 255 - 1, 2 ** 255, 2 ** 256 - 39, 2 ** 256 - 38,
         2 ** 256 - 37, 2 ** 256 - 1, 2 ** 204 - 1, 2 ** 224 + 5, 2 ** 32 - 1, 2 ** 64 - 1]


def w8(x: int):
    return W8(*[(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)])


def val(a) -> int:
    return sum(int(a[i]) << (32 * i) for i in range(8))


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
def test_layout_codes(layout):
    assert H.load(layout).hc_layout() == LAYOUTS[layout]


def test_r32_field_on_any_256_bit_value():
    """Every operation of the eight-word field takes any value below 2^256 (the edge
    values next to p, 2p and 2^256, where the carries fold by 38) and returns a value below
    2^256 that is congruent to the exact result; canonical bytes are the value mod p."""
    L = H.load("_r32")
    rng = random.Random(170)
    vals = EDGES + [rng.getrandbits(256) for _ in range(400)]
    pairs = [(a, b) for a in EDGES for b in EDGES] + [(rng.choice(vals), rng.choice(vals)) for _ in range(3000)]
    for a, b in pairs:
        for op, want in ((0, a + b), (1, a - b), (2, a * b), (5, -a)):
            out = W8()
            L.hc_fe_op_limbs(op, w8(a), w8(b), out)
            r = val(out)
            assert r < 2 ** 256 and r % P == want % P, (op, hex(a), hex(b), hex(r))
    for a in vals:
        out = W8()
        L.hc_fe_to_words(w8(a), out)
        assert val(out) == a % P, hex(a)
    for a in vals[:120]:
        out = W8()
        L.hc_fe_op(4, w8(a), w8(0), out)  # inversion, canonical output
        assert val(out) == pow(a, P - 2, P), hex(a)


def test_r32_field_equals_the_reference_r32():
    """The reference's eight-word field (its host build, MEIKEY_FIELD=8) gives the same
    canonical results on the same 256-bit inputs, including sums and differences fed
    into a multiplication (its ops 5 and 6)."""
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        pytest.skip("the reference kernel's host build (oracle/_ref) is absent")
    fn = R.meikey_b32_l8_test_fe_op_words
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L = H.load("_r32")
    rng = random.Random(171)
    vals = EDGES + [rng.getrandbits(256) for _ in range(200)]
    for i in range(1500):
        a, b = rng.choice(vals), rng.choice(vals)
        for op in range(5):
            ref = ctypes.create_string_buffer(32)
            fn(op, w8(a), w8(b), ref)
            mine = W8()
            L.hc_fe_op(op, w8(a), w8(b), mine)
            assert int.from_bytes(ref.raw, "little") == val(mine), (op, hex(a), hex(b))
        for op, f in ((5, lambda x, y: (x + y) * (y + x)), (6, lambda x, y: (x - y) * (y - x))):
            ref = ctypes.create_string_buffer(32)
            fn(op, w8(a), w8(b), ref)
            assert int.from_bytes(ref.raw, "little") == f(a, b) % P, (op, hex(a), hex(b))


def _entry_fields(words):
    return [int.from_bytes(struct.pack("<8I", *words[8 * f:8 * f + 8]), "little") for f in range(3)]


@pytest.mark.parametrize("layout", PACKED)
@pytest.mark.parametrize("w", [3, 8])
def test_packed_table_matches_oracle(oracle, layout, w):
    """Packed entries: three canonical 256-bit fields ((y+x)/2, (y-x)/2, dxy; row 0 xy) in
    words 0-7 / 8-15 / 16-23, equal to the oracle's entries halved."""
    L = H.load(layout)
    rows, cols = -(-251 // w), (1 << (w - 1)) + 1
    assert L.hc_table_words(w) == rows * cols * 24
    rng = random.Random(172 + w)
    picks = [(0, 0), (0, 1), (rows - 1, 0), (rows - 1, cols - 1)] + [
        (rng.randrange(rows), rng.randrange(cols)) for _ in range(12)]
    half = T.inv(2)
    for row, k in picks:
        out = (ctypes.c_uint32 * 32)()
        L.hc_table_entry(w, row, k, out)
        ypx, ymx, third = _entry_fields(list(out)[:24])
        assert max(ypx, ymx, third) < P  # canonical
        n = oracle.comb_entry_niels(w, row, k)
        assert ypx == int.from_bytes(n[0:32], "little") * half % P
        assert ymx == int.from_bytes(n[32:64], "little") * half % P
        xy2d = int.from_bytes(n[64:96], "little")
        assert third == (xy2d * T.inv(2 * T.D) % P if row == 0 else xy2d * half % P)


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
@pytest.mark.parametrize("w", [4, 7])
def test_layout_full_pk_and_lane(oracle, layout, w):
    """Public keys (x sign included) and a search lane's batched inversion and candidates,
    from each layout's table, equal the oracle's for both seed sources."""
    L = H.load(layout)
    n = L.hc_table_words(w)
    tab = (ctypes.c_uint32 * n)()
    L.hc_table_build(w, tab)
    rng = random.Random(180 + w)
    key = rng.randbytes(32)
    for src in (0, 1):
        first = rng.getrandbits(40)
        K = 5
        pks = oracle.derive(key, src, first, K)
        for j in range(K):
            s = oracle.secret_scalar(oracle.seed(key, src, first + j))
            out = W8()
            L.hc_full_pk(tab, w, H.words(s), out)
            assert H.wbytes(out) == pks[32 * j:32 * j + 32], (layout, src, j)
        needle = T.hay(pks[0:32])[10:13].encode()
        ys = (ctypes.c_uint32 * (8 * K))()
        cand = (ctypes.c_uint8 * K)()
        assert L.hc_lane(tab, w, H.words(key), src, first, K, needle, len(needle), 0, 0, ys, cand) == 0
        for j in range(K):
            pk = pks[32 * j:32 * j + 32]
            assert struct.pack("<8I", *ys[8 * j:8 * j + 8]) == pk[:31] + bytes([pk[31] & 0x7F])
            assert cand[j] == int(oracle.match_unsigned(pk, needle, False, 0))
        assert cand[0] == 1


def k26_value(limbs) -> int:
    return sum(int(l) << (26 * i) for i, l in enumerate(limbs))


@pytest.mark.parametrize("layout", K26)
def test_k26_field_on_values_next_to_p_and_2_255(layout):
    """Karatsuba field ("10k" / "10ku"): every operation on 256-bit inputs (bit 255 kept in
    limb 9) gives the canonical exact result mod p, at the edge values next to p, 2p, 2^255
    and 2^256 and on random values; inversion is a^(p-2)."""
    L = H.load(layout)
    rng = random.Random(260)
    vals = EDGES + [rng.getrandbits(256) for _ in range(300)]
    pairs = [(a, b) for a in EDGES for b in EDGES] + [(rng.choice(vals), rng.choice(vals)) for _ in range(3000)]
    for a, b in pairs:
        for op, want in ((0, a + b), (1, a - b), (2, a * b), (3, a * a), (5, -a)):
            out = W8()
            L.hc_fe_op(op, w8(a), w8(b), out)
            assert val(out) == want % P, (layout, op, hex(a), hex(b))
    for a in vals[:100]:
        out = W8()
        L.hc_fe_op(4, w8(a), w8(0), out)
        assert val(out) == pow(a, P - 2, P), hex(a)


@pytest.mark.parametrize("layout", K26)
def test_k26_mul_at_its_limb_bounds(layout):
    """fe_mul and fe_sq take any limbs below 2^29 on either side (the middle term's signed
    products and the two-part fold of the high columns must not overflow a 64-bit column)
    and return carried limbs (limb 1 below 2^26 + 2^22, the others below 2^26 + 2^13);
    fe_carried_to_words gives the canonical value of limbs below 2^30. Limbs at the bounds
    themselves, zero and random in between."""
    L = H.load(layout)
    rng = random.Random(261)
    top = (1 << 29) - 1
    fixed = [[top] * 10, [0] * 10, [top, 0] * 5, [0, top] * 5, [top] * 5 + [0] * 5, [0] * 5 + [top] * 5]
    cases = [(f, g) for f in fixed for g in fixed] + [
        ([rng.randrange(1 << 29) for _ in range(10)], [rng.randrange(1 << 29) for _ in range(10)]) for _ in range(4000)]
    for f, g in cases:
        out = L10()
        L.hc_fe_op_limbs(2, L10(*f), L10(*g), out)
        assert k26_value(out) % P == k26_value(f) * k26_value(g) % P, (f, g)
        assert out[1] < (1 << 26) + (1 << 22) and all(out[i] < (1 << 26) + (1 << 13) for i in range(10) if i != 1)
        sq = L10()
        L.hc_fe_sq_limbs(L10(*f), sq)
        assert k26_value(sq) % P == k26_value(f) ** 2 % P, f
        assert sq[1] < (1 << 26) + (1 << 22) and all(sq[i] < (1 << 26) + (1 << 13) for i in range(10) if i != 1)
    for _ in range(2000):
        x = [rng.randrange(1 << 30) for _ in range(10)]
        out = W8()
        L.hc_fe_carried_to_words(L10(*x), out)
        assert val(out) == k26_value(x) % P
    for x in ([(1 << 30) - 1] * 10, [0] * 10, [(1 << 26) - 1] * 9 + [(1 << 21) - 1]):
        out = W8()
        L.hc_fe_carried_to_words(L10(*x), out)
        assert val(out) == k26_value(x) % P


def test_k26_field_equals_the_reference_k26():
    """The reference's Karatsuba field (its host build, MEIKEY_FIELD=26) gives the same
    canonical bytes on the same inputs below 2^255 for every one of its test operations:
    sums, differences, products, squares, inversions, and products of sums and of
    differences (its limb-bound cases 5 and 6) and the inversion of a difference (7)."""
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        pytest.skip("the reference kernel's host build (oracle/_ref) is absent")
    fn = R.meikey_b32_l26_test_fe_op_words
    fn.argtypes = [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L = H.load("_k26")
    rng = random.Random(262)
    edges = [e for e in EDGES if e < 2 ** 255]
    vals = edges + [rng.getrandbits(255) for _ in range(200)]
    mine = {0: lambda a, b: (a + b) % P, 1: lambda a, b: (a - b) % P, 2: lambda a, b: a * b % P,
            3: lambda a, b: a * a % P, 4: lambda a, b: pow(a, P - 2, P)}
    for i in range(1500):
        a, b = rng.choice(vals), rng.choice(vals)
        for op in range(8):
            ref = ctypes.create_string_buffer(32)
            fn(op, w8(a), w8(b), ref)
            got = int.from_bytes(ref.raw, "little")
            if op < 5:
                out = W8()
                L.hc_fe_op(op, w8(a), w8(b), out)
                assert got == val(out) == mine[op](a, b), (op, hex(a), hex(b))
            elif op == 5:
                assert got == (a + b) * (b + a) % P
            elif op == 6:
                assert got == (a - b) * (b - a) % P
            else:
                assert got == pow((a - b) % P, P - 2, P)


@pytest.mark.parametrize("w", [3, 8])
def test_k26u_table_holds_canonical_26_bit_limbs(oracle, w):
    """"10ku" entries: three fields of ten 26-bit limbs at the 128-byte entry's word
    positions (limbs 0-7 in words 8f .. 8f+7, limbs 8-9 in words 24+2f, 25+2f), canonical
    (limbs below 2^26, limb 9 below 2^21), equal to the oracle's entries halved."""
    from tests.hostcheck import entry_limbs
    L = H.load("_k26u")
    rows, cols = -(-251 // w), (1 << (w - 1)) + 1
    assert L.hc_table_words(w) == rows * cols * 32
    rng = random.Random(263 + w)
    picks = [(0, 0), (0, 1), (rows - 1, 0), (rows - 1, cols - 1)] + [
        (rng.randrange(rows), rng.randrange(cols)) for _ in range(12)]
    half = T.inv(2)
    for row, k in picks:
        out = (ctypes.c_uint32 * 32)()
        L.hc_table_entry(w, row, k, out)
        words = list(out)
        assert words[30] == 0 and words[31] == 0
        fields = [entry_limbs(words, f) for f in range(3)]
        for limbs in fields:
            assert all(x < (1 << 26) for x in limbs) and limbs[9] < (1 << 21)
        ypx, ymx, third = (k26_value(limbs) for limbs in fields)
        n = oracle.comb_entry_niels(w, row, k)
        assert ypx == int.from_bytes(n[0:32], "little") * half % P
        assert ymx == int.from_bytes(n[32:64], "little") * half % P
        xy2d = int.from_bytes(n[64:96], "little")
        assert third == (xy2d * T.inv(2 * T.D) % P if row == 0 else xy2d * half % P)
