"""The oracle against the reference's own kernel compiled as host C++ (oracle/_ref,
built from kernel.cu by oracle/Makefile where the meikey sources exist): seeds, the
SHA-512 scalar (full and mid-state), signed digits, full public keys from a packed comb
table, the base64 tail and substring / suffix / DFA matching with and without the x sign.
Skipped where the reference build is absent."""
import ctypes
import random
import struct

import pytest

from oracle import oracle as O
from tests import textbook as T

pytestmark = pytest.mark.skipif(O.ref() is None, reason="reference host build (oracle/_ref) not built")


def words(key: bytes):
    return (ctypes.c_uint32 * 8)(*struct.unpack("<8I", key))


def out32():
    return ctypes.create_string_buffer(32)


def test_seeds_and_scalars():
    R = O.ref()
    rng = random.Random(501)
    for _ in range(20):
        key = rng.randbytes(32)
        for i in [0, 1, 2, 2**32 - 1, 2**64 - 1, rng.getrandbits(64)]:
            o = out32()
            R.meikey_test_seed(words(key), i, o)
            assert o.raw == O.seed(key, 0, i), (key.hex(), i)
            R.meikey_test_counter_seed(words(key), i, o)
            assert o.raw == O.seed(key, 1, i)
            s = out32()
            R.meikey_test_secret(o.raw, s)
            assert s.raw == O.secret_scalar(o.raw)
            R.meikey_test_secret_mid(words(key), i, s)
            assert s.raw == O.secret_scalar(O.seed(key, 1, i))


@pytest.mark.parametrize("w", [2, 5, 8, 13, 16, 20])
def test_digits(w):
    R = O.ref()
    rng = random.Random(502 + w)
    rows = -(-251 // w)
    for _ in range(20):
        s = O.secret_scalar(rng.randbytes(32))
        got = (ctypes.c_int32 * rows)()
        R.meikey_test_digits(s, w, got)
        assert list(got) == O.digits(s, w)


def packed_table(w: int):
    """The reference's table layout (UNPACKED = 0, gpu.rs:1455-1480): per entry 24 words,
    (y+x, y-x, 2dxy) as canonical 8-word little-endian values, row 0 holding 2xy."""
    rows, cols = -(-251 // w), (1 << (w - 1)) + 1
    d_inv = T.inv(T.D)
    out = bytearray()
    for row in range(rows):
        for k in range(cols):
            n = O.comb_entry_niels(w, row, k)
            third = n[64:96]
            if row == 0:
                third = (int.from_bytes(third, "little") * d_inv % T.P).to_bytes(32, "little")
            out += n[0:64] + third
    return (ctypes.c_uint32 * (len(out) // 4)).from_buffer_copy(bytes(out))


@pytest.mark.parametrize("w", [4, 8])
def test_derive_batch_equals_oracle(w):
    R = O.ref()
    assert R.meikey_test_entry_words() == 24
    batch = R.meikey_test_batch()
    table = packed_table(w)
    rng = random.Random(503 + w)
    for source in (0, 1):
        key = rng.randbytes(32)
        first = rng.getrandbits(40)
        out = ctypes.create_string_buffer(32 * batch)
        R.meikey_test_derive_batch(table, w, words(key), source, first, out)
        assert out.raw == O.derive(key, source, first, batch), (w, source)


def test_tail_and_matching():
    """Matching with the x sign known and unknown, every mode, on random keys; the DFA
    tables come from this repository's compiler, so this also pins that the reference's
    DFA walk reads them as the oracle does."""
    import meikey_amd as M
    R = O.ref()
    rng = random.Random(504)
    alphabet = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/"
    regexes = [M.Pattern(t, cs, M.MODE_REGEX).needle()[0] for t, cs in
               [("ab", True), ("^[A-P]a|z$", False), ("[0-9][A-F]x", False), ("Q.$", True)]]
    for _ in range(3000):
        pk = O.public_key(rng.randbytes(32))
        tail = ctypes.create_string_buffer(44)
        R.meikey_test_tail_b64(pk, tail)
        assert tail.raw == O.public_b64(pk)[24:68].encode()
        cases = []
        for mode in (0, 1):
            for cs in (True, False):
                n = "".join(rng.choice(alphabet) for _ in range(rng.randint(1, 3)))
                needle = (n if cs else n.lower()).encode()
                cases.append((needle, len(needle), cs, mode))
        for tab, cs in zip(regexes, [True, False, False, True]):
            cases.append((tab, len(tab) // 130, cs, 2))
        for needle, n, cs, mode in cases:
            fold = 0 if (cs or mode == 2) else 1
            want = O.match(O.public_b64(pk), needle, cs, mode)
            assert bool(R.meikey_test_contains(pk, needle, n, fold, mode)) == want, (needle, mode)
            # sign unknown: the key as the search kernel has it, x sign bit clear
            pk_y = pk[:31] + bytes([pk[31] & 0x7F])
            want_u = O.match_unsigned(pk_y, needle, cs, mode)
            assert bool(R.meikey_test_contains_unsigned(pk_y, needle, n, fold, mode)) == want_u, (needle, mode)


def test_device_table_repacks_to_reference_layout(hostcheck):
    """The kernel headers' table (32-word entries of limbs, as the device builds it),
    repacked by ref_table_from_device, equals the reference layout built from the oracle,
    so the reference build can run on the device's own table."""
    import numpy as np
    w = 8
    dev = np.frombuffer(bytes(hostcheck.table(w)), dtype=np.uint32).reshape(-1, 32)
    assert bytes(O.ref_table_from_device(dev)) == bytes(packed_table(w))
