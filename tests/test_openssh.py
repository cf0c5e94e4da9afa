"""OpenSSH output: the authorized_keys line and the openssh-key-v1 private key.

Checked against the known line of the reference (src/model/key.rs:161-169), an
independent PROTOCOL.key parser written here, and ssh-keygen itself, which must derive
the same public key from the private file."""
import base64
import os
import random
import shutil
import struct
import subprocess

import pytest

import meikey_amd as M
from tests import textbook as T

SK1 = bytes.fromhex("9d61b19deffd5a60ba844af492ec2cc44449c5697b326919703bac031cae7f60")
PK1 = bytes.fromhex("d75a980182b10ab7d54bfed3c964073a0ee172f3daa62325af021a68f707511a")


def test_public_line():
    assert M.openssh_public(PK1) == "ssh-ed25519 AAAAC3NzaC1lZDI1NTE5AAAAINdamAGCsQq31Uv+08lkBzoO4XLz2qYjJa8CGmj3B1Ea"
    assert M.openssh_public(PK1, "me@host").endswith(" me@host")
    with pytest.raises(M.MeikeyError):
        M.openssh_public(PK1, "a\nb")


def parse_private(text: str):
    lines = text.strip().split("\n")
    assert lines[0] == "-----BEGIN OPENSSH PRIVATE KEY-----"
    assert lines[-1] == "-----END OPENSSH PRIVATE KEY-----"
    assert all(len(l) <= 70 for l in lines[1:-1]) and all(len(l) == 70 for l in lines[1:-2])
    raw = base64.b64decode("".join(lines[1:-1]))
    assert raw.startswith(b"openssh-key-v1\0")
    p = 15

    def u32():
        nonlocal p
        v = struct.unpack(">I", raw[p:p + 4])[0]
        p += 4
        return v

    def string():
        nonlocal p
        n = u32()
        v = raw[p:p + n]
        p += n
        return v

    cipher, kdf, opts, n = string(), string(), string(), u32()
    pub = string()
    priv = string()
    assert p == len(raw)
    return cipher, kdf, opts, n, pub, priv


def test_private_structure():
    text = M.openssh_private(SK1, PK1, "me@host", 0xDEADBEEF)
    cipher, kdf, opts, n, pub, priv = parse_private(text)
    assert (cipher, kdf, opts, n) == (b"none", b"none", b"", 1)
    assert pub == struct.pack(">I", 11) + b"ssh-ed25519" + struct.pack(">I", 32) + PK1
    c1, c2 = struct.unpack(">II", priv[:8])
    assert c1 == c2 == 0xDEADBEEF
    q = 8
    def s():
        nonlocal q
        k = struct.unpack(">I", priv[q:q + 4])[0]
        v = priv[q + 4:q + 4 + k]
        q += 4 + k
        return v
    assert s() == b"ssh-ed25519" and s() == PK1 and s() == SK1 + PK1 and s() == b"me@host"
    pad = priv[q:]
    assert len(priv) % 8 == 0 and pad == bytes(range(1, len(pad) + 1))


@pytest.mark.skipif(shutil.which("ssh-keygen") is None, reason="ssh-keygen not installed")
def test_ssh_keygen_reads_the_key(tmp_path, oracle):
    rng = random.Random(3)
    for i in range(4):
        seed = rng.randbytes(32)
        pk = oracle.public_key(seed)
        comment = f"vanity-{i}"
        path = tmp_path / f"k{i}"
        path.write_text(M.openssh_private(seed, pk, comment))
        os.chmod(path, 0o600)
        out = subprocess.run(["ssh-keygen", "-y", "-f", str(path)], check=True, capture_output=True, text=True)
        assert out.stdout.split()[:2] == M.openssh_public(pk).split()
        assert T.public_b64(pk) == out.stdout.split()[1]
