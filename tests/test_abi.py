"""The C ABI library loads and exports every entry point declared in include/meikey_hip.h
(no device calls: this runs on machines without a GPU)."""
import ctypes
import os
import re

import meikey_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "meikey_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:mk_status|void|uint64_t|uint32_t|int|size_t|double)\s+(mk_\w+)\s*\(", text, flags=re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "mk_open" in names and "mk_launch_run" in names and "mk_batch_run" in names
    assert sorted(meikey_amd.EXPORTS) == names


def test_library_exports_every_symbol():
    assert os.path.exists(meikey_amd.LIB_PATH), "run __graft_entry__.build() first"
    L = ctypes.CDLL(meikey_amd.LIB_PATH)
    for name in declared():
        assert hasattr(L, name), name


def test_bindings_load():
    L = meikey_amd.lib()
    assert L.mk_keys_per_iter.restype is ctypes.c_uint64


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(meikey_amd, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(meikey_amd, "_lib", None)
    try:
        meikey_amd.lib()
    except meikey_amd.MeikeyError as e:
        assert "missing" in str(e)
    else:
        raise AssertionError("expected MeikeyError")
