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


STRUCTS = {  # C struct -> ctypes mirror in meikey_amd
    "mk_options": "Options", "mk_info": "Info", "mk_hit_rec": "HitRec", "mk_search_spec": "SearchSpec",
    "mk_server_options": "ServerOptions", "mk_server_result": "ServerResult", "mk_stats": "Stats",
    "mk_search_config": "SearchConfig", "mk_outcome": "Outcome", "mk_tune_result": "TuneResult",
}


def test_struct_layouts_match_the_header(tmp_path):
    """sizeof and every field offset of the C structs (gcc on include/meikey_hip.h) equal
    the ctypes mirrors, so C, Python and a Rust #[repr(C)] binding agree on the bytes."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        import pytest
        pytest.skip("gcc not installed")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "meikey_hip.h"', "int main(void) {"]
    for c, py in STRUCTS.items():
        cls = getattr(meikey_amd, py)
        lines.append(f'printf("{c} size %zu\\n", sizeof({c}));')
        for f in cls._fields_:
            lines.append(f'printf("{c} {f[0]} %zu\\n", offsetof({c}, {f[0]}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for c, py in STRUCTS.items():
        cls = getattr(meikey_amd, py)
        assert int(got[f"{c} size"]) == ctypes.sizeof(cls), c
        for f in cls._fields_:
            assert int(got[f"{c} {f[0]}"]) == getattr(cls, f[0]).offset, (c, f[0])


def test_no_device_fails_loudly():
    """Without a GPU every search entry point fails with a status and a message; there is
    no CPU fallback. (Skipped where a device is visible.)"""
    import pytest
    if meikey_amd.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(meikey_amd.MeikeyError) as e:
        meikey_amd.Gpu(0)
    assert e.value.status == 2 and "no HIP device" in str(e.value)
    with pytest.raises(meikey_amd.MeikeyError):
        meikey_amd.fastest_staged(0)


def test_constants_match_the_header():
    """Every numeric #define of include/meikey_hip.h that the Python layer mirrors (seed
    sources, match modes, the five layout codes) has the header's value, and the
    --gpu-limbs codes map onto the layouts one to one (the floating-point "5f" has no code)."""
    text = open(os.path.join(ROOT, "include", "meikey_hip.h")).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define MK_(\w+) (\d+)u?\b", text, flags=re.M)}
    mirrored = [k for k in defs if hasattr(meikey_amd, k)]
    assert sorted(mirrored) == sorted(["SEED_CHACHA20", "SEED_COUNTER", "MODE_CONTAINS", "MODE_SUFFIX", "MODE_REGEX",
                                       "LIMBS_R25", "LIMBS_R25U", "LIMBS_R32", "LIMBS_K26", "LIMBS_K26U"])
    for k in mirrored:
        assert getattr(meikey_amd, k) == defs[k], k
    layouts = {v for k, v in defs.items() if k.startswith("LIMBS_")}
    assert set(meikey_amd.LIMBS_BY_CODE.values()) == layouts and len(meikey_amd.LIMBS_BY_CODE) == len(layouts)
    assert sorted(meikey_amd.LIMBS_BY_CODE) == ["10", "10k", "10ku", "10u", "8"]
    assert set(meikey_amd.ENTRY_WORDS) == layouts
