"""Build the gfx950 search library and its CPU test build, in-tree.

libmeikey_hip.so      hipcc --offload-arch=gfx950: kernels + C ABI (include/meikey_hip.h), plus the
                      host re-derivation of emitted keys compiled by g++
libmeikey_hostcheck.so  g++: the same arithmetic headers compiled for the host, test hooks only
mk_example            gcc: a C client of the ABI (open, pattern, search, OpenSSH line)
mk_valu_probe         hipcc: issue rates of the VALU instructions the search kernel is built from
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmeikey_hip.so")
HOSTCHECK = os.path.join(HERE, "libmeikey_hostcheck.so")
EXAMPLE = os.path.join(HERE, "mk_example")
PROBE = os.path.join(HERE, "mk_valu_probe")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# gfx950 only; packed-FP32 instructions stay off (the kernels are integer code anyway).
# The ILP-first machine scheduler orders the search kernel's multiply-add columns so that
# fewer issue slots go idle: 0.1-0.55% faster than the default schedule, four A/B pairs
# on two boxes (same VGPR count, no scratch).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-sched-strategy=max-ilp"]
# mk_kernels.hip builds the "10u" layout; mk_kernels_p10.hip, mk_kernels_r32.hip,
# mk_kernels_k26.hip and mk_kernels_k26u.hip include it for the packed ten-limb, the
# eight-word and the two Karatsuba layouts.
HIP_SOURCES = [os.path.join(CSRC, f) for f in ("mk_kernels.hip", "mk_kernels_p10.hip", "mk_kernels_r32.hip",
                                                 "mk_kernels_k26.hip", "mk_kernels_k26u.hip", "mk_host.cpp",
                                                 "mk_pattern.cpp", "mk_serve.cpp", "mk_driver.cpp")]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout)
        raise RuntimeError(f"build failed ({res.returncode}): {' '.join(cmd)}")


# Host-only parts of the library, compiled by the host compiler and linked in.
HOST_SOURCES = [os.path.join(CSRC, "mk_confirm_host.cpp")]


def _host_sanitizer_flags(kind: str) -> list[str]:
    # host pass only: GPU code is never built with a sanitizer
    flags = ["-Xarch_host", "-fsanitize=" + kind, "-Xarch_host", "-fno-omit-frame-pointer"]
    if kind == "undefined":
        flags += ["-Xarch_host", "-fno-sanitize-recover=undefined"]
    return flags


def build_hip(out: str = LIB, host_sanitize: str | None = None) -> str:
    """Each source compiled to an object in parallel (every HIP source carries its own gfx950
    code object), then linked by hipcc. host_sanitize ("address" or "undefined") builds the
    library's host code under that sanitizer, for CPU test runs with MEIKEY_HIP_LIB pointing
    at `out` and the compiler's runtime preloaded
    (LD_PRELOAD=/opt/rocm/llvm/lib/clang/<version>/lib/linux/libclang_rt.asan-x86_64.so, or
    libclang_rt.ubsan_standalone-x86_64.so)."""
    from concurrent.futures import ThreadPoolExecutor
    hip_flags = HIP_FLAGS + (_host_sanitizer_flags(host_sanitize) if host_sanitize else [])
    tmp = out + ".tmp"
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    jobs = []
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, os.path.basename(src) + ".host.o")
        jobs.append((["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-c", "-o", obj, src],
                     obj))
    compile_flags = [f for f in hip_flags if f != "-shared"]
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, os.path.basename(src) + ".o")
        jobs.append(([HIPCC, *compile_flags, "-c", "-o", obj, src], obj))
    try:
        workers = max(1, min(len(jobs), os.cpu_count() or 1, 16))
        with ThreadPoolExecutor(workers) as ex:
            list(ex.map(lambda j: _run(j[0]), jobs))
        # the host objects go straight to the linker (hipcc would read them as HIP sources)
        hip_objs = [o for _, o in jobs if not o.endswith(".host.o")]
        host_objs = [o for _, o in jobs if o.endswith(".host.o")]
        _run([HIPCC, *hip_flags, "-o", tmp, *hip_objs, *[a for o in host_objs for a in ("-Xlinker", o)]])
    finally:
        for _, obj in jobs:
            if os.path.exists(obj):
                os.remove(obj)
    os.replace(tmp, out)
    return out


# The CPU test builds of the arithmetic headers, one per layout (as the kernels are built).
HOSTCHECK_LAYOUTS = {"": [], "_p10": ["-DMK_PACKED=1"], "_r32": ["-DMK_FIELD=8", "-DMK_PACKED=1"],
                     "_k26": ["-DMK_FIELD=26", "-DMK_PACKED=1"], "_k26u": ["-DMK_FIELD=26"]}


def hostcheck_path(layout: str = "") -> str:
    return os.path.join(HERE, f"libmeikey_hostcheck{layout}.so")


def build_hostcheck(out: str = HOSTCHECK, sanitize: str | None = None) -> str:
    """sanitize: None, or a host sanitizer of the CPU build ("undefined": any undefined
    behaviour aborts the process; "address": run Python with
    LD_PRELOAD=$(g++ -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0)."""
    extra = []
    if sanitize:
        extra = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=" + sanitize]
        if sanitize == "undefined":
            extra.append("-fno-sanitize-recover=undefined")
    for layout, defs in HOSTCHECK_LAYOUTS.items():
        path = out if layout == "" else hostcheck_path(layout)
        tmp = path + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-function", *extra, *defs, "-o",
              tmp, os.path.join(CSRC, "mk_hostcheck.cpp")])
        os.replace(tmp, path)
    return out


def build_example(out: str = EXAMPLE) -> str:
    """The C client of the ABI (csrc/mk_example.c), linked against the in-tree library."""
    include = os.path.join(os.path.dirname(HERE), "include")
    _run(["gcc", "-O2", "-std=c11", "-Wall", "-I", include, "-o", out, os.path.join(CSRC, "mk_example.c"),
          "-L", HERE, "-lmeikey_hip", "-Wl,-rpath,$ORIGIN"])
    return out


def build_probe(out: str = PROBE) -> str:
    """The VALU issue-rate probe (csrc/mk_valu_probe.hip), a stand-alone program."""
    flags = [f for f in HIP_FLAGS if f not in ("-shared", "-fPIC")]
    _run([HIPCC, *flags, "-o", out, os.path.join(CSRC, "mk_valu_probe.hip")])
    return out


def build_all() -> None:
    build_hostcheck()
    build_hip()
    build_example()
    build_probe()


if __name__ == "__main__":
    if len(sys.argv) == 3 and sys.argv[1] == "--hostcheck-sanitize":
        # the CPU test build only, under a host sanitizer; a plain build restores it
        build_hostcheck(sanitize=sys.argv[2])
        print("built", HOSTCHECK, "with -fsanitize=" + sys.argv[2])
    elif len(sys.argv) == 4 and sys.argv[1] == "--host-sanitize":
        # a copy of the library with its host code under a sanitizer, at the path given
        build_hip(sys.argv[3], host_sanitize=sys.argv[2])
        print("built", sys.argv[3], "with host -fsanitize=" + sys.argv[2])
    else:
        build_all()
        print("built", LIB, HOSTCHECK)
