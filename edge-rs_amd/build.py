"""Builds libnebula_hip.so (HIP kernels for gfx950 + the C++ host driver behind
the C ABI in include/nebula_hip.h) into edge-rs_amd/nebula_hip/.

    python edge-rs_amd/build.py            # full rebuild
    python edge-rs_amd/build.py --incremental
    python edge-rs_amd/build.py --sanitize  # host code under ASan + UBSan, into
                                            # edge-rs_amd/_build_asan/libnebula_hip.so

The sanitizer build instruments host code only (each -fsanitize= follows
-Xarch_host; device code is built as usual).  Load it with NH_LIBRARY pointing
at it and the clang ASan runtime preloaded (tests/run_sanitized.sh).

Every HIP source is compiled with -packed-fp32-ops disabled (no v_pk_*_f32).
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "nebula_hip", "libnebula_hip.so")

SOURCES = [
    "kernels.hip",
    "stage_kernels.hip",
    "persistent.hip",
    "mlp.hip",
    "noalloc.hip",
    "resident.cpp",
    "capi_resident.cpp",
    "driver.cpp",
    "sc_test.cpp",
    "zero_tables.cpp",
    "shrink.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NH_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function",
          "-Wno-unknown-pragmas", "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=all"]


def newest_dep_mtime():
    m = 0.0
    for f in os.listdir(CSRC):
        if f.endswith((".hpp", ".h")):
            m = max(m, os.path.getmtime(os.path.join(CSRC, f)))
    m = max(m, os.path.getmtime(os.path.join(ROOT, "include", "nebula_hip.h")))
    return m


def compile_one(src, incremental, build=BUILD, extra=()):
    path = os.path.join(CSRC, src)
    obj = os.path.join(build, src + ".o")
    if incremental and os.path.exists(obj):
        if os.path.getmtime(obj) >= max(os.path.getmtime(path), newest_dep_mtime()):
            return obj, 0.0
    cmd = [HIPCC] + COMMON + DEVICE + list(extra) + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + COMMON + DEVICE + list(extra) + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {src}")
    return obj, 1.0


def build(incremental=False, jobs=8, sanitize=False):
    build_dir = os.path.join(HERE, "_build_asan") if sanitize else BUILD
    out = os.path.join(build_dir, "libnebula_hip.so") if sanitize else OUT
    extra = SANITIZE if sanitize else []
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, incremental, build_dir, extra)[0], srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    if sanitize:
        cmd += ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                "-shared-libasan"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args()
    print(build(a.incremental, a.j, a.sanitize))
