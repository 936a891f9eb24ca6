"""Builds puck_amd/libpuck.so (gfx950 device code + host engine) with hipcc, in-tree.

Usage: python -m puck_amd.build  (or puck_amd.build.build_library() from Python).
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpuck.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PUCK_OFFLOAD_ARCH", "gfx950")

HOST_SOURCES = ["grammar.cpp", "grz.cpp", "synth.cpp", "server.cpp", "batch.cpp", "capi.cpp"]
DEVICE_SOURCES = ["engine.hip"]

COMMON_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]
# Packed FP32 instructions are kept out of every device build; contraction is off so that rule
# arithmetic keeps the reference's association and rounding.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


LLVM_BIN = os.environ.get("PUCK_LLVM_BIN", "/opt/rocm/llvm/bin")
PACKED_FP32 = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # The x86-64 host pass does not know the packed-fp32-ops feature and says so; its messages name
    # no target, so the device code itself is checked instead (check_no_packed_fp32).
    out = "\n".join(l for l in res.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def check_no_packed_fp32(obj, tmp):
    """Fails the build when the gfx950 code object inside `obj` holds a packed FP32 add, mul or fma
    (the device builds disable the feature; this proves it took effect)."""
    fb = os.path.join(tmp, os.path.basename(obj) + ".fatbin")
    co = os.path.join(tmp, os.path.basename(obj) + ".co")
    _run([os.path.join(LLVM_BIN, "llvm-objcopy"), "--dump-section", f".hip_fatbin={fb}", obj])
    _run([os.path.join(LLVM_BIN, "clang-offload-bundler"), "--unbundle", "--type=o", f"--input={fb}",
          f"--targets=hipv4-amdgcn-amd-amdhsa--{ARCH}", f"--output={co}"])
    dis = subprocess.run([os.path.join(LLVM_BIN, "llvm-objdump"), "-d", co], stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True, check=True).stdout
    hits = [l.strip() for l in dis.splitlines() if PACKED_FP32.search(l)]
    if hits:
        raise RuntimeError(f"{obj}: {len(hits)} packed FP32 instructions in the {ARCH} code, e.g. {hits[0]}")
    if "s_endpgm" not in dis:
        raise RuntimeError(f"{obj}: no {ARCH} kernel code found to check")


def build_library(verbose=False):
    """Compiles every source and links libpuck.so next to this file. Always rebuilds."""
    tmp = tempfile.mkdtemp(prefix="puck_build_")
    try:
        objs = []
        for src in HOST_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            _run([HIPCC, "-x", "c++", *COMMON_FLAGS, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                  "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        for src in DEVICE_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            out = _run([HIPCC, *COMMON_FLAGS, *DEVICE_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
            if verbose and out.strip():
                print(out)
            check_no_packed_fp32(obj, tmp)
            objs.append(obj)
        tmp_lib = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp_lib])
        os.replace(tmp_lib, LIB)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
