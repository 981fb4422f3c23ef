"""Builds libhpsint_amd.so (HIP, gfx950) in-tree.

    python -m hpsint_amd.build          # or hpsint_amd.build.build()

Every translation unit under hpsint_amd/csrc/*.hip is compiled with hipcc
for gfx950 in parallel, then linked into hpsint_amd/libhpsint_amd.so; the
C++ adapter (hpsint_amd/adapter/) is compiled against the C ABI into
hpsint_amd/libhpsint_adapter_check.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libhpsint_amd.so")
ADAPTER_LIB = os.path.join(HERE, "libhpsint_adapter_check.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed-FP32 instructions stay off (scalar/float2 code only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-8000:]))
    return obj


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(len(srcs), os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC",
           "-Wl,-soname,libhpsint_amd.so", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-8000:])
    os.replace(tmp, LIB)
    build_adapter_check()
    if verbose:
        print("built", LIB, "and", ADAPTER_LIB)
    return LIB


def build_adapter_check():
    """the C++ adapter of adapter/sintering_operator_mi355x.h compiled
    against include/hpsint_amd.h and linked to the library, with a C entry
    point per adapter call (adapter/adapter_check.cpp) for the in-process
    test"""
    tmp = ADAPTER_LIB + ".tmp"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-shared", "-fPIC", "-o", tmp,
           os.path.join(HERE, "adapter", "adapter_check.cpp"), "-L" + HERE, "-l:libhpsint_amd.so",
           "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("adapter build failed:\n" + r.stderr[-8000:])
    os.replace(tmp, ADAPTER_LIB)


if __name__ == "__main__":
    build(jobs=int(sys.argv[1]) if len(sys.argv) > 1 else None)
