"""Build libpir_hip.so (HIP kernels + C++ host runtime + C ABI) for gfx950.

Usage: python -m piranha_amd.build [-j N]
Objects go to piranha_amd/_build/, the library to piranha_amd/libpir_hip.so.
"""
import argparse
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpir_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "ring_gemm.hip",
    "fused_rss.hip",
    "elementwise.hip",
    "tpc_fpc_compare.hip",
    "noalloc.hip",
    "peaks.hip",
    "prg.hip",
    "runtime.cpp",
    "protocol.cpp",
    "tpc_fpc.cpp",
    "layers.cpp",
    "model_io.cpp",
    "capi.cpp",
]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # packed FP32 instructions are kept off for all HIP code on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{msgs}")
    return out, msgs


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out, msgs in ex.map(_compile, SOURCES):
            objs.append(out)
            if verbose and msgs.strip():
                print(msgs, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.j, a.v))
