"""Compile libwavemoth_hip.so (gfx950) in-tree with hipcc.

Usage: python -m wavemoth_amd.build   (or wavemoth_amd.build.build()).
The library has a plain C ABI (include/wavemoth.h) and no torch dependency.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBNAME = "libwavemoth_hip.so"
SOURCES = ["plan.hip", "batch.hip", "legendre.hip", "mixed.hip", "ring.hip", "ring16.hip",
           "adjoint.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 ops are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path():
    return os.path.join(HERE, LIBNAME)


def build(verbose=False, extra_flags=()):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    out = lib_path()
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + list(extra_flags) + ["-I", os.path.join(ROOT, "include"),
                                                 "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass does not know the packed-fp32 feature; that warning is harmless
    lines = [ln for ln in res.stdout.splitlines()
             if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if res.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("hipcc failed building %s" % LIBNAME)
    if verbose and lines:
        print("\n".join(lines))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
