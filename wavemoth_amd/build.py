"""Compile libwavemoth_hip.so (gfx950) in-tree with hipcc.

Usage: python -m wavemoth_amd.build   (or wavemoth_amd.build.build()).
The library has a plain C ABI (include/wavemoth.h) and no torch dependency.
Each source is compiled to its own object in parallel (no device symbol
crosses a translation unit), then linked into the shared library.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBNAME = "libwavemoth_hip.so"
SOURCES = ["plan.hip", "batch.hip", "legendre.hip", "mixed.hip", "ring.hip", "ring16.hip",
           "ring_long.hip",
           "adjoint.hip", "strip.hip", "resource.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 ops are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def lib_path():
    return os.path.join(HERE, LIBNAME)


def _filter(out):
    # the host pass does not know the packed-fp32 feature; that warning is harmless
    return [ln for ln in out.splitlines()
            if "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]


def _run(cmd):
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    return res.returncode, _filter(res.stdout)


def build(verbose=False, extra_flags=(), jobs=None):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    out = lib_path()
    tmp = out + ".tmp"
    inc = ["-I", os.path.join(ROOT, "include")]
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with tempfile.TemporaryDirectory(prefix="wm_build_") as td:
        objs, cmds = [], []
        for s in srcs:
            o = os.path.join(td, os.path.basename(s) + ".o")
            objs.append(o)
            lang = ["-x", "hip"] if s.endswith(".hip") else []
            cmds.append([HIPCC] + FLAGS + list(extra_flags) + inc + ["-c"] + lang + [s, "-o", o])
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(jobs) as ex:
            results = list(ex.map(_run, cmds))
        failed = False
        for s, (rc, lines) in zip(srcs, results):
            if rc != 0:
                failed = True
                sys.stderr.write("== %s\n%s\n" % (os.path.basename(s), "\n".join(lines)))
            elif verbose and lines:
                print("\n".join(lines))
        if failed:
            raise RuntimeError("hipcc failed building %s" % LIBNAME)
        rc, lines = _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
        if rc != 0:
            sys.stderr.write("\n".join(lines) + "\n")
            raise RuntimeError("hipcc failed linking %s" % LIBNAME)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
