"""Build ``libtmx.so`` (the gfx950 HIP library behind the C ABI in include/tmx.h).

Run ``python -m topomax_amd.build`` (or call :func:`build`).  The library is built in-tree
(``topomax_amd/libtmx.so``) so it travels with the source tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtmx.so")
SOURCES = ["tmx_kernels.hip", "tmx_band.hip", "tmx_p2.hip", "tmx_dct.hip", "tmx_solver.hip",
           "tmx_api.hip", "tmx_batch.hip", "tmx_noalloc.hip"]
ARCH = os.environ.get("TMX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-pthread",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build libtmx.so")
    return path


def build(verbose: bool = False, jobs: int = 3) -> str:
    """Compile every HIP source for gfx950 and link ``libtmx.so``; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    while procs:
        _wait(procs.pop(0))
    tmp = LIB + ".tmp"
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    if verbose:
        print(" ".join(link), flush=True)
    _run(link)
    os.replace(tmp, LIB)
    return LIB


def _filter(out: str) -> str:
    keep = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    return "\n".join(keep)


def _wait(item):
    src, p = item
    out, _ = p.communicate()
    text = _filter(out.decode(errors="replace"))
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{text}")
    if text.strip():
        print(text, file=sys.stderr)


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    text = _filter(p.stdout.decode(errors="replace"))
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{text}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
