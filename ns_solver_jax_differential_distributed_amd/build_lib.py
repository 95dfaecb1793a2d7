"""Build libnsk.so (gfx950) in-tree with hipcc.

The row/column kernels are compiled once per grid size (-DNSK_N=<N>) into
separate objects, in parallel, and linked with the C ABI object into
``libnsk.so`` next to this file.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnsk.so")
SIZES = (8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096)
ARCH = os.environ.get("NSK_OFFLOAD_ARCH", "gfx950")


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (set HIPCC or install ROCm)")


def _flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
        # packed FP32 (v_pk_*_f32) stays off for all device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
    ]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = _hipcc()
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(len(SIZES) + 1, max(1, (os.cpu_count() or 4)), 16)
    deps = [os.path.join(CSRC, f) for f in ("nsk_fft.h", "nsk_internal.h")]
    deps.append(os.path.join(HERE, "..", "include", "nsk.h"))
    cmds = []
    objs = []
    for n in SIZES:
        obj = os.path.join(BUILD, f"nsk_rowcol_{n}.o")
        objs.append(obj)
        cmds.append([hipcc, *_flags(), f"-DNSK_N={n}", "-c",
                     os.path.join(CSRC, "nsk_rowcol.hip"), "-o", obj])
    api = os.path.join(BUILD, "nsk_api.o")
    objs.append(api)
    cmds.append([hipcc, *_flags(), "-c", os.path.join(CSRC, "nsk_api.hip"), "-o", api])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
