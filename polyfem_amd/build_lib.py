"""Compile libpfgpu.so (HIP kernels + C-ABI) for gfx950, in-tree.

Usage: ``python -m polyfem_amd.build_lib`` or ``polyfem_amd.build_lib.build()``.
Every source is compiled with hipcc; objects go to ``polyfem_amd/build/`` and
the shared library to ``polyfem_amd/libpfgpu.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpfgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PFGPU_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++20", f"--offload-arch={ARCH}", "-fPIC",
    f"-I{os.path.join(ROOT, 'include')}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _compile(src: str, obj: str) -> None:
    cmd = [HIPCC, *FLAGS, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")


def build(verbose: bool = True) -> str:
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")) + glob.glob(os.path.join(HERE, "csrc", "*.cpp")))
    out_dir = os.path.join(HERE, "build")
    os.makedirs(out_dir, exist_ok=True)
    objs = [os.path.join(out_dir, os.path.splitext(os.path.basename(s))[0] + ".o") for s in srcs]
    workers = min(8, os.cpu_count() or 1)
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda so: _compile(*so), zip(srcs, objs)))
    tmp = LIB + ".tmp"
    r = subprocess.run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libpfgpu.so failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
