"""CPU check of the fused kernel's patch tables and index arithmetic.

Builds tests/cpp/fused_check.cpp (host code only) and runs it on small
meshes: every index the kernel derives from the tables is checked against the
buffers it addresses, every (row, element) pair must be scheduled exactly
once, and a serial FP64 replay of the kernel's phases must reproduce a direct
element-by-element assembly to 1e-12.
"""
import os
import subprocess

import numpy as np
import pytest

from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "fused_check")


@pytest.fixture(scope="module")
def exe():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    srcs = [os.path.join(ROOT, "tests", "cpp", "fused_check.cpp"),
            os.path.join(ROOT, "polyfem_amd", "csrc", "pattern.cpp"),
            os.path.join(ROOT, "polyfem_amd", "csrc", "fused_tables.cpp")]
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++20", f"-I{ROOT}/include", *srcs, "-o", EXE, "-lpthread"]
    subprocess.run(cmd, check=True, capture_output=True)
    return EXE


def write_mesh(path, m, x, lam, mu):
    with open(path, "wb") as f:
        np.array([m.order, m.n_elements, m.n_nodes, m.vertices.shape[0]], dtype=np.int32).tofile(f)
        m.elem_nodes.astype(np.int32).tofile(f)
        m.elem_vertices.astype(np.int32).tofile(f)
        m.vertices.astype(np.float64).tofile(f)
        x.astype(np.float64).tofile(f)
        lam.astype(np.float64).tofile(f)
        mu.astype(np.float64).tofile(f)


@pytest.mark.parametrize("order,grid,groups,chunk,icap,cap", [
    (2, (3, 3, 4), 28, 32, 320, 4000),       # several patches
    (2, (4, 3, 3), 28, 8, 40, 12000),        # several chunks per patch, item-capped chunks
    (2, (10, 10, 10), 28, 32, 320, 21000),   # the FP32 P2 launch shape
    (2, (6, 6, 6), 30, 16, 160, 10400),      # the FP64 P2 launch shape
    (1, (6, 5, 4), 70, 128, 512, 3000),
    (1, (12, 12, 12), 70, 128, 512, 27000),  # the FP32 P1 launch shape
])
def test_fused_tables(exe, tmp_path, order, grid, groups, chunk, icap, cap):
    m = kuhn_box(*grid, order=order)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    path = str(tmp_path / "mesh.bin")
    write_mesh(path, m, x, lam, mu)
    r = subprocess.run([exe, path, str(groups), str(chunk), str(icap), str(cap), "256"], capture_output=True, text=True)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout
