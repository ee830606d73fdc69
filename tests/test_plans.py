"""CPU checks of the host-built index tables under AddressSanitizer and
UndefinedBehaviorSanitizer (no GPU).

Builds tests/cpp/plan_check.cpp with pattern.cpp, plans.cpp and
fused_tables.cpp, instrumented, and runs it on the full-size meshes of the
benchmarked and transient workloads: every index the row-unit gather derives
from its plan must stay inside the buffers it addresses, the stage positions
of a chunk must be a bijection, and each block's contributions must be in
element order.
"""
import os
import subprocess

import numpy as np
import pytest

from meshes import star_mesh
from polyfem_amd.mesh import kuhn_box

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "build", "plan_check")


@pytest.fixture(scope="module")
def exe():
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    srcs = [os.path.join(ROOT, "tests", "cpp", "plan_check.cpp"),
            os.path.join(ROOT, "polyfem_amd", "csrc", "pattern.cpp"),
            os.path.join(ROOT, "polyfem_amd", "csrc", "plans.cpp"),
            os.path.join(ROOT, "polyfem_amd", "csrc", "fused_tables.cpp")]
    # host code only, compiled by the host compiler with the sanitizers
    cmd = ["g++", "-O1", "-g", "-std=c++20", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT}/include", "-I/opt/rocm/include",
           *srcs, "-o", EXE, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return EXE


def write_mesh(path, m):
    with open(path, "wb") as f:
        np.array([m.order, m.n_elements, m.n_nodes, m.vertices.shape[0]], dtype=np.int32).tofile(f)
        m.elem_nodes.astype(np.int32).tofile(f)
        m.elem_vertices.astype(np.int32).tofile(f)
        m.vertices.astype(np.float64).tofile(f)


@pytest.mark.parametrize("grid,order", [((40, 40, 40), 2), ((20, 20, 20), 1), ((3, 2, 2), 2)])
def test_plans_under_sanitizers(exe, tmp_path, grid, order):
    m = kuhn_box(*grid, order=order)
    p = tmp_path / "mesh.bin"
    write_mesh(p, m)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(p)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "OK (0 failures)" in r.stdout


@pytest.mark.parametrize("order", [1, 2])
def test_plans_high_valence(exe, tmp_path, order):
    """A centre vertex in 60 tets: blocks of up to 60 contributions, split
    into pieces."""
    m = star_mesh(30, order=order)
    p = tmp_path / "mesh.bin"
    write_mesh(p, m)
    r = subprocess.run([exe, str(p)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "OK (0 failures)" in r.stdout
    assert "gs 64 lanes 64: " in r.stdout and "gs 64: a row has too many units" not in r.stdout
