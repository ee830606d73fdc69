"""Benchmark: P2 NeoHookean energy + gradient + Hessian assembly on one MI355X.

Workload: 40^3 box grid, Kuhn 6-tet split, P2 (384,000 tets, 531,441 nodes,
1,594,323 dofs, 134,475,849 Hessian non-zeros), 4-point quadrature, per-element
E ~ logNormal(ln 1e6, 0.5), nu ~ U[0.3, 0.45], smooth random displacement;
mixed precision (FP32 element kernels, FP64 energy/gradient sums, FP32 Hessian
values).  One step = one assembly of energy, gradient and Hessian values at
the current state, inputs resident in HBM.

Prints one JSON line (rank 0).  ``--gpus`` above 1 is not supported: this
build runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(n_dofs, n_elements, n_local, n_vertices, nnz, value_bytes, lame_bytes):
    """Compulsory traffic of one E+G+H assembly, each datum touched once:
    reads x (8 B/dof), connectivity (4 B per element node), vertex
    coordinates (24 B/vertex), per-element Lame pair; writes the gradient
    (8 B/dof) and the Hessian values."""
    reads = n_dofs * 8 + n_elements * n_local * 4 + n_vertices * 24 + n_elements * 2 * lame_bytes
    writes = n_dofs * 8 + nnz * value_bytes
    return reads + writes


def cpu_baseline_sample(threads: int):
    """The reference algorithm restated in C (threaded, per-thread storage,
    slot-map Hessian) timed on a 20^3 x 6 P2 sample of the same workload."""
    import numpy as np

    from oracle import pyoracle as po
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement
    po.build()
    m = kuhn_box(20, 20, 20, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    o = po.Oracle(m, lam, mu)
    b = po.Baseline(o, threads)
    t0 = time.perf_counter()
    nnz = b.hessian(x)  # first call: triplets -> CSC + slot map
    first = time.perf_counter() - t0
    vals = np.zeros(nnz)
    g = np.zeros(m.n_dofs)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        b.energy(x)
        po.lib().orc_baseline_gradient(b.b, x, g)
        b.hessian(x, vals)
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    eqp = m.n_elements * 4
    return {"value": eqp / t, "unit": "elem*qp/s", "cores": b.threads, "kind": "port",
            "sample": f"20^3x6 P2 NeoHookean ({m.n_elements} tets, {eqp} elem*qp), FP64 energy+gradient+Hessian, "
                      f"median of 3 steady-state calls; first call (pattern + slot map) {first:.2f} s",
            "seconds_per_step": t}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="fused", choices=["fused", "plain"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--grid", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "element slabs with halo exchange, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from polyfem_amd import NLAssembler
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

    n = args.grid
    m = kuhn_box(n, n, n, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x_host = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=args.precision, path=args.path)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    a.set_stream(stream.cuda_stream)
    vdt = torch.float32 if args.precision == "mixed" else torch.float64
    x = torch.from_numpy(x_host).to(dev)
    g = torch.empty(a.n_dofs, dtype=torch.float64, device=dev)
    e = torch.empty(1, dtype=torch.float64, device=dev)
    vals = torch.empty(a.nnz, dtype=vdt, device=dev)
    torch.cuda.synchronize()

    def step():
        a.energy_gradient_hessian_dev(x.data_ptr(), e.data_ptr(), g.data_ptr(), vals.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = statistics.mean(s.elapsed_time(t) for s, t in zip(starts, ends))
    ms_per_step = wall * 1000.0 / args.steps
    eqp = m.n_elements * a.n_quad
    value = eqp / (ms_per_step / 1000.0)

    vb = 4 if args.precision == "mixed" else 8
    alg = algorithmic_bytes(a.n_dofs, m.n_elements, 10, m.vertices.shape[0], a.nnz, vb, vb)
    achieved = alg / (kernel_ms / 1000.0) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_egh_p2_{n}_{args.precision}_{args.path}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")

    # Hv throughput on the same state (matrix-free)
    v = torch.ones(a.n_dofs, dtype=torch.float64, device=dev)
    y = torch.empty_like(v)
    for _ in range(2):
        a.hvp_dev(x.data_ptr(), v.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()
    hs, he = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    hs.record(stream)
    for _ in range(args.steps):
        a.hvp_dev(x.data_ptr(), v.data_ptr(), y.data_ptr())
    he.record(stream)
    torch.cuda.synchronize()
    hv_ms = hs.elapsed_time(he) / args.steps

    out = {
        "metric": "elem*qp/s for NeoHookean E/grad/Hess + Hv",
        "value": value,
        "unit": "elem*qp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "mixed" else "f64",
        "data": "synthetic (Kuhn box mesh, seeded displacement and Lame parameters)",
        "config": {"workload": f"P2 NeoHookean E+G+H assembly, {n}^3x6 tets ({m.n_elements}), 4 qp, "
                               f"{args.precision} precision, {args.path} path",
                   "n_elements": m.n_elements, "n_dofs": a.n_dofs, "nnz": a.nnz, "quadrature_points": a.n_quad},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg, "kernel_ms": kernel_ms},
        "hvp": {"ms": hv_ms, "elem_qp_per_s": eqp / (hv_ms / 1000.0)},
    }
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline_sample(threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
