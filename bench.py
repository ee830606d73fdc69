"""Benchmark: P2 NeoHookean energy + gradient + Hessian assembly on one MI355X.

Workload: 40^3 box grid, Kuhn 6-tet split, P2 (384,000 tets, 531,441 nodes,
1,594,323 dofs, 134,475,849 Hessian non-zeros), 4-point quadrature, per-element
E ~ logNormal(ln 1e6, 0.5), nu ~ U[0.3, 0.45], smooth random displacement;
mixed precision (FP32 element kernels, FP64 energy/gradient sums, FP32 Hessian
values).  One step = one assembly of energy, gradient and Hessian values at
the current state, inputs resident in HBM.

Also reported on the same workload: the matrix-free Hv time; the relative
error of the mixed-precision energy, gradient, Hessian and Hv against the FP64
GPU mode; the adjoint-gradient time (static adjoint CG solve with Hv plus the
material VJP) in mixed precision against the FP64 plain form; and a
Newton-CG solve (bottom face clamped, top face moved 0.1 z, matrix-free CG).

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
METRIC = "elem\u00b7qp/s for NeoHookean E/grad/Hess + Hv; adjoint-grad time vs FP64 GPU"


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


def extras(args, m, lam, mu, x_host, a, x, g, vals, y, eqp):
    """Mixed-vs-FP64 errors, adjoint-gradient times, Newton-CG solve."""
    import ctypes as C

    import numpy as np
    import torch

    from polyfem_amd import NLAssembler, NewtonSolver, NLProblem, _ffi
    from polyfem_amd.mesh import face_nodes

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    res = {}
    # FP64 GPU reference on the same state (plain form)
    a64 = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64", path="plain")
    a64.set_stream(stream.cuda_stream)
    g64 = torch.empty_like(g)
    e64 = torch.empty(1, dtype=torch.float64, device=dev)
    v64 = torch.empty(a64.nnz, dtype=torch.float64, device=dev)
    a64.energy_gradient_hessian_dev(x.data_ptr(), e64.data_ptr(), g64.data_ptr(), v64.data_ptr())
    e32 = torch.empty(1, dtype=torch.float64, device=dev)
    a.energy_gradient_hessian_dev(x.data_ptr(), e32.data_ptr(), g.data_ptr(), vals.data_ptr())
    v = torch.from_numpy(np.random.default_rng(5).normal(size=a.n_dofs)).to(dev)
    y64 = torch.empty_like(v)
    a64.hvp_dev(x.data_ptr(), v.data_ptr(), y64.data_ptr())
    a.hvp_dev(x.data_ptr(), v.data_ptr(), y.data_ptr())
    torch.cuda.synchronize()

    def rel(u, w):
        return float((u.double() - w).abs().max() / w.abs().max())

    res["mixed_vs_fp64_rel_error"] = {"energy": rel(e32, e64), "gradient": rel(g, g64), "hessian": rel(vals, v64),
                                      "hvp": rel(y, y64)}
    # Dirichlet setup of the Newton and adjoint solves
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    vals_bc = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, 0.1], top.size)])
    prob = NLProblem(m.n_dofs, fixed, vals_bc)
    # start of the solves: the seeded displacement plus the linear ramp 0.1 z e_z,
    # which meets both faces' Dirichlet values without folding the elements
    # next to them
    x_bc = x_host.copy()
    x_bc[2::3] += 0.1 * m.nodes[:, 2]
    x_bc[prob.fixed] = prob.fixed_values
    # adjoint solve at the uniform 10 % stretch between the faces (a state with
    # a positive definite Hessian, so CG runs its full iteration count)
    x_adj = np.zeros(m.n_dofs)
    x_adj[2::3] = 0.1 * m.nodes[:, 2]
    xd = torch.from_numpy(x_adj).to(dev)
    mask = torch.from_numpy(prob.mask).to(dev)
    rhs = torch.from_numpy(np.random.default_rng(7).normal(size=m.n_dofs)).to(dev)
    p = torch.empty_like(rhs)
    out2 = torch.empty(2 * m.n_elements, dtype=torch.float64, device=dev)

    def adjoint_time(asm):
        it, rr = C.c_int32(0), C.c_double(0)

        def once():
            _ffi.check(_ffi.lib().pfgpu_cg_solve_dev(asm._h, C.c_void_p(xd.data_ptr()), C.c_void_p(mask.data_ptr()),
                                                     C.c_void_p(rhs.data_ptr()), C.c_void_p(p.data_ptr()),
                                                     args.cg_iterations, 0.0, 1, C.byref(it), C.byref(rr)))
            asm.material_vjp_dev(xd.data_ptr(), p.data_ptr(), out2.data_ptr())
        once()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        once()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3, it.value

    t_mixed, it_m = adjoint_time(a)
    t_fp64, it_d = adjoint_time(a64)
    res["adjoint"] = {"what": f"static adjoint: {args.cg_iterations} CG iterations on H(x)_ff p = r (matrix-free Hv) "
                              "+ material VJP, at 10 % stretch between the clamped faces",
                      "mixed_ms": t_mixed, "fp64_plain_ms": t_fp64,
                      "speedup_vs_fp64_plain": t_fp64 / t_mixed, "cg_iterations": it_m}
    del a64
    # Newton-CG from the seeded displacement with the Dirichlet values applied
    if args.newton_iterations > 0:
        solver = NewtonSolver(max_iterations=args.newton_iterations, grad_norm=0.0, cg_max_iterations=200,
                              cg_tolerance=1e-4)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, rep = solver.minimize(a, x_bc, prob)
        t = time.perf_counter() - t0
        res["newton_cg"] = {"iterations": rep.iterations, "cg_iterations": rep.cg_iterations, "seconds": t,
                            "ms_per_newton_iteration": t * 1e3 / max(rep.iterations, 1),
                            "final_grad_norm": rep.grad_norm, "energy": rep.energy,
                            "line_search_halvings": rep.line_search_halvings}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="auto", choices=["auto", "fused", "plain"],
                    help="auto: the launch-shape autotuner picks path and workgroup sizes (untimed setup)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--grid", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the error, adjoint and Newton measurements")
    ap.add_argument("--newton-iterations", type=int, default=10)
    ap.add_argument("--cg-iterations", type=int, default=100)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 1 GiB device to device five times (the reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the path shards by "
              "element slabs with halo exchange, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")  # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 4, "copies": 5}))
        return

    from polyfem_amd import NLAssembler
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

    n = args.grid
    m = kuhn_box(n, n, n, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x_host = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=args.precision,
                    path="plain" if args.path == "auto" else args.path)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    a.set_stream(stream.cuda_stream)
    vdt = torch.float32 if args.precision == "mixed" else torch.float64
    x = torch.from_numpy(x_host).to(dev)
    g = torch.empty(a.n_dofs, dtype=torch.float64, device=dev)
    e = torch.empty(1, dtype=torch.float64, device=dev)
    vals = torch.empty(a.nnz, dtype=vdt, device=dev)
    torch.cuda.synchronize()

    tuning = a.autotune_dev(x.data_ptr()) if args.path == "auto" else None

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
    tfile = os.path.join(ROOT, "profiles", f"traffic_egh_p2_{n}_{args.precision}_{a.path}.json")
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
        "metric": METRIC,
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
                               f"{args.precision} precision, {a.path} path",
                   "tuning": tuning,
                   "n_elements": m.n_elements, "n_dofs": a.n_dofs, "nnz": a.nnz, "quadrature_points": a.n_quad},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg, "kernel_ms": kernel_ms},
        "hvp": {"ms": hv_ms, "elem_qp_per_s": eqp / (hv_ms / 1000.0)},
    }
    if not args.no_extras:
        out.update(extras(args, m, lam, mu, x_host, a, x, g, vals, y, eqp))
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline_sample(threads)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
