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


def host_ram_bytes() -> int:
    """Available host memory (MemAvailable), or 0 when unknown."""
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    return int(line.split()[1]) * 1024
    except OSError:
        pass
    return 0


def cpu_baseline_sample(cores: int):
    """The reference algorithm restated in C (static std::thread chunking of
    utils/par_for.cpp:10-31, per-thread gradient vectors and full-length
    Hessian value arrays with a serial reduction, MatrixCache slot map),
    timed on a 20^3 x 6 P2 sample of the same workload: one warm-up, the
    median of 5 steady-state E+G+H calls, with every host core (capped so
    that the per-thread value arrays fit the host memory) and with one
    thread.  The first Hessian call (triplets -> CSC + slot map) is reported
    apart.  Bounded to a sample: the first call alone takes ~100 s at the
    full 40^3 x 6 size."""
    import numpy as np

    from oracle import pyoracle as po
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement
    po.build()
    m = kuhn_box(20, 20, 20, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    o = po.Oracle(m, lam, mu)
    eqp = m.n_elements * 4

    def timed(threads):
        b = po.Baseline(o, threads)
        t0 = time.perf_counter()
        nnz = b.hessian(x)  # first call: triplets -> CSC + slot map
        first = time.perf_counter() - t0
        vals = np.zeros(nnz)
        g = np.zeros(m.n_dofs)

        def step():
            t0 = time.perf_counter()
            b.energy(x)
            po.lib().orc_baseline_gradient(b.b, x, g)
            b.hessian(x, vals)
            return time.perf_counter() - t0
        step()  # warm-up
        t = statistics.median(step() for _ in range(5))
        used = b.threads
        del b
        return t, first, used, nnz

    # RAM cap: the reference's per-thread value arrays cost nnz x 8 B per thread
    t1, first1, _, nnz = timed(1)
    ram = host_ram_bytes()
    cap = max(1, int(0.5 * ram // (nnz * 8))) if ram else cores
    threads = max(1, min(cores, cap))
    print(f"cpu baseline: {cores} host cores, RAM cap {cap} threads ({ram / 2**30:.0f} GiB available, "
          f"{nnz * 8 / 2**20:.0f} MiB of values per thread) -> {threads} threads", file=sys.stderr, flush=True)
    tn, firstn, used, _ = timed(threads)
    return {"value": eqp / tn, "unit": "elem*qp/s", "cores": used, "kind": "port",
            "sample": f"20^3x6 P2 NeoHookean ({m.n_elements} tets, {eqp} elem*qp), FP64 energy+gradient+Hessian, "
                      f"reference-algorithm restatement (oracle/oracle.c baseline), median of 5 after a warm-up; "
                      f"first call (pattern + slot map) {firstn:.1f} s",
            "seconds_per_step": tn,
            "one_thread": {"value": eqp / t1, "seconds_per_step": t1, "first_call_s": first1},
            "threads_cap_by_ram": cap}


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
    # PSD projection of the element Hessians (host entry point, so both timings
    # include the same staging and copy of the values)
    xh = x.cpu().numpy()
    tp = []
    for psd in (False, True, False, True):
        t0 = time.perf_counter()
        a.hessian_values(xh, project_to_psd=psd)
        tp.append(time.perf_counter() - t0)
    res["project_to_psd"] = {"extra_ms": 1e3 * (tp[3] - tp[2]),
                             "what": "one E+G+H-sized Hessian with every element matrix projected onto the PSD cone, "
                                     "minus the same call without projection (host API)"}
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
    # Newton-CG from the seeded displacement with the Dirichlet values applied,
    # in mixed precision and, for the trajectory comparison, in FP64
    if args.newton_iterations > 0:
        # regularized Newton-CG (H + delta D on non-positive curvature) to a
        # relative gradient tolerance: 1e-5 in mixed precision (the FP32
        # element gradients carry ~1e-6 relative noise), 1e-8 in FP64
        def newton(asm, rtol):
            solver = NewtonSolver(max_iterations=args.newton_iterations, grad_norm=0.0, relative_grad_norm=rtol,
                                  cg_max_iterations=500, cg_tolerance=1e-3, regularization=(1e-2, 1e6))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xs, rep = solver.minimize(asm, x_bc, prob)
            t = time.perf_counter() - t0
            return xs, {"converged": rep.converged, "relative_grad_norm_tolerance": rtol,
                        "iterations": rep.iterations, "cg_iterations": rep.cg_iterations, "seconds": t,
                        "ms_per_newton_iteration": t * 1e3 / max(rep.iterations, 1),
                        "initial_grad_norm": rep.initial_grad_norm, "final_grad_norm": rep.grad_norm,
                        "grad_norm_ratio": rep.grad_norm / rep.initial_grad_norm, "energy": rep.energy,
                        "line_search_halvings": rep.line_search_halvings, "gradient_steps": rep.gradient_steps,
                        "regularized_solves": rep.regularized_solves,
                        "cg_stops": {"tolerance": rep.cg_stops_tolerance, "curvature": rep.cg_stops_curvature,
                                     "cap": rep.cg_stops_cap}}

        xm, res["newton_cg"] = newton(a, 1e-5)
        a64n = NLAssembler("NeoHookean", m, lame=(lam, mu), precision="fp64")
        xd64, r64 = newton(a64n, 1e-8)
        del a64n
        u0 = x_bc[prob.free]
        res["newton_cg"]["fp64"] = r64
        res["newton_cg"]["mixed_vs_fp64_solution_rel"] = float(
            np.abs((xm - xd64)[prob.free]).max() / max(np.abs((xd64[prob.free] - u0)).max(), 1e-300))
    # the 256-problem heterogeneous batch (the suite's batch workload)
    if args.batch_problems > 0:
        res["batch"] = batch_run(args.batch_problems, args.batch_newton, args.batch_rtol, args.batch_cg,
                                 lambda msg: print(msg, file=sys.stderr, flush=True))
    return res


_ARGS = None


def _clock():
    import torch
    torch.cuda.synchronize()
    return time.perf_counter()


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def suite(args):
    """Further workloads at full size, one JSON line each (timings in seconds
    unless named otherwise)."""
    global _ARGS
    import torch
    torch.cuda.is_available()
    _ARGS = args
    parts = set(args.suite_parts.split(","))
    if "step" in parts:
        suite_step()
    if "batch" in parts:
        suite_batch()
    if "cantilever" in parts:
        suite_cantilever()
    if "transient" in parts:
        suite_transient()


def suite_step():
    import numpy as np
    from polyfem_amd import NewtonSolver, NLAssembler, NLProblem
    from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, smooth_displacement
    clock = _clock

    # ---- P1 20^3 x 6, FP64, E = 1e6, nu = 0.4: E+G+H, then one Newton step on the
    # free dofs (bottom clamped, top moved by +0.1 z) with alpha = 1, and the same
    # step with the backtracking line search
    m = kuhn_box(20, 20, 20, order=1)
    a = NLAssembler("NeoHookean", m, young=1e6, poisson=0.4, precision="fp64")
    x = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    bot, top = face_nodes(m, 2, 0.0), face_nodes(m, 2, 1.0)
    fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
    prob = NLProblem(m.n_dofs, fixed, np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, 0.1], top.size)]))
    for _ in range(3):
        a.assemble_energy_gradient_hessian(x)
    ts = []
    for _ in range(20):
        t0 = clock()
        a.assemble_energy_gradient_hessian(x)
        ts.append(clock() - t0)
    x0 = x.copy()
    x0[prob.fixed] = prob.fixed_values
    g0 = float(np.linalg.norm(a.assemble_gradient(x0)[prob.free]))
    steps = {}
    for name, ls in (("alpha_1", "None"), ("backtracking", "Backtracking")):
        t0 = clock()
        x1, rep = NewtonSolver(max_iterations=1, grad_norm=0.0, cg_max_iterations=20000, cg_tolerance=1e-10,
                               line_search=ls).minimize(a, x0, prob)
        tn = clock() - t0
        e1 = a.assemble_energy_per_element(x1) if name == "backtracking" else None
        steps[name] = {"seconds": tn, "cg_iterations": rep.cg_iterations, "step_length": rep.last_alpha,
                       "grad_norm_after": rep.grad_norm if np.isfinite(rep.grad_norm) else None,
                       "valid_after": bool(np.isfinite(rep.grad_norm))}
        if e1 is not None:
            steps[name]["energy_after"] = float(e1.sum())
    print(json.dumps({"workload": "P1 NeoHookean 20^3x6 (48,000 tets, 27,783 dofs), FP64, 1 qp",
                      "e_g_h_ms_host_api": 1e3 * statistics.median(ts),
                      "elem_qp_per_s_host_api": m.n_elements / statistics.median(ts),
                      "grad_norm_before": g0, "newton_step": steps}), flush=True)
    a.close()


def batch_run(n_problems, newton_cap, rtol, cg_cap, log):
    """The heterogeneous batch: n_problems box meshes with tet counts
    log-uniform in [1k, 200k], problem i P1 if i is even, LinearElasticity if
    i // 2 is even; per problem E log-normal (mean ln 1e6, sigma 0.5), nu
    uniform [0.3, 0.45], bottom clamped and top pulled by 2-8 % of the height;
    mixed precision, one batch handle (pfgpu_batch_create: every element
    kernel and node gather one launch over all problems), graph-replayed CG.
    Mesh generation and the handle's setup are untimed."""
    import numpy as np
    from polyfem_amd import BatchSolver, NewtonSolver
    from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, lame_from_young

    rng = np.random.default_rng(BASE_SEED + 21)
    probs = []
    t_setup = time.perf_counter()
    for i in range(n_problems):
        order = 1 + (i % 2)
        material = "LinearElasticity" if (i // 2) % 2 == 0 else "NeoHookean"
        cells = np.exp(rng.uniform(np.log(1000), np.log(200000))) / 6
        aspect = rng.uniform(1.0, 3.0)
        nx = max(1, int(round((cells / aspect) ** (1 / 3))))
        nz = max(1, int(round(cells / (nx * nx))))
        mb = kuhn_box(nx, nx, nz, order=order, size=(1.0, 1.0, nz / nx))
        lam_i, mu_i = lame_from_young(np.exp(rng.normal(np.log(1e6), 0.5)), rng.uniform(0.3, 0.45))
        pull = rng.uniform(0.02, 0.08) * (nz / nx)
        bot, top = face_nodes(mb, 2, 0.0), face_nodes(mb, 2, nz / nx)
        fixed = np.concatenate([3 * bot[:, None] + np.arange(3), 3 * top[:, None] + np.arange(3)]).ravel()
        vals = np.concatenate([np.zeros(3 * bot.size), np.tile([0.0, 0.0, pull], top.size)])
        x0 = np.zeros(mb.n_dofs)
        x0[2::3] = pull * mb.nodes[:, 2] / (nz / nx)
        probs.append(dict(mesh=mb, material=material, lame=(float(lam_i), float(mu_i)), fixed_dofs=fixed,
                          fixed_values=vals, x0=x0))
    n_tets = sum(p["mesh"].n_elements for p in probs)
    newton = NewtonSolver(max_iterations=newton_cap, grad_norm=0.0, relative_grad_norm=rtol,
                          cg_max_iterations=cg_cap, cg_tolerance=1e-6)
    bs = BatchSolver(probs, precision="mixed", newton=newton)
    t_setup = time.perf_counter() - t_setup
    log(f"batch: {len(probs)} problems, {n_tets} tets, setup {t_setup:.1f} s")
    t0 = _clock()
    xs, reps = bs.solve()
    t_batch = _clock() - t0
    red = np.array([r.grad_norm / r.initial_grad_norm if r.initial_grad_norm > 0 else 0.0 for r in reps])
    res = {
        "workload": f"batch of {len(probs)} box meshes, {n_tets} tets in all (1k-200k each, half P1 / half P2, "
                    "half LinearElasticity / half NeoHookean), mixed precision, one batch handle",
        "newton": {"max_iterations": newton_cap, "relative_grad_norm": rtol,
                   "cg_max_iterations": cg_cap, "cg_tolerance": 1e-6},
        "converged": int(sum(r.converged for r in reps)), "newton_iterations_max": max(r.iterations for r in reps),
        "solve_s": t_batch, "problems_per_s": len(probs) / t_batch, "setup_s": t_setup,
        "cg_iterations_total": int(sum(r.cg_iterations for r in reps)),
        "grad_reduction": {"median": float(np.median(red)), "max": float(red.max()),
                           "below_1e-6": int((red <= 1e-6).sum())}}
    bs.close()
    del bs, probs
    return res


def suite_batch():
    args = _ARGS
    print(json.dumps(batch_run(args.batch_problems, args.batch_newton, args.batch_rtol, args.batch_cg, _log)),
          flush=True)


def suite_cantilever():
    import numpy as np
    from polyfem_amd import NewtonSolver, NLAssembler, NLProblem, TransientSimulation, cg_solve
    from polyfem_amd.mesh import BASE_SEED, face_nodes, kuhn_box, random_lame, random_smooth_field
    args, clock, log = _ARGS, _clock, _log

    # ---- cantilever: 160 x 40 x 40 cells over [0,4]x[0,1]x[0,1], P1 NeoHookean, FP64,
    # clamped at x = 0, gravity rho g (rho = 1000, g = 0.2 m/s^2 along -z, a dead load:
    # the nodal vector f = M g of the initial geometry); Lame parameters log-normal
    # (sigma 0.3); J = sum over top-face vertices of |X + u - X*|^2 with X* = X + u* and u*
    # the equilibrium of a second Lame draw.  dJ/d(lambda, mu) and dJ/dX by the static
    # adjoint, checked by central differences (h = 1e-6 relative for the Lame
    # parameters, 1e-6 absolute for the vertices) through full Newton solves.
    m = kuhn_box(160, 40, 40, order=1, size=(4.0, 1.0, 1.0))
    E = m.n_elements
    lam, mu = random_lame(E, seed=BASE_SEED + 31, E_sigma=0.3)
    lam2, mu2 = random_lame(E, seed=BASE_SEED + 32, E_sigma=0.3)
    clamp = face_nodes(m, 0, 0.0)
    prob = NLProblem(m.n_dofs, (3 * clamp[:, None] + np.arange(3)).ravel())
    v2n = np.empty(m.vertices.shape[0], dtype=np.int64)
    v2n[m.elem_vertices.ravel()] = m.elem_nodes[:, :4].ravel()
    S = np.flatnonzero(np.abs(m.vertices[:, 2] - 1.0) < 1e-12)
    Sd = (3 * v2n[S][:, None] + np.arange(3)).ravel()
    newton = NewtonSolver(max_iterations=25, grad_norm=1e-8, cg_max_iterations=20000, cg_tolerance=1e-10)
    a = NLAssembler("NeoHookean", m, lame=(lam2, mu2), precision="fp64")
    g = np.zeros(m.n_dofs)
    g[2::3] = -0.2
    f = TransientSimulation(a, rho=1000.0).mass_matvec(g)
    V0 = m.vertices.copy()

    def equilibrium(u0):
        a.set_load(f)
        t0 = clock()
        u, rep = newton.minimize(a, u0, prob)
        if not rep.converged:
            raise RuntimeError(f"cantilever Newton did not converge: {rep}")
        return u, rep, clock() - t0

    ustar = equilibrium(np.zeros(m.n_dofs))[0]
    Xstar = V0[S] + ustar[Sd].reshape(-1, 3)
    a.update_lame_params(lam, mu)
    u, rep, t_fwd = equilibrium(np.zeros(m.n_dofs))
    log(f"cantilever forward: {t_fwd:.2f} s, {rep.iterations} Newton, {rep.cg_iterations} CG")

    def J(V, uu):
        return float(np.sum((V[S] + uu[Sd].reshape(-1, 3) - Xstar) ** 2))

    r = V0[S] + u[Sd].reshape(-1, 3) - Xstar
    dJdu = np.zeros(m.n_dofs)
    dJdu[Sd] = 2 * r.ravel()
    t0 = clock()
    p, adj_it, adj_rel = cg_solve(a, u, dJdu, prob.mask, max_iterations=20000, tolerance=1e-10)
    t_adj = clock() - t0
    t0 = clock()
    dtheta = a.force_material_derivative(u, p)
    shape_term = a.force_shape_derivative(u, p)
    t_vjp = clock() - t0
    explicit = np.zeros_like(V0)
    explicit[S] = 2 * r
    dX = explicit.ravel() - shape_term
    rng = np.random.default_rng(BASE_SEED + 33)
    h = 1e-6
    fd_m, fd_s = [], []
    t_fd = time.perf_counter()
    for _ in range(args.fd_directions):
        dl, dm = rng.normal(size=E) * lam, rng.normal(size=E) * mu
        a.update_lame_params(lam + h * dl, mu + h * dm)
        jp = J(V0, equilibrium(u)[0])
        a.update_lame_params(lam - h * dl, mu - h * dm)
        jm = J(V0, equilibrium(u)[0])
        lin = float(dtheta[:E] @ dl + dtheta[E:] @ dm)
        fd_m.append(abs((jp - jm) / (2 * h) - lin) / abs(lin))
    a.update_lame_params(lam, mu)
    for _ in range(args.fd_shape_directions):
        dV = random_smooth_field(V0, 1.0, int(rng.integers(1 << 62))).reshape(-1, 3)
        dV[m.vertices[:, 0] < 1e-12] = 0.0  # the clamped face stays where it is
        a.set_vertices(V0 + h * dV)
        jp = J(V0 + h * dV, equilibrium(u)[0])
        a.set_vertices(V0 - h * dV)
        jm = J(V0 - h * dV, equilibrium(u)[0])
        lin = float(dX @ dV.ravel())
        fd_s.append(abs((jp - jm) / (2 * h) - lin) / abs(lin))
    a.set_vertices(V0)
    t_fd = time.perf_counter() - t_fd
    log(f"cantilever FD: material max rel err {max(fd_m, default=0):.2e}, shape {max(fd_s, default=0):.2e}")
    print(json.dumps({
        "workload": f"P1 NeoHookean cantilever 160x40x40 cells ({E} tets, {m.n_dofs} dofs), FP64, gravity",
        "forward_newton_s": t_fwd, "newton_iterations": rep.iterations, "newton_cg_iterations": rep.cg_iterations,
        "objective": J(V0, u), "adjoint_cg_s": t_adj, "adjoint_cg_iterations": adj_it,
        "adjoint_rel_residual": adj_rel, "material_and_shape_vjp_s": t_vjp,
        "grad_norms": {"dJ_dtheta": float(np.linalg.norm(dtheta)), "dJ_dX": float(np.linalg.norm(dX))},
        "fd_check": {"h": h, "material_directions": len(fd_m), "material_max_rel_err": max(fd_m, default=None),
                     "shape_directions": len(fd_s), "shape_max_rel_err": max(fd_s, default=None),
                     "seconds": t_fd}}), flush=True)
    a.close()


def suite_transient():
    import numpy as np
    from polyfem_amd import NewtonSolver, NLAssembler, TransientSimulation
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, random_smooth_field
    args, clock, log = _ARGS, _clock, _log

    # ---- implicit Euler: 80^3 x 6 P1 tets, rho = 1000, dt = 1e-3, 100 steps, checkpoints
    # every 10, v0 a seeded random smooth field of amplitude 0.5; J = 1/2 v_T^T M v_T +
    # E(u_T); dJ/dv0 and dJ/dmu (the adjoint's implicit part plus the explicit
    # dE(u_T)/dmu).  Timed in mixed precision and FP64; the FP64 gradient is checked
    # by central differences (h = 1e-6) on a few random directions.
    m = kuhn_box(80, 80, 80, order=1)
    E = m.n_elements
    lam, mu = random_lame(E, seed=BASE_SEED + 41)
    v0 = random_smooth_field(m.nodes, 0.5, BASE_SEED + 42)

    def run(asm, sim, vv):
        xT, vT = sim.forward(np.zeros(m.n_dofs), vv)
        MvT = sim.mass_matvec(vT)
        return xT, vT, MvT, float(0.5 * vT @ MvT + asm.assemble_energy(xT))

    out = {"workload": f"implicit Euler, P1 NeoHookean 80^3x6 ({E} tets, {m.n_dofs} dofs), dt 1e-3, "
                       f"{args.transient_steps} steps, checkpoints every 10"}
    tol = {"mixed": 5e-8, "fp64": 1e-11}
    for precision in args.transient_precisions.split(","):
        gn = tol[precision]
        asm = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=precision)
        sim = TransientSimulation(asm, rho=1000.0, dt=1e-3, time_steps=args.transient_steps, checkpoint_every=10,
                                  newton=NewtonSolver(max_iterations=20, grad_norm=gn, cg_max_iterations=5000,
                                                      cg_tolerance=1e-10))
        # untimed: one assembly builds the handle's tables (pattern maps, gather
        # plans), once per mesh as for the other workloads
        asm.hessian_values(np.zeros(m.n_dofs))
        t0 = clock()
        xT, vT, MvT, Jv = run(asm, sim, v0)
        t_fwd = clock() - t0
        t0 = clock()
        dth, _, dv0 = sim.adjoint(asm.assemble_gradient(xT), MvT, cg_max_iterations=5000, cg_tolerance=1e-10)
        t_adj = clock() - t0
        dmu = dth[E:] + asm.energy_lame_gradient(xT)[E:]
        log(f"transient {precision}: forward {t_fwd:.2f} s, adjoint {t_adj:.2f} s")
        out[precision] = {"forward_s": t_fwd, "adjoint_s_with_recomputation": t_adj, "objective": Jv,
                          "work": sim.stats(),
                          "grad_norms": {"dJ_dv0": float(np.linalg.norm(dv0)), "dJ_dmu": float(np.linalg.norm(dmu))}}
        if precision == "mixed":
            g_mixed = (dv0, dmu)
            asm.close()
            continue
        if "mixed" in out:
            out["mixed_vs_fp64_gradient_rel_err"] = {
                "dJ_dv0": float(np.linalg.norm(g_mixed[0] - dv0) / np.linalg.norm(dv0)),
                "dJ_dmu": float(np.linalg.norm(g_mixed[1] - dmu) / np.linalg.norm(dmu))}
        rng = np.random.default_rng(BASE_SEED + 43)
        h = 1e-6
        fd = []
        for k in range(args.fd_transient_directions):
            if k % 2 == 0:
                dv = random_smooth_field(m.nodes, 0.5, int(rng.integers(1 << 62)))
                jp, jm = run(asm, sim, v0 + h * dv)[3], run(asm, sim, v0 - h * dv)[3]
                lin, what = float(dv0 @ dv), "v0"
            else:
                dm = rng.normal(size=E) * mu
                asm.update_lame_params(lam, mu + h * dm)
                jp = run(asm, sim, v0)[3]
                asm.update_lame_params(lam, mu - h * dm)
                jm = run(asm, sim, v0)[3]
                asm.update_lame_params(lam, mu)
                lin, what = float(dmu @ dm), "mu"
            fd.append({"wrt": what, "fd": (jp - jm) / (2 * h), "adjoint": lin,
                       "rel_err": abs((jp - jm) / (2 * h) - lin) / abs(lin)})
            log(f"transient FD direction: {fd[-1]}")
        out["fd_check_fp64"] = {"h": h, "directions": fd}
        asm.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--path", default="auto", choices=["auto", "fused", "plain", "factored"],
                    help="auto: the launch-shape autotuner picks path and workgroup sizes (untimed setup)")
    ap.add_argument("--gather-threads", type=int, default=0, choices=[0, 64, 128, 256, 512],
                    help="with --path plain/factored: the gather workgroup size (0: the library default, 256)")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--grid", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the error, adjoint and Newton measurements")
    ap.add_argument("--newton-iterations", type=int, default=60)
    ap.add_argument("--cg-iterations", type=int, default=100)
    ap.add_argument("--suite", action="store_true",
                    help="instead of the headline line: one JSON line per further configuration "
                         "(P1 20^3 FP64 E+G+H and Newton step; 1.5M-tet cantilever static adjoint; "
                         "3M-tet implicit-Euler adjoint), timed at full size")
    ap.add_argument("--suite-parts", default="step,batch,cantilever,transient",
                    help="--suite: comma-separated subset of step, batch, cantilever, transient")
    ap.add_argument("--batch-problems", type=int, default=256,
                    help="problems in the heterogeneous batch (an extra of the default run, and --suite); 0 skips it")
    ap.add_argument("--batch-newton", type=int, default=20, help="--suite: Newton iteration cap of the batch")
    ap.add_argument("--batch-rtol", type=float, default=1e-4,
                    help="--suite: per-problem gradient reduction at which a batch problem has converged")
    ap.add_argument("--batch-cg", type=int, default=2000, help="--suite: CG iteration cap of the batch")
    ap.add_argument("--fd-directions", type=int, default=64,
                    help="--suite: random Lame directions of the cantilever's finite-difference check")
    ap.add_argument("--fd-shape-directions", type=int, default=4,
                    help="--suite: random vertex directions of the cantilever's finite-difference check")
    ap.add_argument("--transient-steps", type=int, default=100, help="--suite: implicit-Euler steps")
    ap.add_argument("--transient-precisions", default="mixed,fp64",
                    help="--suite: precisions of the transient run (FD check on fp64, after mixed)")
    ap.add_argument("--fd-transient-directions", type=int, default=4,
                    help="--suite: random directions of the transient finite-difference check")
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

    if args.suite:
        suite(args)
        return

    from polyfem_amd import NLAssembler
    from polyfem_amd.mesh import BASE_SEED, kuhn_box, random_lame, smooth_displacement

    n = args.grid
    m = kuhn_box(n, n, n, order=2)
    lam, mu = random_lame(m.n_elements, seed=BASE_SEED + 1)
    x_host = smooth_displacement(m.nodes, 0.05, 1e-3, BASE_SEED)
    a = NLAssembler("NeoHookean", m, lame=(lam, mu), precision=args.precision,
                    path=None if args.path == "auto" else args.path)
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
    if args.path in ("plain", "factored") and args.gather_threads:
        a.set_tuning(args.path, args.gather_threads, 256)

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
        # every host core of this process's share (16 per GPU on the MI355X boxes)
        cores = min(int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1), os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline_sample(cores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
