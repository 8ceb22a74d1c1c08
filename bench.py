#!/usr/bin/env python
"""Benchmark: P1 kappa-Laplace assembly + Jacobi-PCG on one MI355X.

Workload: 3D variable-coefficient Poisson -div(kappa grad u) = 1 on the unit
cube, 128^3 Kuhn grid x 6 P1 tetrahedra (12,582,912 elements, 2,146,689
nodes, nnz 31,802,497), kappa = exp(g) with g a seeded Gaussian random field
(length 0.1, sigma 1, seed 0), zero Dirichlet.  A step is one assembly of the
stiffness values from resident coordinates / kappa into the precomputed CSR
pattern (FP64 geometry and accumulation, FP32 values).  The solve leg then
condenses and runs Jacobi-PCG to relative residual 1e-8 (FP64 vectors).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model, sockets = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return model, max(1, len(sockets))


def _cpu_phases(n: int, threads: int, repeats: int, solve: str, pcg_iters: int = 30):
    """Torch-CPU float64 restatement of the reference pipeline on the n^3 Kuhn
    grid with ``threads`` threads: pattern (SciPy, once), Map + Reduce and
    condensation (median of ``repeats`` after one warm-up), Jacobi-PCG
    (``solve`` = "full": to 1e-8 once; "iters": ``pcg_iters`` iterations)."""
    import numpy as np
    import torch
    import warnings
    from oracle import pattern as op
    from oracle import torch_cpu as tc
    from tensormesh_amd.dataset import lognormal_kappa_cube
    from tensormesh_amd.mesh import kuhn_cube_arrays
    torch.set_num_threads(threads)
    pts, conn, bnd = kuhn_cube_arrays(n)
    kappa = lognormal_kappa_cube(n, 0.1, 1.0, 0)
    t0 = time.perf_counter()
    u, v, pe = op.pattern_and_slots(conn, len(pts))
    t_pat = time.perf_counter() - t0
    P, C, K, PE = (torch.from_numpy(a) for a in (pts, conn, kappa, pe))
    med = lambda ts: sorted(ts)[len(ts) // 2]
    ta = []
    for it in range(repeats + 1):
        t0 = time.perf_counter()
        vals = tc.assemble_kappa_laplace(P, C, K, PE, len(u))
        if it:
            ta.append(time.perf_counter() - t0)
    row, col = torch.from_numpy(u.astype(np.int64)), torch.from_numpy(v.astype(np.int64))
    mask = torch.from_numpy(bnd)
    rhs = torch.ones(len(pts), dtype=torch.float64)
    tcd = []
    for it in range(repeats + 1):
        t0 = time.perf_counter()
        kv, kr, kc, ni, fi = tc.condense(vals, row, col, mask, rhs)
        if it:
            tcd.append(time.perf_counter() - t0)
    rp = torch.zeros(ni + 1, dtype=torch.int64)
    rp[1:] = torch.cumsum(torch.bincount(kr, minlength=ni), 0)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        A = torch.sparse_csr_tensor(rp, kc, kv, (ni, ni))
    d = torch.zeros(ni, dtype=torch.float64)
    dm = kr == kc
    d[kr[dm]] = kv[dm]
    t0 = time.perf_counter()
    _, its = tc.jacobi_pcg(A, fi, 1.0 / d, rtol=1e-8, maxiter=100000 if solve == "full" else pcg_iters)
    t_s = time.perf_counter() - t0
    out = {"grid": f"{n}^3", "elements": len(conn), "threads": threads,
           "pattern_s": round(t_pat, 3), "assembly_s": round(med(ta), 4),
           "assembly_melem_per_s": round(len(conn) / med(ta) / 1e6, 4), "condense_s": round(med(tcd), 4),
           "pcg_iterations": its, "pcg_s_per_iter": round(t_s / max(its, 1), 5)}
    if solve == "full":
        out["solve_to_1e-8_s"] = round(t_s, 3)
    return out


def cpu_baseline(n_cpu: int, n_single: int = 48, repeats: int = 3):
    """The reference pipeline restated in float64 torch on the host CPU: all
    threads of the process's affinity set on the n_cpu^3 grid, and one thread
    on the bounded n_single^3 sample; each phase (pattern, assembly,
    condensation, Jacobi-PCG) timed separately."""
    import torch
    k = len(os.sched_getaffinity(0))
    prev = torch.get_num_threads()
    try:
        full = _cpu_phases(n_cpu, k, repeats, "full")
        one = _cpu_phases(n_single, 1, repeats, "iters")
    finally:
        torch.set_num_threads(prev)
    model, sockets = cpu_info()
    return {"value": full["assembly_melem_per_s"], "unit": "Melem/s", "cores": k, "kind": "port",
            "sample": (f"float64 torch-CPU restatement of the reference (SciPy pattern, einsum Map, index_add "
                       f"Reduce, boolean condensation, Jacobi-PCG with torch CSR), kappa-Laplace on the "
                       f"{n_cpu}^3 Kuhn grid with {k} threads (assembly and condensation: median of {repeats} "
                       f"after 1 warm-up; pattern and the solve to 1e-8 once), and on the {n_single}^3 grid "
                       f"with 1 thread (30 PCG iterations)"),
            "all_threads": full, "one_thread": one, "cpu": model, "sockets": sockets}


def learned_leg(n: int, seeds, log_fn):
    """Learned-preconditioner FCG vs Jacobi-PCG to 1e-8 on held-out log-normal
    Poisson problems (n^3 Kuhn grid, GRF length 0.05-0.2, sigma 0.5-2.3)."""
    import numpy as np
    import torch
    from tensormesh_amd.precond import GNNPreconditioner, LearnedFCG
    from tensormesh_amd.precond.problem import PoissonCube, field_params
    from tensormesh_amd.precond.train import plain_solve
    dev = torch.device("cuda:0")
    wpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tensormesh_amd", "precond", "weights",
                         f"poisson{n}_sym.npz")
    if not os.path.exists(wpath):
        return {"skipped": f"no trained weights at {os.path.relpath(wpath)}"}
    t0 = time.perf_counter()
    net = GNNPreconditioner.load(wpath, dev)
    prob = PoissonCube(n, dev, n_levels=net.n_levels)
    solver = LearnedFCG(net, prob.hier, prob.csr, chunk=8, m=1)
    log_fn(f"learned: {n}^3 set-up {time.perf_counter() - t0:.1f}s, levels {[l.n for l in prob.hier.levels]}")
    # warm-up (kernel loading, graph capture) on a training-range seed
    kappa = prob.kappa(999)
    values, b = prob.system(kappa)
    solver.set_system(values, prob.log_kappa_feature(kappa))
    solver.solve(b, rtol=1e-8)
    plain_solve(prob, values, maxiter=50)
    rows = []
    for sd in seeds:
        kappa = prob.kappa(sd)
        values, b = prob.system(kappa)
        torch.cuda.synchronize(dev)
        t1 = time.perf_counter()
        _, itj, _, okj = plain_solve(prob, values)
        torch.cuda.synchronize(dev)
        tj = time.perf_counter() - t1
        t1 = time.perf_counter()
        solver.set_system(values, prob.log_kappa_feature(kappa))
        _, info = solver.solve(b, rtol=1e-8)
        tl = time.perf_counter() - t1
        ell, sig = field_params(sd)
        rows.append({"seed": sd, "length": round(ell, 3), "sigma": round(sig, 3), "jacobi_iters": itj,
                     "jacobi_s": round(tj, 5), "fcg_iters": info["fcg_iterations"],
                     "fallback_iters": info["fallback_iterations"], "fcg_s": round(tl, 5),
                     "relres": info["relres"], "speedup": round(tj / tl, 3)})
    # one network application, HIP events on the launch stream
    st = torch.cuda.current_stream(dev)
    r = torch.randn(prob.n_inner, dtype=torch.float64, device=dev)
    lk = prob.log_kappa_feature(kappa)
    prog = solver.prog
    prog.forward(r, lk)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(st)
    for _ in range(20):
        prog.forward(r, lk)
    ev1.record(st)
    torch.cuda.synchronize(dev)
    t_app = ev0.elapsed_time(ev1) / 1e3 / 20
    # MFMA work: [n x 128] x [128 x 64] per layer; on the fused finest level a
    # level-0 layer pair is rank-2 vector work (no MFMA), so it is not counted
    mfma_levels = [lev for lev in net.schedule if not (lev == 0 and prog.fusable_sym0)]
    flops = sum(2 * prob.hier.levels[lev].n * 128 * 64 for lev in mfma_levels)
    sp = np.array([r_["speedup"] for r_ in rows])
    return {"workload": f"{n}^3 Kuhn log-normal Poisson, held-out seeds {seeds[0]}-{seeds[-1]}",
            "network": {"schedule": list(net.schedule), "symmetric": net.symmetric, "width": 64,
                        "apply_ms": round(t_app * 1e3, 4), "mfma_layer_levels": mfma_levels,
                        "gemm_gflop_per_apply": round(flops / 1e9, 3),
                        "mfma_tflops": round(flops / t_app / 1e12, 2),
                        "mfma_util_vs_2500_tflops_bf16": round(flops / t_app / 2.5e15, 5)},
            "seeds": len(rows),
            "median_speedup_vs_jacobi": float(np.median(sp)),
            "p10_speedup_vs_jacobi": float(np.percentile(sp, 10)),
            "min_speedup_vs_jacobi": float(sp.min()),
            "median_jacobi_iters": float(np.median([r_["jacobi_iters"] for r_ in rows])),
            "median_fcg_iters": float(np.median([r_["fcg_iters"] for r_ in rows])),
            "max_fcg_iters": int(max(r_["fcg_iters"] for r_ in rows)),
            "median_jacobi_s": float(np.median([r_["jacobi_s"] for r_ in rows])),
            "median_fcg_s": float(np.median([r_["fcg_s"] for r_ in rows])),
            "max_true_relres": float(max(r_["relres"] for r_ in rows)),
            "fallbacks": int(sum(r_["fallback_iters"] > 0 for r_ in rows)),
            "worst_rows": sorted(rows, key=lambda r_: r_["speedup"])[:5],
            "speedups": [r_["speedup"] for r_ in rows]}


def batched_leg(n_mesh: int, log_fn):
    """Batched 2D Poisson: n_mesh uneven jittered square meshes (2e3-2e5 elements,
    log-uniform), each with its own log-normal kappa; assembly + condensation +
    Jacobi-PCG to 1e-8 for all of them in one captured hipGraph, tiles autotuned.
    Compared with solving the same condensed systems one at a time with the
    single-system device PCG (the loop a caller writes without the batched path)."""
    import warnings
    import numpy as np
    import torch
    from tensormesh_amd.batched import BatchedPoisson
    from tensormesh_amd.dataset import batched_square_meshes
    from tensormesh_amd.sparse.ops import CSR, pcg_raw
    from tensormesh_amd.tune import autotune_batched
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    ms = batched_square_meshes(n_mesh, seed=0)
    bp = BatchedPoisson([m[:3] for m in ms], dev)
    bp.kappa.copy_(torch.tensor(np.concatenate([m[3] for m in ms]), dtype=torch.float32, device=dev))
    t_setup = time.perf_counter() - t0
    log_fn(f"batched: {n_mesh} meshes, {bp.n_elements} elements, {bp.n_inner} unknowns ({t_setup:.1f}s set-up)")
    # two workgroups per CU measured 1.2-14x slower at this size (the grid barrier has twice the
    # arrivals): not searched here; the tuner's GPU test covers that candidate
    space = {"rows_per_tile": (1024, 2048, 4096), "lanes": (4, 8), "blocks_per_cu": (1,)}
    log_fn("batched: autotuning " + ", ".join(f"{k} {v}" for k, v in space.items()))
    cfg = autotune_batched(bp, space=space, use_cache=False, maxiter=10000, verbose=True)
    log_fn(f"batched: tuned {[(k, cfg[k]) for k in ('rows_per_tile', 'lanes', 'blocks')]}")
    g = bp.capture(maxiter=10000)
    st = torch.cuda.current_stream(dev)
    g.replay()
    torch.cuda.synchronize(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(5):
        ev0.record(st)
        g.replay()
        ev1.record(st)
        torch.cuda.synchronize(dev)
        times.append(ev0.elapsed_time(ev1) / 1e3)
    _, status, iters, rel = bp.results()
    t_graph = sorted(times)[len(times) // 2]
    log_fn(f"batched: graph {t_graph * 1e3:.1f} ms, converged {int((status == 1).sum())}/{len(status)}")
    # the same condensed systems, one device PCG call each
    rp = bp.rp_in.cpu().numpy()
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for k in range(bp.n_mesh):
            a, z = int(bp.inner_off[k]), int(bp.inner_off[k + 1])
            ea, ez = int(rp[a]), int(rp[z])
            csr = CSR(torch.as_tensor(rp[a:z + 1] - rp[a], dtype=torch.int32, device=dev),
                      (bp.ci_in[ea:ez] - a).contiguous(), z - a, z - a)
            pcg_raw(csr, bp.vals_in[ea:ez], bp.b_in[a:z], rtol=1e-8)
    torch.cuda.synchronize(dev)
    t_loop = time.perf_counter() - t1
    return {"workload": f"{n_mesh} P1 triangle meshes of the unit square, elements log-uniform 2e3-2e5 "
                        f"(seed 0), jitter 0.2 h, kappa=exp(GRF l=U[0.05,0.2])",
            "elements": int(bp.n_elements), "unknowns": int(bp.n_inner), "nnz": int(bp.nnz_inner),
            "graph_s": round(t_graph, 5), "melem_per_s": round(bp.n_elements / t_graph / 1e6, 2),
            "all_converged": bool((status == 1).all()), "max_relres": float(rel.max()),
            "median_iters": float(np.median(iters)), "max_iters": int(iters.max()),
            "tuned": {k: cfg[k] for k in ("rows_per_tile", "lanes", "blocks")},
            "loop_of_single_solves_s": round(t_loop, 4), "speedup_vs_loop": round(t_loop / t_graph, 2),
            "setup_s": round(t_setup, 1)}


def hyperelastic_leg(n, log_fn):
    """Neo-Hookean 64^3 leg: tangent + residual assembly, linear-elastic stiffness and the
    adjoint action (HIP events on the launch stream), then Newton with block Jacobi-PCG."""
    import numpy as np
    import torch
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays
    dev = torch.device("cuda:0")
    pts, conn, bnd = kuhn_cube_arrays(n)
    E_np = 1.0 + np.random.default_rng(1).uniform(-0.1, 0.1, len(conn))
    mesh = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    Et = torch.tensor(E_np, dtype=torch.float32, device=dev)
    model = tm.NeoHookeanModel.from_mesh(mesh, E=Et, nu=0.3).to(dev)
    lin = tm.LinearElasticityElementAssembler.from_assembler(model, E=Et, nu=0.3).to(dev)
    lin.value_dtype = torch.float32
    z = torch.tensor(pts[:, 2], device=dev)
    fixed = (z == 0) | (z == 1)
    u0 = torch.zeros((len(pts), 3), dtype=torch.float64, device=dev)
    u0[:, 2] = 0.1 * z
    E, N, nnz = len(conn), len(pts), model.nnz
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn, reps=20):
        for _ in range(20):   # clocks ramp up over the first launches
            fn()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(reps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / 1e3 / reps

    t_tr = timed(lambda: model._launch(u0, Et, want_k=True, want_r=True))
    t_lin = timed(lambda: lin.stiffness(E=Et))
    lam = torch.rand((N, 3), dtype=torch.float64, device=dev)
    t_adj = timed(lambda: model._tangent_action(u0, Et, lam, want_dE=True))
    # conn + coords + u + E + row_ptr + col_idx + 3x3 f32 blocks + residual
    b_tr = 16 * E + 24 * N + 24 * N + 4 * E + 4 * (N + 1) + 4 * nnz + 36 * nnz + 24 * N
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    u, info = model.newton(u0, fixed, rtol=1e-8)
    torch.cuda.synchronize()
    t_newton = time.perf_counter() - t0
    log_fn(f"hyperelastic {n}^3: tangent+residual {t_tr * 1e6:.1f} us, linear stiffness {t_lin * 1e6:.1f} us, "
           f"adjoint action {t_adj * 1e6:.1f} us; Newton {info['iterations']} its in {t_newton:.2f}s")
    return {"workload": f"compressible neo-Hookean, {n}^3 Kuhn grid x 6 P1 tets, E=1+U(-0.1,0.1) per element "
                        f"(seed 1), nu=0.3, bottom clamped, top u_z=0.1, Newton from u_z=0.1 z",
            "elements": E, "nodes": N, "block_nnz": nnz,
            "tangent_residual_us": round(t_tr * 1e6, 1), "tangent_residual_melem_per_s": round(E / t_tr / 1e6, 1),
            "tangent_residual_GBps": round(b_tr / t_tr / 1e9, 1), "tangent_residual_bytes": b_tr,
            "linear_elastic_stiffness_us": round(t_lin * 1e6, 1), "adjoint_action_us": round(t_adj * 1e6, 1),
            "newton_iterations": info["iterations"], "newton_converged": bool(info["converged"]),
            "newton_relres": info["residual_history"][-1], "pcg_iterations": info["linear_iterations"],
            "newton_s": round(t_newton, 3)}


def reference_poisson_leg(sizes, log_fn):
    """The reference's published 3D Poisson benchmark setting (unit cube, P1 tetrahedra, f = 1,
    constant coefficient, Jacobi-preconditioned Krylov to rtol = atol = 1e-10) on Kuhn grids of
    about the same unknown counts: wall time of assembly + load + condensation + Jacobi-PCG with
    the mesh, pattern and plans set up beforehand (one untimed pass first)."""
    import torch
    import tensormesh_amd as tm
    dev = torch.device("cuda:0")
    rows = []
    for n in sizes:
        t0 = time.perf_counter()
        mesh = tm.kuhn_cube(n)
        asm = tm.LaplaceElementAssembler.from_mesh(mesh).to(dev)
        asm.value_dtype = torch.float32
        load = tm.const_node_assembler(1.0).from_mesh(mesh).to(dev)
        cond = tm.Condenser(mesh.boundary_mask.to(dev))
        asm.topology.device_plan(dev)
        torch.cuda.synchronize(dev)
        t_setup = time.perf_counter() - t0

        def run():
            K = asm()
            b = load()
            Ki, fi = cond(K, b)
            x = Ki.solve(fi, tol=1e-10, atol=1e-10, maxiter=20000)
            return Ki, fi, x

        run()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        Ki, fi, x = run()
        torch.cuda.synchronize(dev)
        t_run = time.perf_counter() - t0
        info = Ki.last_solve_info
        rel = ((fi.double() - Ki @ x).norm() / fi.double().norm()).item()
        rows.append({"n": n, "unknowns": Ki.shape[0], "elements": int(mesh.n_elements()),
                     "assemble_condense_solve_s": round(t_run, 4), "pcg_iterations": info["iterations"],
                     "relres": rel, "setup_s": round(t_setup, 2)})
        log_fn(f"reference-style 3D Poisson n={n}: {Ki.shape[0]} unknowns, {t_run * 1e3:.1f} ms "
               f"({info['iterations']} its, relres {rel:.1e})")
        del mesh, asm, load, cond, Ki, fi, x
        torch.cuda.empty_cache()
    return {"workload": "3D Poisson -lap u = 1, unit cube, Kuhn P1 tets, homogeneous Dirichlet, constant "
                        "coefficient; Jacobi-PCG to ||r|| <= max(1e-10 ||b||, 1e-10) (the reference's published "
                        "benchmark: BiCGSTAB + Jacobi, rtol = atol = 1e-10, gmsh meshes)",
            "rows": rows}


def reference_elasticity_leg(sizes, log_fn):
    """The reference's published 3D linear-elasticity benchmark setting (3 DOF per node, Jacobi-
    preconditioned Krylov to rtol = atol = 1e-10) on unit-cube Kuhn grids of about the same DOF
    counts: E = 1, nu = 0.3, bottom face clamped, unit downward body force.  Wall time of the
    stiffness assembly + load + block condensation + block Jacobi-PCG (set-up excluded, one
    untimed pass first)."""
    import torch
    import tensormesh_amd as tm
    from tensormesh_amd.assemble.hyperelastic import BlockCondenser
    from tensormesh_amd.sparse.ops import pcg_raw
    dev = torch.device("cuda:0")
    rows = []
    for n in sizes:
        t0 = time.perf_counter()
        mesh = tm.kuhn_cube(n)
        asm = tm.LinearElasticityElementAssembler.from_mesh(mesh, E=1.0, nu=0.3).to(dev)
        asm.value_dtype = torch.float32
        load = tm.const_node_assembler(1.0).from_mesh(mesh).to(dev)
        free = mesh.points[:, 2].to(dev) > 0.0
        cond = BlockCondenser(asm, free)
        asm.stiffness()   # builds the block plan
        torch.cuda.synchronize(dev)
        t_setup = time.perf_counter() - t0

        def run():
            K = asm()
            f = torch.zeros((mesh.n_points, 3), dtype=torch.float64, device=dev)
            f[:, 2] = -load()
            vals, csr = cond.condense(K.values)
            rhs = f[free].reshape(-1)
            return vals, csr, rhs, pcg_raw(csr, vals, rhs, rtol=1e-10, atol=1e-10, maxiter=50000, block_size=3)

        run()
        torch.cuda.synchronize(dev)
        t0 = time.perf_counter()
        vals, csr, rhs, (x, it, rel, ok) = run()
        torch.cuda.synchronize(dev)
        t_run = time.perf_counter() - t0
        from tensormesh_amd.sparse.ops import _spmv_raw
        true_rel = ((rhs - _spmv_raw(csr, vals, x, 3)).norm() / rhs.norm()).item()
        rows.append({"n": n, "dofs": int(rhs.numel()), "elements": int(mesh.n_elements()),
                     "assemble_condense_solve_s": round(t_run, 4), "pcg_iterations": int(it), "converged": bool(ok),
                     "relres": true_rel, "setup_s": round(t_setup, 2)})
        log_fn(f"reference-style 3D elasticity n={n}: {rhs.numel()} DOFs, {t_run * 1e3:.1f} ms "
               f"({it} its, relres {true_rel:.1e})")
        del mesh, asm, load, cond, vals, csr, rhs, x
        torch.cuda.empty_cache()
    return {"workload": "3D linear elasticity, unit cube, Kuhn P1 tets, E=1 nu=0.3, bottom clamped, body force "
                        "(0,0,-1); block Jacobi-PCG to ||r|| <= max(1e-10 ||b||, 1e-10) (the reference's published "
                        "benchmark: hollow cube from gmsh, BiCGSTAB + Jacobi, rtol = atol = 1e-10)",
            "rows": rows}


def convection_diffusion_leg(m, log_fn):
    """Non-symmetric leg: -lap u + beta . grad u = 1 on the unit square (m x m x 2 P1 triangles,
    zero Dirichlet), assembled by the general constant-form kernel and solved by Jacobi-BiCGSTAB
    to 1e-8 (set-up and one untimed pass first).  Deterministic assembly mode for the matrix and
    the load: the iteration count is the same in every run."""
    import torch
    import tensormesh_amd as tm
    from tensormesh_amd.mesh import structured_rectangle
    dev = torch.device("cuda:0")
    mesh = structured_rectangle(m)
    beta = torch.tensor([20.0, -10.0], dtype=torch.float64)

    class ConvDiff(tm.ElementAssembler):
        def forward(self, gradu, gradv, v):
            return gradu @ gradv + (beta.to(gradu.dtype) @ gradu) * v

    asm = ConvDiff.from_mesh(mesh).to(dev)
    asm.value_dtype = torch.float32
    # deterministic assembly: the matrix repeats bit for bit, so does BiCGSTAB's iteration count
    asm.deterministic = True
    load = tm.const_node_assembler(1.0).from_mesh(mesh).to(dev)
    load.deterministic = True
    cond = tm.Condenser(mesh.boundary_mask.to(dev))

    def run():
        K = asm()
        Ki, fi = cond(K, load())
        x = Ki.solve(fi, tol=1e-8, atol=0.0, maxiter=50000)
        return Ki, fi, x

    run()
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    Ki, fi, x = run()
    torch.cuda.synchronize(dev)
    t_run = time.perf_counter() - t0
    info = Ki.last_solve_info
    rel = ((fi.double() - Ki @ x).norm() / fi.double().norm()).item()
    it = max(info["iterations"], 1)
    # per iteration: 2 SpMVs (4(N+1) + 8 nnz + 16 N each) + vector passes (24 N bytes read/written per
    # vector touched: r, r0h, p, v, s, t, phat, shat, minv, x)
    n, nnz = Ki.shape[0], Ki.nnz
    b_it = 2 * (4 * (n + 1) + 8 * nnz + 16 * n) + 8 * n * 22
    log_fn(f"convection-diffusion {m}x{m}: {n} unknowns, {info['method']} {info['iterations']} its, "
           f"{t_run * 1e3:.1f} ms, relres {rel:.1e}")
    return {"workload": f"-lap u + (20,-10) . grad u = 1, unit square, {m}x{m}x2 P1 triangles, zero Dirichlet; "
                        "general constant-form kernel (deterministic mode) + Jacobi-BiCGSTAB to 1e-8",
            "unknowns": n, "nnz": nnz, "method": info["method"], "iterations": info["iterations"],
            "assemble_condense_solve_s": round(t_run, 4), "relres": rel,
            "ms_per_iter": round(t_run / it * 1e3, 4), "iter_GBps_est": round(b_it / (t_run / it) / 1e9, 1)}


def _event_time(fn, reps, dev):
    """Mean seconds per call of ``fn`` between HIP events on the current stream (one untimed call first)."""
    import torch
    fn()
    torch.cuda.synchronize(dev)
    stream = torch.cuda.current_stream(dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize(dev)
    return e0.elapsed_time(e1) / 1e3 / reps


def forms_leg(mesh, asm, pd, t_default, log_fn):
    """The other assembly paths on the headline mesh: the deterministic mode of the kappa-Laplace
    kernel, a spatially varying bilinear form (moments kernel), a spatially varying load
    (``func_node_assembler``) and a facet load over the boundary.  End-to-end call times include
    the per-call probe of the user's ``forward`` on the quadrature points."""
    import torch
    import tensormesh_amd as tm
    from tensormesh_amd.assemble import forms
    dev = torch.device("cuda:0")
    out = {}
    asm.deterministic = True
    try:
        t_det = _event_time(lambda: asm(point_data=pd), 20, dev)
    finally:
        asm.deterministic = False
    out["kappa_laplace_deterministic_us"] = round(t_det * 1e6, 1)
    out["kappa_laplace_us"] = round(t_default * 1e6, 1)
    log_fn(f"forms: deterministic kappa-Laplace {t_det * 1e6:.1f} us ({t_det / t_default:.2f}x the default)")

    class VaryingDiffusion(tm.ElementAssembler):
        def forward(self, gradu, gradv, x):
            return (1.0 + x[0] * x[1]) * (gradu @ gradv) + x[2] * (gradu[0] * gradv[1])

    vd = VaryingDiffusion.from_assembler(asm)
    vd.value_dtype = torch.float32
    pts = vd.points.to(dev)
    conn = vd.elements["tetra"].to(dev)
    t_call = _event_time(lambda: vd(), 5, dev)
    mom, _ = forms.probe_bilinear_q(vd.forward, pts, conn, "tetra", vd.quadrature_order, {"x": pts}, {}, {})
    p64 = pts.to(torch.float64).contiguous()
    t_kernel = _event_time(lambda: vd._launch_moments(p64, mom, vd.value_dtype), 20, dev)
    E = conn.shape[0]
    # moments (49 doubles per element) + coordinates + values, the kernel's algorithmic bytes
    b_mom = 8 * 49 * E + 24 * pts.shape[0] + 4 * asm.nnz
    out["moments_form_call_ms"] = round(t_call * 1e3, 2)
    out["moments_kernel_us"] = round(t_kernel * 1e6, 1)
    out["moments_kernel_GBps"] = round(b_mom / t_kernel / 1e9, 1)
    log_fn(f"forms: spatially varying bilinear form {t_call * 1e3:.1f} ms per call, moments kernel "
           f"{t_kernel * 1e6:.1f} us ({b_mom / t_kernel / 1e9:.0f} GB/s on its moments)")

    load = tm.func_node_assembler(lambda x: torch.sin(3.0 * x[0]) * x[1] + x[2]).from_assembler(
        tm.const_node_assembler(1.0).from_mesh(mesh)).to(dev)
    t_load = _event_time(lambda: load(), 5, dev)
    out["func_load_call_ms"] = round(t_load * 1e3, 2)

    class Traction(tm.FacetAssembler):
        def forward(self, v, x):
            return (x[0] + 2.0 * x[1]) * v

    fa = Traction.from_mesh(mesh).to(dev)
    t_facet = _event_time(lambda: fa(), 5, dev)
    out["facet_load_call_ms"] = round(t_facet * 1e3, 2)
    out["boundary_facets"] = fa.n_facets
    log_fn(f"forms: func_node load {t_load * 1e3:.1f} ms, facet load over {fa.n_facets} facets "
           f"{t_facet * 1e3:.1f} ms per call")
    del mom
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=100,
                    help="untimed assembly steps first: the clocks ramp up over the first ~60 launches")
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-hyper", action="store_true")
    ap.add_argument("--no-refpoisson", action="store_true")
    ap.add_argument("--no-forms", action="store_true")
    ap.add_argument("--hyper-n", type=int, default=64)
    ap.add_argument("--batched-meshes", type=int, default=256)
    ap.add_argument("--learned-n", type=int, default=96)
    ap.add_argument("--learned-seeds", type=int, default=200)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="MiB of a plain device copy run before the timed region (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the assembly and solver kernels here run on a single "
              "device (no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import tensormesh_amd as tm
    from tensormesh_amd.dataset import lognormal_kappa_cube
    from tensormesh_amd.mesh import Mesh, kuhn_cube_arrays
    from tensormesh_amd.sparse.ops import pcg_raw

    dev = torch.device("cuda:0")
    n = args.n
    t0 = time.perf_counter()
    pts, conn, bnd = kuhn_cube_arrays(n)
    kappa_np = lognormal_kappa_cube(n, 0.1, 1.0, 0)
    mesh = Mesh(torch.from_numpy(pts), {"tetra": torch.from_numpy(conn)}, {"is_boundary": torch.from_numpy(bnd)})
    log(f"mesh {n}^3: {len(conn)} tets, {len(pts)} nodes ({time.perf_counter() - t0:.1f}s)")
    t0 = time.perf_counter()
    asm = tm.KappaLaplaceElementAssembler.from_mesh(mesh)
    t_pattern = time.perf_counter() - t0
    asm = asm.to(dev)
    asm.value_dtype = torch.float32           # FP32 values, FP64 geometry and accumulation
    t0 = time.perf_counter()
    plan = asm.topology.device_plan(dev, edge=True)
    t_plan = time.perf_counter() - t0
    log(f"pattern {t_pattern:.1f}s (nnz {asm.nnz}), plan {t_plan:.1f}s "
        f"({plan.n_blocks} blocks, {int(plan.info[2])} records, lds {int(plan.info[4])} B, "
        f"edge records: {int(plan.edge_info[0]) * 4 if plan.has_edge else 16} B, "
        f"{int(plan.edge_info[1]) if plan.has_edge else 0} mirrored entries)")
    kappa = torch.tensor(kappa_np, dtype=torch.float32, device=dev)
    pd = {"kappa": kappa}
    E, N, nnz, nb = len(conn), len(pts), asm.nnz, 4
    # the same pattern built on the GPU from device connectivity, checked against the host one
    from tensormesh_amd.assemble.topology import device_element_slots, device_pattern
    conn_dev = torch.from_numpy(conn).to(dev)
    device_pattern(conn_dev, N)   # warm-up (code-object load)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    rp_d, ci_d, pwork = device_pattern(conn_dev, N)
    torch.cuda.synchronize(dev)
    t_pattern_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    slots_d = device_element_slots(conn_dev, rp_d, ci_d, pwork, N)
    torch.cuda.synchronize(dev)
    t_slots_dev = time.perf_counter() - t0
    pattern_exact = bool(np.array_equal(rp_d.cpu().numpy(), asm.topology.row_ptr)
                         and np.array_equal(ci_d.cpu().numpy(), asm.topology.col_idx))
    log(f"device pattern {t_pattern_dev * 1e3:.1f} ms (bit-exact vs host: {pattern_exact}), "
        f"element slots {t_slots_dev * 1e3:.1f} ms")
    del rp_d, ci_d, pwork, slots_d, conn_dev

    if args.calibrate_copy:
        src = torch.empty(args.calibrate_copy << 18, dtype=torch.float32, device=dev).uniform_()
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst
    for _ in range(args.warmup):
        K = asm(point_data=pd)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        K = asm(point_data=pd)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    t_step = ev0.elapsed_time(ev1) / 1e3 / args.steps
    log(f"assembly: {t_step * 1e6:.1f} us/step (events), {wall / args.steps * 1e6:.1f} us/step (wall)")
    b_asm = 4 * nb * E + 24 * N + 4 * N + 4 * (N + 1) + 4 * nnz + 4 * nnz
    achieved = b_asm / t_step / 1e9

    traffic = None
    tpath = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "asm_traffic.json")
    if os.path.exists(tpath) and n == 128:
        try:
            traffic = json.load(open(tpath)).get("hbm_bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    out = {
        "metric": "assembly Melem/s + HBM GB/s; time-to-1e-8 residual, learned-FPCG vs Jacobi-PCG",
        "value": round(E / t_step / 1e6, 2), "unit": "Melem/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(t_step * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 compute, f32 values", "data": "synthetic",
        "config": {"workload": f"3D variable-coefficient Poisson, {n}^3 Kuhn grid x 6 P1 tets, "
                               f"kappa=exp(GRF l=0.1 sigma=1 seed 0), kappa-Laplace assembly into CSR",
                   "elements": E, "nodes": N, "nnz": nnz, "values_stored": "f32",
                   "geometry_and_accumulation": "f64", "setup_pattern_s": round(t_pattern, 2),
                   "setup_plan_s": round(t_plan, 2), "setup_pattern_device_s": round(t_pattern_dev, 4),
                   "setup_element_slots_device_s": round(t_slots_dev, 4),
                   "pattern_device_bit_exact": pattern_exact},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": b_asm},
    }

    if not args.no_solve:
        b = tm.const_node_assembler(1.0).from_mesh(mesh).to(dev)()
        cond = tm.Condenser(mesh.boundary_mask.to(dev))
        Ki, fi = cond(K, b)
        csr = Ki.csr()
        ni, nnzi = Ki.shape[0], Ki.nnz
        # SpMV bandwidth
        x = torch.rand(ni, dtype=torch.float64, device=dev)
        from tensormesh_amd.sparse.ops import _spmv_raw
        for _ in range(3):
            _spmv_raw(csr, Ki.values, x)
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(20):
            _spmv_raw(csr, Ki.values, x)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_spmv = ev0.elapsed_time(ev1) / 1e3 / 20
        b_spmv = 4 * (ni + 1) + 8 * nnzi + 16 * ni
        import warnings
        with warnings.catch_warnings():                              # the warm-up stops early on purpose
            warnings.simplefilter("ignore")
            pcg_raw(csr, Ki.values, fi, rtol=1e-8, maxiter=50)       # warm-up (kernels, allocator)
        torch.cuda.synchronize()
        ev0.record(stream)
        xs, it, rel, ok = pcg_raw(csr, Ki.values, fi, rtol=1e-8, maxiter=20000)
        ev1.record(stream)
        torch.cuda.synchronize()
        t_solve = ev0.elapsed_time(ev1) / 1e3
        true_rel = ((fi.double() - Ki @ xs).norm() / fi.double().norm()).item()
        from tensormesh_amd.sparse.ops import col_delta16
        d16 = col_delta16(csr) is not None      # int16 column offsets: 6 instead of 8 bytes per entry
        b_iter = (6 if d16 else 8) * nnzi + 4 * (ni + 1) + 96 * ni
        out["solve"] = {"method": "jacobi-pcg", "columns": "int16 offsets" if d16 else "int32",
                        "n_inner": ni, "nnz_inner": nnzi, "iterations": it,
                        "converged": ok, "relres_recursive": rel, "relres_true_fp32_matrix": true_rel,
                        "time_to_1e-8_s": round(t_solve, 4), "ms_per_iter": round(t_solve / max(it, 1) * 1e3, 4),
                        "pcg_iter_GBps": round(b_iter / (t_solve / max(it, 1)) / 1e9, 1),
                        "spmv_us": round(t_spmv * 1e6, 2), "spmv_GBps": round(b_spmv / t_spmv / 1e9, 1)}
        log(f"solve: {it} its, {t_solve:.3f}s, true relres {true_rel:.2e}, spmv {t_spmv * 1e6:.1f} us")

    if not args.no_forms:
        out["forms"] = forms_leg(mesh, asm, pd, t_step, log)

    if not args.no_learned:
        seeds = list(range(5000, 5000 + args.learned_seeds))
        out["learned_fpcg"] = learned_leg(args.learned_n, seeds, log)
        lf = out["learned_fpcg"]
        if "median_speedup_vs_jacobi" in lf:
            log(f"learned FCG: {lf['seeds']} held-out seeds, speedup over Jacobi-PCG median "
                f"{lf['median_speedup_vs_jacobi']:.2f}x, 10th percentile {lf['p10_speedup_vs_jacobi']:.2f}x, "
                f"min {lf['min_speedup_vs_jacobi']:.2f}x; iterations {lf['median_fcg_iters']} vs "
                f"{lf['median_jacobi_iters']}, fallbacks {lf['fallbacks']}")

    if not args.no_batched:
        out["batched"] = batched_leg(args.batched_meshes, log)
        bt = out["batched"]
        log(f"batched: {bt['graph_s'] * 1e3:.1f} ms for all systems ({bt['speedup_vs_loop']}x the loop), "
            f"iterations median {bt['median_iters']}, max {bt['max_iters']}")

    if not args.no_hyper:
        out["hyperelastic"] = hyperelastic_leg(args.hyper_n, log)

    if not args.no_refpoisson:
        out["reference_poisson_3d"] = reference_poisson_leg((47, 114, 151), log)
        out["reference_elasticity_3d"] = reference_elasticity_leg((30, 58, 73), log)
        out["convection_diffusion_2d"] = convection_diffusion_leg(1024, log)

    if not args.no_cpu:
        log(f"cpu baseline on {args.cpu_n}^3 ...")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n)
        cb = out["cpu_baseline"]
        log(f"cpu baseline ({cb['cores']} threads, {args.cpu_n}^3): pattern {cb['all_threads']['pattern_s']} s, "
            f"assembly {cb['all_threads']['assembly_s']} s = {cb['value']} Melem/s, condense "
            f"{cb['all_threads']['condense_s']} s, PCG {cb['all_threads']['pcg_iterations']} its in "
            f"{cb['all_threads'].get('solve_to_1e-8_s')} s; 1 thread ({cb['one_thread']['grid']}): "
            f"{cb['one_thread']['assembly_melem_per_s']} Melem/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
