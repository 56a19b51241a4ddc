"""Benchmark of the hot path: assembly + equilibrated-flux estimator (DOF/s) and
MINRES (s/solve), mixed precision vs FP64, on one MI355X.

A step is one pass of the hot path over the workload, i.e. one inner
iteration of the adaptive loop: assemble K, f, g -> MINRES to a true relative
residual of 1e-10 -> estimator: slab reconstruction (every slab's
A_slab lt = f_slab - (B^T u)_slab solved to 1e-13), patch problems in FP32 with
two FP64 refinement steps, per-cell indicators.  `value` is (n_X + n_Y) divided
by the mean time of assembly + estimator per step (device time of the assembly,
patch and indicator kernels; host wall time of the slab solve, which
synchronises per iteration); MINRES seconds per solve are reported beside it.
Inputs are resident in HBM when the timed region starts (the mesh, pattern and
coefficients are uploaded once).

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c3g|c4|c5|adapt1d]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "assembly+estimator DOF/s & MINRES s/solve; mixed vs FP64 at 1e-10 rel err"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per launch (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, the factors a
# plain copy kernel of known size reads) from the rocprofv3 --pmc runs of
# `bench.py --profile`, summarised in profiles/.
# HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (calibrated on a copy kernel)
MEASURED_TRAFFIC = {"c2": {"spmv": 2.11e9, "mg_step": 9.41e7, "solve": 2.49e9, "factor": 0.28e9}}


def workload(name):
    import doubleadapparabolicfem_amd as apf
    from doubleadapparabolicfem_amd import meshes

    if name == "c2":
        m = meshes.uniform_mesh(2, 5)
        # MINRES preconditioned with FP32 V-cycles: 54 iterations in 0.049 s against 45 in 0.053 s
        # with FP64 ones (MINRES and its true-residual acceptance stay FP64)
        return dict(mesh=m, p=2, coarse=meshes.uniform_hierarchy(2, 5), coeff=meshes.smooth_initial(2),
                    precond=apf.APF_FP32_REFINE,
                    anchor=dict(lambda_y2=2.42088e-05, eta2=2.42111e-05, flux2=6.64391e-09),
                    desc="2+1D heat eq smooth_initial (M=I, ell=0, u0=sin(pi x)sin(pi y)), p=2, uniform "
                         "union-jack 64x64 squares x 32 intervals = 262144 prisms")
    if name == "c1":
        m = meshes.uniform_mesh(1, 3)
        return dict(mesh=m, p=1, coarse=meshes.uniform_hierarchy(1, 3), coeff=meshes.smooth_initial(1),
                    anchor=dict(lambda_y2=0.00767078, eta2=0.00779326, flux2=0.000766702),
                    desc="1+1D heat eq smooth_initial, p=1, uniform 16x16 space-time rectangles")
    if name == "c4":
        m = meshes.union_jack_mesh(64, np.linspace(0.0, 1.0, 17))
        coeff = meshes.smooth_initial(2)
        coeff = apf.Coefficients(meshes.checkerboard_kappa(m, 8, 4), 0.0, coeff.u0)
        return dict(mesh=m, p=3, coarse=[meshes.union_jack_mesh(n, [0.0, 1.0]) for n in (32, 16, 8, 4, 2)],
                    coeff=coeff, anchor=None, precond=apf.APF_FP32_REFINE,
                    desc="2+1D p=3, union-jack 64x64 x 16 intervals, kappa piecewise constant 10^U(-1,1) "
                         "on an 8x8 checkerboard")
    if name == "c5":
        nx, kap, coef = meshes.segment_batch_problems(256, 32, 5000, multiple=64)
        m = meshes.segment_batch_mesh(nx, np.linspace(0.0, 1.0, 33))
        return dict(mesh=m, p=1, coarse=meshes.segment_batch_hierarchy(nx, 6),
                    coeff=meshes.segment_batch_coefficients(m, nx, kap, coef), anchor=None,
                    desc="256 independent 1+1D heat problems in one system (disjoint segments), p=1, "
                         "n_x log-uniform for 1e3-1e5 X DOFs rounded to multiples of 64, 32 shared time "
                         "intervals, kappa 10^U(-1,1), u0 = 5 random sine modes, seeds 5000+i")
    raise SystemExit(f"unknown config {name}")


def run_c3(args):
    """Fan-patch workload: q = 8 squares per side, side splits uniform in 1..16
    (seed 3), so the 64 fan centres have K in 4..64 cells; time graded toward
    t = 0 (intervals 2^-10 .. 2^-2, then 1/4); u0 = indicator of [0.25,0.75]^2.
    The same mesh at p = 1, 2, 3 (every patch of each degree); value is the sum
    of DOFs over the sum of assembly + mixed estimator device times."""
    import doubleadapparabolicfem_amd as apf
    from doubleadapparabolicfem_amd import meshes

    m = meshes.fan_mesh(8, meshes.graded_times(2, 10), seed=3)
    coeff = meshes.indicator_initial()
    ctx = apf.Context(0)
    per_p, tot_dofs, tot_ms, wall = {}, 0, 0.0, 0.0
    for p in (1, 2, 3):
        sp = apf.Space(m, p, coarse=meshes.fan_hierarchy(8))
        sysm = apf.System(ctx, sp)
        nf, ns_, nc, bd = sp.patch_sizes()
        order = nf + ns_ + (bd == 0)
        rec = []
        # one untimed step whatever --warmup says: a system's first estimate also chooses its
        # tiles and, per size bucket, the factor precision of the mixed path
        warm = max(args.warmup, 1)
        for i in range(warm + args.steps):
            t0 = time.perf_counter()
            sysm.assemble(coeff)
            t_asm = sysm.timings()["assembly_ms"]
            t1 = time.perf_counter()
            x, st = sysm.minres(tol=1e-10, max_it=20000, check_interval=20, download=False)
            t_minres = time.perf_counter() - t1
            est = sysm.estimate(apf.APF_FP32_REFINE, 2, per_cell=False, slab=apf.APF_SLAB_SOLVE)
            if i >= warm:
                wall += time.perf_counter() - t0
                rec.append((t_asm, t_minres, st["iterations"],
                            est["seconds_total"] * 1e3,
                            sysm.phase_timings(), est))
        sysm.estimate(apf.APF_FP64, 0, per_cell=False, slab=apf.APF_SLAB_SOLVE)  # untimed: tunes the tiles
        est64 = sysm.estimate(apf.APF_FP64, 0, per_cell=False, slab=apf.APF_SLAB_SOLVE)
        ph64 = sysm.phase_timings()
        t_asm = float(np.mean([r[0] for r in rec]))
        t_est = float(np.mean([r[3] for r in rec]))
        ph = {k: float(np.mean([r[4][k] for r in rec])) for k in rec[0][4]}
        est = rec[-1][5]
        n_dofs = sp.nX + sp.nY
        tot_dofs += n_dofs
        tot_ms += t_asm + t_est
        per_p[f"p{p}"] = {
            "n_X": sp.nX, "n_Y": sp.nY, "patches_per_slab": int(order.size), "max_patch_cells": int(nc.max()),
            "max_patch_order": int(order.max()), "assembly_ms": t_asm, "estimator_ms_mixed": t_est,
            "estimator_ms_fp64": est64["seconds_total"] * 1e3,
            "slab_solve_ms": float(np.mean([r[5]["seconds_slab"] for r in rec])) * 1e3,
            "phases_mixed_ms": ph, "phases_fp64_ms": ph64,
            "minres_s_per_solve": float(np.mean([r[1] for r in rec])), "minres_iterations": int(rec[-1][2]),
            "minres_preconditioner": "multigrid over criss-cross 8x8, union-jack 8x8 .. 2x2",
            "mixed_vs_fp64_rel_err_flux2": abs(est["flux2"] - est64["flux2"]) / abs(est64["flux2"]),
            "max_patch_residual": est["max_patch_residual"], "extra_refinements": est["extra_refinements"],
            "roofline_patch_phases_mixed": phase_rooflines(sp, ph, 2, True)}
    dom_p = max(per_p, key=lambda k: per_p[k]["estimator_ms_mixed"])
    roof = per_p[dom_p]["roofline_patch_phases_mixed"]
    dom = max(("factor", "solve"), key=lambda k: roof[k]["avg_launch_ms"])
    out = {
        "metric": METRIC, "value": tot_dofs / (tot_ms * 1e-3), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": warm, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "2+1D fan-patch mesh: 8x8 squares, side splits U{1..16} (seed 3), fan centres "
                               "with K in 4..64 cells; 12 intervals graded toward t=0 (2^-10 .. 1/4); "
                               "u0 = indicator of [0.25,0.75]^2; p = 1, 2, 3 in turn",
                   "space_cells": int(m.cells.shape[0]), "intervals": int(m.times.size - 1)},
        "roofline": dict(roof[dom], degree=dom_p, traffic=None),
        "per_degree": per_p,
    }
    print(json.dumps(out))


def run_c3_graded(args):
    """The graded half of C3: the reference's split policy applied 5 times (generations 2..10)
    to the cells with t0 = 0 of a 2x2 union-jack mesh on [0, 1]: time layers [1/2, 1] at
    4x4 squares down to [0, 1/32] and [1/32, 1/16] at 64x64, hanging X nodes at every layer
    interface; u0 = indicator of [0.25, 0.75]^2; p = 1, 2, 3 in turn.  A step: element-
    quadrature assembly (apf_general_assemble) and the estimator over the slab groups (slab
    solve, patches, indicators) on the solution of one MINRES solve per degree (caller's-CSR
    path, Chebyshev blocks; timed once, outside the steps).  value = sum of DOFs / sum of
    assembly + estimator times."""
    import doubleadapparabolicfem_amd as apf
    from doubleadapparabolicfem_amd import layered as L, meshes

    xl = L.graded_layers(2, 5)
    u0 = meshes.indicator_initial().u0
    ctx = apf.Context(0)
    per_p, tot_dofs, tot_ms, wall = {}, 0, 0.0, 0.0
    warm = max(args.warmup, 1)
    for p in (1, 2, 3):
        t0 = time.perf_counter()
        lm = L.layered_mesh(2, p, xl)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        gs = apf.GeneralSystem(ctx, lm)
        t_create = time.perf_counter() - t0
        groups = []
        for g in L.slab_groups(lm):
            coarse, k = [], g["n"] // 2
            while k >= 2:
                coarse.append(meshes.union_jack_mesh(k, [0.0, 1.0]))
                k //= 2
            groups.append((meshes.union_jack_mesh(g["n"], g["times"]), coarse, g["first_y_cell"]))
        ge = apf.GeneralEstimator(gs, groups)
        ge.assemble(u0)
        K = gs.matrix()
        gs.assemble(u0)
        b = gs.rhs()
        print(f"c3g p={p}: {gs.n} unknowns, nnz {gs.nnz}; MINRES ...", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        H = gs.graph_norm()  # Hhat = -C + B_dt diag(A_y)^-1 B_dt^T + A_x on the device
        t_hhat = time.perf_counter() - t1
        x, st = K.minres(b, gs.nY, hhat=H, tol=1e-10, max_it=40000, check_interval=50, degree=16,
                         precond=apf.APF_FP64)
        t_minres = time.perf_counter() - t1
        del H
        print(f"c3g p={p}: MINRES {st['iterations']} iterations, {t_minres:.2f} s", file=sys.stderr, flush=True)
        rec = []
        for i in range(warm + args.steps):
            t0 = time.perf_counter()
            t_asm = gs.assemble(u0)
            est = ge.estimate(x, apf.APF_FP32_REFINE, 2, per_cell=False)
            if i >= warm:
                wall += time.perf_counter() - t0
                rec.append((t_asm, est))
        print(f"c3g p={p}: steps done", file=sys.stderr, flush=True)
        t_asm = float(np.mean([r[0] for r in rec]))
        t_est = float(np.mean([r[1]["seconds_total"] * 1e3 for r in rec]))
        est = rec[-1][1]
        n_dofs = gs.n
        tot_dofs += n_dofs
        tot_ms += t_asm + t_est
        hanging = int(np.sum(np.diff(lm.X.restr_ptr) > 1))
        # element assembly work per Y cell (the reference's q-loop): 10 n^2 n_q + 2 n n_q flops,
        # n = (p+1)^2 (p+2) / 2 prism basis functions, n_q = (2p+2)^2 (2p+1) points; the kernel
        # evaluates the same sums in sum-factorised form, so it executes fewer
        nb, nq = (p + 1) ** 2 * (p + 2) // 2, (2 * p + 2) ** 2 * (2 * p + 1)
        flops = float(lm.Y.n_cells) * (10 * nb * nb * nq + 2 * nb * nq)
        roof_asm = {"bound": "mfma", "kernel": "k_gen_cells + k_gen_bottom (FP64)",
                    "achieved": flops / (t_asm * 1e-3) / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                    "frac": flops / (t_asm * 1e-3) / 1e12 / 78.6, "algorithmic_flops_per_launch": flops,
                    "work_unit": "10 n^2 n_q + 2 n n_q per Y cell"}
        per_p[f"p{p}"] = {
            "n_X": lm.X.n_dofs, "n_Y": lm.Y.n_dofs, "nnz_K": gs.nnz, "prisms": int(lm.Y.n_cells),
            "hanging_x_rows": hanging, "slab_groups": len(groups),
            "assembly_ms": t_asm, "assembly_kernel": "k_gen_cells (element quadrature) + k_gen_bottom",
            "roofline_assembly": roof_asm,
            "estimator_ms_mixed": t_est,
            "slab_solve_ms": float(np.mean([r[1]["seconds_slab"] for r in rec])) * 1e3,
            "minres_s_one_solve": t_minres, "minres_iterations": int(st["iterations"]),
            "minres_true_rel_residual": st["true_relative_residual"],
            "minres_hhat_setup_s": t_hhat,
            "minres_preconditioner": ("Chebyshev blocks of degree 16, FP64: Jacobi on A_y, and Jacobi on the "
                                      "graph norm Hhat formed on the device (apf_general_graph_norm; the solve "
                                      "time includes forming it)"),
            "host_mesh_build_s": t_build, "host_pattern_and_upload_s": t_create,
            "estimator": {k: est[k] for k in ("lambda_y2", "eta2", "flux2", "recon2", "y_est2", "max_patch_residual")}}
    out = {
        "metric": METRIC, "value": tot_dofs / (tot_ms * 1e-3), "unit": "DOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": warm, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "2+1D graded toward t=0 by 5 double rounds of the split policy from a 2x2 "
                               "union-jack mesh: 6 time layers (1/2..1 at 4x4 squares .. 0..1/32 at 64x64), "
                               "hanging X nodes at the layer interfaces; u0 = indicator of [0.25,0.75]^2; "
                               "p = 1, 2, 3 in turn"},
        "per_degree": per_p,
    }
    print(json.dumps(out))


def run_adaptive_1d(args):
    """The 1+1D adaptive driver with its inner Y loop (the published adaptive_standard
    smooth_initial runs: rho = 1, theta_x = theta_y = 0.4) from the initial mesh through
    `--steps` outer iterations (default 22) at p = 1, 2, 3, timed as a whole run: host
    refinement and numbering, GPU assembly, Hhat and MINRES, X indicators, Y estimator, and
    (p <= 2) G estimator on every inner iteration.  Reports DOF/s = sum over solves of
    (n_X + n_Y) / wall time, with the per-phase split and the last iteration's sizes; MINRES
    starts from the previous solution (runs) and from zero (runs_zero_start)."""
    import doubleadapparabolicfem_amd as apf
    from doubleadapparabolicfem_amd import adaptive1d as A

    ctx = apf.Context(0)
    outer = args.steps if args.steps != 5 else 22
    result = {}
    degrees = [int(v) for v in args.adapt_p.split(",")]
    starts = [v.strip() for v in args.adapt_starts.split(",")]
    for key, initial_guess in (("runs", True), ("runs_zero_start", False)):
        if ("guess" if initial_guess else "zero") not in starts:
            continue
        out = {}
        for p in degrees:
            run = A.AdaptiveX1D(ctx, p=p, tol=1e-10, max_it=200000, inner_y_loop=True, with_g_estimator=p <= 2,
                                initial_guess=initial_guess)
            t0 = time.perf_counter()
            solves = dofs = 0
            its = 0
            ph = dict(minres_s=0.0, estimator_s=0.0, g_s=0.0)
            for k in range(outer):
                rec = run.step()
                for r in rec["inner"]:
                    solves += 1
                    dofs += r["x_true_dofs"] + r["y_true_dofs"]
                    its += r["minres_iterations"]
                    ph["minres_s"] += r["minres_seconds"]
                    ph["estimator_s"] += r["estimator_seconds"]
                    ph["g_s"] += r.get("g_seconds", 0.0)
                print(f"adapt1d {key} p={p} outer={k} X={rec['x_true_dofs']} Y={rec['y_true_dofs']} "
                      f"inner={len(rec['inner'])}", file=sys.stderr, flush=True)
            wall = time.perf_counter() - t0
            out[f"p{p}"] = dict(outer_iterations=outer, solves=solves, wall_s=wall, dof_per_s=dofs / wall,
                                minres_iterations=its, last_x_true_dofs=rec["x_true_dofs"],
                                last_y_true_dofs=rec["y_true_dofs"], eta_squared=rec["eta_squared"],
                                y_estimator_squared=rec["y_estimator_squared"],
                                other_s=wall - sum(ph.values()), **ph)
        result[key] = out
    first = next(iter(result.values()))
    print(json.dumps({"metric": "1+1D adaptive driver DOF/s (whole run, inner Y loop)",
                      "value": first.get("p2", next(iter(first.values())))["dof_per_s"], "unit": "DOF/s",
                      "n_gpus": 1, "steps": outer,
                      "warmup": 0, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
                      "data": "synthetic",
                      "config": {"workload": "1+1D heat eq smooth_initial, adaptive_standard driver (rho = 1, "
                                             "theta_x = theta_y = 0.4), p = 1, 2, 3, from the initial mesh"},
                      **result}))


def patch_work(sp, steps):
    """Work of one estimate's patch phases.  n = order of a spatial vertex patch problem
    of one Legendre time mode (flux + scalar + mean-zero multiplier); on a tensor mesh
    its matrix is shared by every slab and mode, so each patch is factored once and
    applied to cols = slabs x modes right-hand sides.
    Algorithmic (the per-unit work of the roofline): factor (2/3) n^3 (an LU of the
    mode matrix), solve 2 n^2 cols (one application of the inverse per right-hand side).
    Executed: the factor is a Gauss-Jordan inversion, 2 n^3; the solve applies the
    inverse (1 + steps) times and forms (1 + steps) FP64 residuals, 2 n^2 cols each;
    the products beyond the first are refinement overhead."""
    nf, ns_, _, bd = sp.patch_sizes()
    n = (nf + ns_ + (bd == 0)).astype(np.float64)
    cols = sp.n_intervals * (sp.p + 1)
    prod = float(np.sum(2.0 * n ** 2 * cols))
    return {"factor_alg": float(np.sum(2.0 / 3.0 * n ** 3)), "factor_exec": float(np.sum(2.0 * n ** 3)),
            "solve_alg": prod, "solve_ft": (1 + steps) * prod, "solve_fp64": (1 + steps) * prod}


def phase_rooflines(sp, ph, steps, mixed):
    """Roofline entries of the patch kernels from their measured times, on the
    algorithmic work of patch_work against the peak of the inverse's precision
    (FP32 157.3 TFLOP/s in the mixed path, FP64 78.6 in the plain one)."""
    w = patch_work(sp, steps)
    ft_peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    out = {}
    t = ph["factor_ms"] * 1e-3
    ach = w["factor_alg"] / t / 1e12
    name = ("k_patch_factor_blocked (Gauss-Jordan: FP%d VALU panels, MFMA rank-16 updates)"
            % (32 if mixed else 64))
    out["factor"] = {"bound": "mfma", "kernel": name, "achieved": ach, "peak": ft_peak, "unit": "TFLOP/s",
                     "frac": ach / ft_peak, "algorithmic_flops_per_launch": w["factor_alg"],
                     "work_unit": "(2/3) n^3 per patch mode matrix",
                     "executed_flops_per_launch": w["factor_exec"], "avg_launch_ms": ph["factor_ms"]}
    t = ph["solve_ms"] * 1e-3
    ach = w["solve_alg"] / t / 1e12
    out["solve"] = {"bound": "mfma", "kernel": "k_patch_solve (MFMA 16x16x4: " + ("FP32" if mixed else "FP64")
                    + " products, FP64 residuals)", "achieved": ach, "peak": ft_peak, "unit": "TFLOP/s",
                    "frac": ach / ft_peak, "algorithmic_flops_per_launch": w["solve_alg"],
                    "work_unit": "2 n^2 per patch and right-hand side (one inverse application)",
                    "executed_flops_per_launch": w["solve_ft"] + w["solve_fp64"],
                    "overhead": "%d refinement rounds: %d more products in the inverse's precision and %d FP64 "
                                "residual products" % (steps, steps, 1 + steps),
                    "avg_launch_ms": ph["solve_ms"]}
    out["rhs"] = {"kernel": "k_patch_rhs", "avg_launch_ms": ph["rhs_ms"]}
    return out


def g_estimator_run(apf, ctx, W, sysm, y_est2):
    """G-estimator of the final solution, outside the timed steps: the degree-(p+1)
    system on the same mesh is assembled once; its first call also sets up the
    multigrid and the interpolation, the second is the one reported."""
    if W["mesh"].dim == 2 and W["p"] + 1 > 3:
        return {"skipped": "degree p+1 = %d in 2D is beyond the degrees this build assembles" % (W["p"] + 1)}
    sq = apf.Space(W["mesh"], W["p"] + 1, coarse=W["coarse"])
    gsys = apf.System(ctx, sq)
    gsys.assemble(W["coeff"])
    t_asm = gsys.timings()["assembly_ms"]
    first = sysm.g_estimate(gsys)
    g = sysm.g_estimate(gsys)
    return {"g_true_dofs": g["g_true_dofs"], "nnz_K_degree_p_plus_1": sq.nnz,
            "lambda_difference_squared": g["lambda_difference_squared"],
            "g_over_y_estimator": float(np.sqrt(g["lambda_difference_squared"] / y_est2)) if y_est2 > 0 else None,
            "seconds": g["seconds"], "first_call_seconds": first["seconds"],
            "solver_iterations": g["solver_iterations"], "solver_relative_residual": g["solver_relative_residual"],
            "assembly_ms_degree_p_plus_1": t_asm, "repeat_bit_identical": g["raw"] == first["raw"]}


def cpu_baseline(apf, ctx, sysm, sp, gpu):
    """FP64 C++/OpenMP restatement (oracle/cpu_baseline.cpp) of the C2 step's CPU work on
    this host's cores, all of them and the reference's 8-thread policy for p = 2
    (README_SNELLIUS.md:53-61): the element assembly q-loop + scatter on the full
    workload; the coupled (slab, vertex) patch LU solves with the 1e-10 acceptance on
    one slab's patches (the estimate is that times the slab count); MINRES iterations
    with the Chebyshev-Jacobi block preconditioner on the GPU-assembled K (the GPU runs
    the same algorithm, apf_csr_minres_solve, for the per-iteration ratio)."""
    from oracle import cases, cpu_baseline as CB

    host = CB.host_info()
    t_all = int(host["OMP_NUM_THREADS"] or host["nproc"])
    t0 = time.perf_counter()
    stm = cases.uniform_mesh(2, 5)
    tc = CB.TensorCase(stm, 2)
    orders = tc.patch_orders()
    vals, b = sysm.download()
    rowptr, col = sysm.download_pattern()
    setup_s = time.perf_counter() - t0
    m = stm.n_intervals
    n_dofs = sp.nX + sp.nY
    iters = 10
    phases = {}
    for th in sorted({t_all, 8}, reverse=True):
        t_asm = tc.assemble(th, 1)
        t_slab_patches, stat = CB.patch_solves(orders, th)
        t_mr, it, _ = CB.minres(rowptr, col, vals, b, sp.nY, th, iters)
        phases[th] = {"threads": th, "assembly_s": t_asm, "patch_solves_one_slab_s": t_slab_patches,
                      "patch_solves_s": t_slab_patches * m, "patch_max_acceptance_stat": stat,
                      "minres_s_per_iteration": t_mr / max(it, 1),
                      "dof_per_s": n_dofs / (t_asm + t_slab_patches * m)}
    # the CPU assembly against the GPU values (every entry outside the t = 0 block C,
    # which the CPU q-loop does not assemble)
    nY = sp.nY
    rows = np.repeat(np.arange(sp.n, dtype=np.int64), np.diff(rowptr))
    keep = ~((rows >= nY) & (col >= nY))
    asm_err = float(np.max(np.abs(tc.val[keep] - vals[keep])) / np.max(np.abs(vals)))
    # the same MINRES algorithm on the GPU, for the per-iteration ratio
    K = apf.CsrMatrix(ctx, outer=rowptr, inner=col, values=vals) if rowptr[-1] < 2 ** 31 else None
    gpu_it_s = None
    if K is not None:
        _, st = K.minres(b, nY, tol=1e-30, max_it=iters, check_interval=iters, degree=8, raise_on_fail=False)
        gpu_it_s = st["seconds"] / max(st["iterations"], 1)
    best = phases[t_all]
    return {"value": best["dof_per_s"], "unit": "DOF/s", "cores": t_all, "kind": "port",
            "sample": "FP64 C++/OpenMP restatement on the host: element assembly (q-loop + scatter) of the whole C2 "
                      "mesh; coupled (slab, vertex) patch LU + solve + 1e-10 acceptance at the C2 orders "
                      f"(N = {sorted(set(orders.tolist()))}) for one slab's {len(orders)} patches, times {m} slabs; "
                      f"{iters} MINRES iterations on the GPU-assembled K; patch matrices synthetic (LU work depends "
                      "on N only); value = DOFs / (assembly + patch solves)",
            "host": host, "phases": {str(k): v for k, v in phases.items()},
            "assembly_max_rel_diff_vs_gpu": asm_err,
            "gpu_same_minres_algorithm_s_per_iteration": gpu_it_s,
            "gpu_multigrid_minres_s_per_iteration": gpu["minres_s_per_solve"] / max(gpu["minres_iterations"], 1),
            "setup_s_untimed": setup_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-g", action="store_true", help="skip the G-estimator block (untimed, outside `value`)")
    ap.add_argument("--adapt-p", default="1,2,3", help="adapt1d: the degrees to run (comma-separated)")
    ap.add_argument("--adapt-starts", default="guess,zero",
                    help="adapt1d: MINRES starts to run (guess: previous solution, zero)")
    ap.add_argument("--profile", action="store_true",
                    help="one assembly, MINRES solve, two mixed and two FP64 estimates (slab solve), 20 launches each of "
                         "the CSR SpMV, the matrix-free product and the multigrid step, and a 1 GiB "
                         "device copy (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only "
                                   "(time-slab row partition), --gpus must be 1"}))
        sys.exit(2)

    import doubleadapparabolicfem_amd as apf

    if args.config == "c3":
        run_c3(args)
        return
    if args.config == "c3g":
        run_c3_graded(args)
        return
    if args.config == "adapt1d":
        run_adaptive_1d(args)
        return
    W = workload(args.config)
    sp = apf.Space(W["mesh"], W["p"], coarse=W["coarse"])
    ctx = apf.Context(0)
    sysm = apf.System(ctx, sp)
    n_dofs = sp.nX + sp.nY

    if args.profile:
        # counter runs: one pass of each hot-path kernel, then the calibration copy
        sysm.assemble(W["coeff"])
        sysm.minres(tol=1e-10, max_it=5000, check_interval=10, precond=W.get("precond", apf.APF_FP64))
        # each estimate twice: a system's first call in a precision also times its tile
        # candidates, the second runs the chosen tiles only
        for _ in range(2):
            sysm.estimate(apf.APF_FP32_REFINE, 2, per_cell=False, slab=apf.APF_SLAB_SOLVE)
        for _ in range(2):
            sysm.estimate(apf.APF_FP64, 0, per_cell=False, slab=apf.APF_SLAB_SOLVE)
        ms = sysm.bench_spmv(20)
        sysm.bench_kron_apply(20)
        sysm.bench_mg_step(20, W.get("precond", apf.APF_FP64))
        gbs = ctx.bench_stream(1 << 30, 3)
        print(json.dumps({"spmv_ms": ms, "spmv_algorithmic_bytes": 12 * sp.nnz + 20 * sp.n, "nnz": sp.nnz,
                          "n": sp.n, "copy_kernel_bytes_read": 1 << 30, "copy_kernel_bytes_written": 1 << 30,
                          "copy_GBs": gbs, "patch_work": patch_work(sp, 2)}))
        return

    def step(prec=apf.APF_FP32_REFINE):
        sysm.assemble(W["coeff"])
        t_asm = sysm.timings()["assembly_ms"]
        t0 = time.perf_counter()
        x, st = sysm.minres(tol=1e-10, max_it=5000, check_interval=10, precond=W.get("precond", apf.APF_FP64),
                            download=False)
        t_minres = time.perf_counter() - t0
        est = sysm.estimate(prec, 2, per_cell=False, slab=apf.APF_SLAB_SOLVE)
        # device span of the whole estimate (slab solve, patches, indicators; the factor phase
        # runs beside the slab solve after the first estimate)
        t_est = est["seconds_total"] * 1e3
        return t_asm, t_minres, st, t_est, est, sysm.phase_timings()

    # at least one untimed step whatever --warmup says: a system's first estimate also chooses
    # its tiles and, per size bucket, the factor precision of the mixed path
    for _ in range(max(args.warmup, 1)):
        step()
    rec = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rec.append(step())
    wall = time.perf_counter() - t0
    t_asm = np.mean([r[0] for r in rec])
    t_minres = np.mean([r[1] for r in rec])
    t_est = np.mean([r[3] for r in rec])
    its = int(np.mean([r[2]["iterations"] for r in rec]))
    t_slab = float(np.mean([r[4]["seconds_slab"] for r in rec])) * 1e3
    est_mixed = rec[-1][4]
    # FP64 estimator path on the same solution for the mixed-vs-FP64 error and speed ratio (one
    # untimed call first: a system's first estimate in a precision times its tile candidates)
    sysm.estimate(apf.APF_FP64, 0, per_cell=False, slab=apf.APF_SLAB_SOLVE)
    est64 = sysm.estimate(apf.APF_FP64, 0, per_cell=False, slab=apf.APF_SLAB_SOLVE)
    t_est64 = est64["seconds_total"] * 1e3
    rel = {k: abs(est_mixed[k] - est64[k]) / abs(est64[k]) for k in ("flux2", "eta2", "lambda_y2", "y_est2")}
    value = n_dofs / ((t_asm + t_est) * 1e-3)

    # rooflines of the patch phases (the dominant kernels of `value`)
    ph_mixed = {k: float(np.mean([r[5][k] for r in rec])) for k in rec[0][5]}
    ph64 = sysm.phase_timings()
    roof = phase_rooflines(sp, ph_mixed, 2, True)
    roof64 = phase_rooflines(sp, ph64, 0, False)
    # HBM roofline of the CSR SpMV (MINRES true residuals, the general path's operator):
    # algorithmic bytes 12 nnz + 20 n
    spmv_ms = sysm.bench_spmv(50)
    stream_gbs = ctx.bench_stream(1 << 30, 10)
    spmv_bytes = 12 * sp.nnz + 20 * sp.n
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
    # the MINRES Lanczos product, matrix-free: x read and y written once, the spatial
    # pattern (int column, K and M values) and row pointer
    kron_ms = sysm.bench_kron_apply(50)
    nnz_sp, nf_sp = sp.nnz_spatial, sp.nf
    kron_bytes = 16 * sp.n + 20 * nnz_sp + 4 * nf_sp
    kron_gbs = kron_bytes / (kron_ms * 1e-3) / 1e9
    # the dominant kernel of a step: the multigrid's fine-level Chebyshev step, in the
    # precision of the MINRES preconditioner
    mg_prec = W.get("precond", apf.APF_FP64)
    mg_ms, mg_bytes = sysm.bench_mg_step(50, mg_prec)
    mg_gbs = mg_bytes / (mg_ms * 1e-3) / 1e9
    # assembly: the value array (8 nnz bytes) is its only large stream; a plain
    # store-only kernel over the same bytes gives the write ceiling beside the peak
    asm_gbs = 8 * sp.nnz / (t_asm * 1e-3) / 1e9
    fill_gbs = ctx.bench_store(8 * sp.nnz, 10)
    traffic = MEASURED_TRAFFIC.get(args.config, {})
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(args.warmup, 1),
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "dtype_note": "FP64 assembly, MINRES (matrix-free Lanczos products, CSR true residuals, inner products; "
                      "FP32 multigrid V-cycles as preconditioner) and slab solve (FP32 V-cycles in its "
                      "preconditioner); patch inverses and products FP32 with FP64 residual refinement",
        "data": "synthetic",
        "config": {"workload": W["desc"], "p": W["p"], "n_X": sp.nX, "n_Y": sp.nY, "nnz_K": sp.nnz,
                   "space_time_cells": sp.n_space_cells * sp.n_intervals,
                   "vertex_patches": sp.n_patches * sp.n_intervals},
        # the kernel that dominates a step: the multigrid's fine-level Chebyshev step (16 launches
        # per MINRES iteration over 161 or 65 columns, plus the slab solve's)
        "roofline": {"bound": "hbm", "kernel": "k_mg_cheb_fused (fine-level Chebyshev step of the multigrid, "
                                               + ("FP32" if mg_prec == apf.APF_FP32_REFINE else "FP64")
                                               + ", all columns)",
                     "achieved": mg_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": mg_gbs / HBM_PEAK_GBS,
                     "traffic": traffic.get("mg_step"), "algorithmic_bytes_per_launch": mg_bytes,
                     "avg_launch_ms": mg_ms, "measured_copy_GBs": stream_gbs,
                     "frac_of_measured_copy": mg_gbs / stream_gbs},
        "roofline_csr_spmv": {"bound": "hbm", "kernel": "k_spmv (CSR SpMV, FP64 values, int32 columns: MINRES "
                                                        "true residuals, the caller's-CSR and general paths)",
                              "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "traffic": traffic.get("spmv"),
                              "algorithmic_bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms,
                              "frac_of_measured_copy": achieved / stream_gbs},
        "roofline_minres_product": {"bound": "hbm", "kernel": "k_kron_apply (K x from the Kronecker factors)",
                                    "achieved": kron_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": kron_gbs / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": kron_bytes,
                                    "avg_launch_ms": kron_ms, "csr_spmv_ms": spmv_ms},
        "roofline_patch_phases_mixed": roof,
        "roofline_patch_phases_fp64": roof64,
        "roofline_assembly": {"bound": "hbm", "kernel": "assembly (k_spatial_elements + k_fill_kron + k_fill_rhs)",
                              "achieved": asm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": asm_gbs / HBM_PEAK_GBS,
                              "algorithmic_bytes_per_launch": 8 * sp.nnz, "avg_launch_ms": t_asm,
                              "measured_store_fill_GBs": fill_gbs, "frac_of_measured_fill": asm_gbs / fill_gbs},
        "minres_s_per_solve": t_minres,
        "minres_iterations": its,
        "minres_mg_setup_s": sysm.mg_setup_seconds(),
        "minres_true_rel_residual": rec[-1][2]["true_relative_residual"],
        "assembly_ms": t_asm,
        "estimator_ms_mixed": t_est,
        "slab_solve_ms": t_slab,
        "slab_solve_iterations": est_mixed["slab_iterations"],
        "slab_solve_relative_residual": est_mixed["slab_relative_residual"],
        "value_without_slab_solve": n_dofs / ((t_asm + t_est - t_slab) * 1e-3),
        "estimator_ms_fp64": t_est64,
        "patch_kernel_ms_mixed": est_mixed["seconds_patches"] * 1e3,
        "patch_kernel_ms_fp64": est64["seconds_patches"] * 1e3,
        "mixed_vs_fp64_rel_err": rel,
        "max_patch_residual": est_mixed["max_patch_residual"],
        "estimator": {k: est_mixed[k] for k in ("lambda_y2", "eta2", "flux2", "y_est2", "div_residual_l2")},
    }
    if W["anchor"]:
        out["published_anchor_rel_err"] = {k: abs(est_mixed[k] - v) / v for k, v in W["anchor"].items()}
    if not args.no_g:
        out["g_estimator"] = g_estimator_run(apf, ctx, W, sysm, est_mixed["y_est2"])
    if not args.no_cpu_baseline and args.config == "c2":
        out["cpu_baseline"] = cpu_baseline(apf, ctx, sysm, sp, out)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
