"""Benchmark of the hot path: assembly + equilibrated-flux estimator (DOF/s) and
MINRES (s/solve), mixed precision vs FP64, on one MI355X.

A step is one pass of the hot path over the workload, i.e. one inner
iteration of the adaptive loop: assemble K, f, g -> MINRES to a true relative
residual of 1e-10 -> estimator (patch problems in FP32 with two FP64
refinement steps, per-cell indicators).  `value` is (n_X + n_Y) divided by the
mean device time of assembly + estimator per step; MINRES seconds per solve
are reported beside it.  Inputs are resident in HBM when the timed region
starts (the mesh, pattern and coefficients are uploaded once).

    python bench.py [--steps K] [--warmup W] [--config c2|c1|c3|c4|c5]

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
MEASURED_TRAFFIC = {"c2": {"spmv": 2.11e9, "solve": 2.49e9, "factor": 0.28e9}}


def workload(name):
    import doubleadapparabolicfem_amd as apf
    from doubleadapparabolicfem_amd import meshes

    if name == "c2":
        m = meshes.uniform_mesh(2, 5)
        return dict(mesh=m, p=2, coarse=meshes.uniform_hierarchy(2, 5), coeff=meshes.smooth_initial(2),
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
            est = sysm.estimate(apf.APF_FP32_REFINE, 2, per_cell=False)
            if i >= warm:
                wall += time.perf_counter() - t0
                rec.append((t_asm, t_minres, st["iterations"], (est["seconds_patches"] + est["seconds_diagnostics"]) * 1e3,
                            sysm.phase_timings(), est))
        sysm.estimate(apf.APF_FP64, 0, per_cell=False)  # untimed: the first call tunes the tiles
        est64 = sysm.estimate(apf.APF_FP64, 0, per_cell=False)
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
            "estimator_ms_fp64": (est64["seconds_patches"] + est64["seconds_diagnostics"]) * 1e3,
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
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2+1D fan-patch mesh: 8x8 squares, side splits U{1..16} (seed 3), fan centres "
                               "with K in 4..64 cells; 12 intervals graded toward t=0 (2^-10 .. 1/4); "
                               "u0 = indicator of [0.25,0.75]^2; p = 1, 2, 3 in turn",
                   "space_cells": int(m.cells.shape[0]), "intervals": int(m.times.size - 1)},
        "roofline": dict(roof[dom], degree=dom_p, traffic=None),
        "per_degree": per_p,
    }
    print(json.dumps(out))


def patch_work(sp, steps):
    """Algorithmic work of one estimate's patch phases (n = order of a spatial
    vertex patch problem of one Legendre time mode: flux + scalar + mean-zero
    multiplier; the matrix is shared by every slab and mode of a tensor mesh,
    so each patch is inverted once and applied to cols = slabs x modes
    right-hand sides):
      factor: Gauss-Jordan inversion, 2 n^3 flops;
      solve:  X = K^-1 B and `steps` refinement rounds, each one FP64
              residual product and one correction product, plus the final
              FP64 residual: (1 + steps) products in the inverse's precision
              and (1 + steps) FP64 products, 2 n^2 cols flops each."""
    nf, ns_, _, bd = sp.patch_sizes()
    n = (nf + ns_ + (bd == 0)).astype(np.float64)
    cols = sp.n_intervals * (sp.p + 1)
    prod = float(np.sum(2.0 * n ** 2 * cols))
    return {"factor": float(np.sum(2.0 * n ** 3)), "solve_ft": (1 + steps) * prod, "solve_fp64": (1 + steps) * prod,
            "lu_and_solve": float(np.sum(2.0 / 3.0 * n ** 3 + 2.0 * n ** 2 * cols))}


def phase_rooflines(sp, ph, steps, mixed):
    """Roofline entries of the patch kernels from their measured times."""
    w = patch_work(sp, steps)
    ft_peak = FP32_PEAK_TFLOPS if mixed else FP64_PEAK_TFLOPS
    out = {}
    t = ph["factor_ms"] * 1e-3
    ach = w["factor"] / t / 1e12
    name = ("k_patch_factor_blocked (Gauss-Jordan: FP32 VALU panels, MFMA rank-16 updates)" if mixed
            else "k_patch_factor_full (register Gauss-Jordan, FP64 VALU)")
    out["factor"] = {"bound": "mfma", "kernel": name, "achieved": ach, "peak": ft_peak, "unit": "TFLOP/s",
                     "frac": ach / ft_peak,
                     "algorithmic_flops_per_launch": w["factor"], "avg_launch_ms": ph["factor_ms"]}
    t = ph["solve_ms"] * 1e-3
    flops = w["solve_ft"] + w["solve_fp64"]
    # time at peak: products in the inverse's precision at its MFMA peak, FP64 residuals at the FP64 peak
    t_ideal = w["solve_ft"] / (ft_peak * 1e12) + w["solve_fp64"] / (FP64_PEAK_TFLOPS * 1e12)
    peak = flops / t_ideal / 1e12
    ach = flops / t / 1e12
    out["solve"] = {"bound": "mfma", "kernel": "k_patch_solve (MFMA 16x16x4: " + ("FP32" if mixed else "FP64")
                    + " products, FP64 residuals)", "achieved": ach, "peak": peak, "unit": "TFLOP/s",
                    "frac": ach / peak, "algorithmic_flops_per_launch": flops, "avg_launch_ms": ph["solve_ms"]}
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


def cpu_baseline(min_s=10.0):
    """The oracle (numpy/scipy FP64 restatement, one process) on a bounded
    sample of the same workload: 2+1D p=2 smooth_initial on the uniform
    4x4-square x 2-interval mesh, assembly + estimator, repeated until at least
    `min_s` seconds of CPU work are measured."""
    from oracle import assembly as OA, cases, estimator as OE, mesh as OM

    stm = cases.uniform_mesh(2, 1)
    d = OM.spacetime_dofs(stm, 2)
    prob = cases.smooth_initial(stm)
    S = OA.assemble(stm, d, prob)
    x = OE.main_solve(S)
    t_total, reps = 0.0, 0
    while t_total < min_s:
        t0 = time.perf_counter()
        S = OA.assemble(stm, d, prob)
        OE.estimate(stm, d, prob, S, x)
        t_total += time.perf_counter() - t0
        reps += 1
    dofs = d.nX + d.nY
    return {"value": dofs * reps / t_total, "unit": "DOF/s", "cores": 1, "kind": "port",
            "sample": f"oracle assembly+estimator, 2+1D p=2 uniform 4x4 squares x 2 intervals "
                      f"({dofs} DOFs) x {reps} reps in {t_total:.1f} s, single-threaded numpy/scipy"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-g", action="store_true", help="skip the G-estimator block (untimed, outside `value`)")
    ap.add_argument("--profile", action="store_true",
                    help="one assembly, MINRES solve, two mixed and two FP64 estimates, 20 SpMVs and a 1 GiB "
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
    W = workload(args.config)
    sp = apf.Space(W["mesh"], W["p"], coarse=W["coarse"])
    ctx = apf.Context(0)
    sysm = apf.System(ctx, sp)
    n_dofs = sp.nX + sp.nY

    if args.profile:
        # counter runs: one pass of each hot-path kernel, then the calibration copy
        sysm.assemble(W["coeff"])
        sysm.minres(tol=1e-10, max_it=5000, check_interval=10)
        # each estimate twice: a system's first call in a precision also times its tile
        # candidates, the second runs the chosen tiles only
        for _ in range(2):
            sysm.estimate(apf.APF_FP32_REFINE, 2, per_cell=False)
        for _ in range(2):
            sysm.estimate(apf.APF_FP64, 0, per_cell=False)
        ms = sysm.bench_spmv(20)
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
        est = sysm.estimate(prec, 2, per_cell=False)
        t_est = (est["seconds_patches"] + est["seconds_diagnostics"]) * 1e3
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
    est_mixed = rec[-1][4]
    # FP64 estimator path on the same solution for the mixed-vs-FP64 error and speed ratio (one
    # untimed call first: a system's first estimate in a precision times its tile candidates)
    sysm.estimate(apf.APF_FP64, 0, per_cell=False)
    est64 = sysm.estimate(apf.APF_FP64, 0, per_cell=False)
    t_est64 = (est64["seconds_patches"] + est64["seconds_diagnostics"]) * 1e3
    rel = {k: abs(est_mixed[k] - est64[k]) / abs(est64[k]) for k in ("flux2", "eta2", "lambda_y2", "y_est2")}
    value = n_dofs / ((t_asm + t_est) * 1e-3)

    # rooflines of the patch phases (the dominant kernels of `value`)
    ph_mixed = {k: float(np.mean([r[5][k] for r in rec])) for k in rec[0][5]}
    ph64 = sysm.phase_timings()
    roof = phase_rooflines(sp, ph_mixed, 2, True)
    roof64 = phase_rooflines(sp, ph64, 0, False)
    dom = max(("factor", "solve"), key=lambda k: roof[k]["avg_launch_ms"])
    # HBM roofline of the MINRES SpMV: algorithmic bytes 12 nnz + 20 n
    spmv_ms = sysm.bench_spmv(50)
    stream_gbs = ctx.bench_stream(1 << 30, 10)
    spmv_bytes = 12 * sp.nnz + 20 * sp.n
    achieved = spmv_bytes / (spmv_ms * 1e-3) / 1e9
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": W["desc"], "p": W["p"], "n_X": sp.nX, "n_Y": sp.nY, "nnz_K": sp.nnz,
                   "space_time_cells": sp.n_space_cells * sp.n_intervals,
                   "vertex_patches": sp.n_patches * sp.n_intervals},
        "roofline": dict(roof[dom], traffic=traffic.get(dom)),
        "roofline_patch_phases_mixed": roof,
        "roofline_patch_phases_fp64": roof64,
        "roofline_minres_spmv": {"bound": "hbm", "kernel": "k_spmv (MINRES CSR SpMV, FP64 values, int32 columns)",
                                 "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": achieved / HBM_PEAK_GBS, "traffic": traffic.get("spmv"),
                                 "algorithmic_bytes_per_launch": spmv_bytes, "avg_launch_ms": spmv_ms,
                                 "measured_copy_GBs": stream_gbs, "frac_of_measured_copy": achieved / stream_gbs},
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
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
