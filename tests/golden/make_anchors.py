"""Extract the published uniform-run estimator totals used as golden anchors.

Reads the reference's published CSV histories
(algorithm_data_plot/{1d,2d}/smooth_initial.csv, record_type == "outer",
mode == "uniform_standard") and writes tests/golden/anchors.json.  The values
are data (6 significant digits as published); no reference code is run.

Also writes tests/golden/g_anchors.json: the G-estimator values
(g_lambda_difference_squared) of the adaptive runs' first record on the initial
mesh, identified by X true DOFs equal to the uniform run's outer iteration 0, G
true DOFs equal to the degree-(p+1) uniform Y true DOFs at outer iteration 0, and
the same eta^2 and lambda_y^2 as that uniform record.

Also writes tests/golden/adaptive1d_anchors.json: the 1+1D smooth_initial adaptive runs
(modes adaptive_standard and adaptive_force_effective_rho_no_inner_refinement, p = 1, 2, 3),
one record per outer iteration while the inner Y loop accepted its first iterate; and
tests/golden/adaptive1d_inner_anchors.json: every inner-loop record of the adaptive_standard runs, and
the G values (g_true_dofs, g_lambda_difference_squared) of the adaptive_g_effective_rho runs (p = 1, 2),
whose meshes follow the same history (inner iteration 0 with Y true DOFs unchanged and the
next record starting the next outer iteration): X true DOFs before and after the X
refinement, Y true DOFs, eta^2, lambda_Y^2, y_estimator^2 and its threshold.

Usage:  python tests/golden/make_anchors.py [/root/reference]
"""
import csv
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
COLS = ["x_true_dofs", "y_true_dofs_after", "eta_squared", "lambda_y_squared",
        "y_estimator_squared", "y_flux_squared", "y_reconstruction_squared",
        "divergence_residual_l2", "iteration_seconds"]


def main():
    out = []
    for dim in ("1d", "2d"):
        path = os.path.join(REF, "algorithm_data_plot", dim, "smooth_initial.csv")
        with open(path) as fh:
            for i, row in enumerate(csv.DictReader(fh)):
                if row["record_type"] != "outer" or row["mode"] != "uniform_standard":
                    continue
                rec = {"dim": int(dim[0]), "p": int(row["p_int"]),
                       "outer": int(row["outer_iteration"]),
                       "source": f"algorithm_data_plot/{dim}/smooth_initial.csv:{i + 2}"}
                for c in COLS:
                    rec[c] = float(row[c])
                out.append(rec)
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "anchors.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {len(out)} anchors to {dst}")
    first = {(a["dim"], a["p"]): a for a in out if a["outer"] == 0}
    gout = []
    for dim in ("1d", "2d"):
        path = os.path.join(REF, "algorithm_data_plot", dim, "smooth_initial.csv")
        seen = set()
        with open(path) as fh:
            for i, row in enumerate(csv.DictReader(fh)):
                if row["g_estimator_computed"] != "1" or not row["g_lambda_difference_squared"]:
                    continue
                d, p = int(dim[0]), int(row["p_int"])
                u, uq = first.get((d, p)), first.get((d, p + 1))
                if u is None or uq is None or (d, p) in seen:
                    continue
                if (int(float(row["x_true_dofs"])) != u["x_true_dofs"]
                        or int(float(row["g_true_dofs"])) != uq["y_true_dofs_after"]
                        or float(row["eta_squared"]) != u["eta_squared"]
                        or float(row["lambda_y_squared"]) != u["lambda_y_squared"]):
                    continue
                seen.add((d, p))
                gout.append({"dim": d, "p": p, "outer": 0, "source": f"algorithm_data_plot/{dim}/smooth_initial.csv:{i + 2}",
                             "mode": row["mode"], "g_true_dofs": int(float(row["g_true_dofs"])),
                             "g_lambda_difference_squared": float(row["g_lambda_difference_squared"])})
    gdst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "g_anchors.json")
    with open(gdst, "w") as fh:
        json.dump(gout, fh, indent=1)
    print(f"wrote {len(gout)} G anchors to {gdst}")
    aout = []
    path = os.path.join(REF, "algorithm_data_plot", "1d", "smooth_initial.csv")
    with open(path) as fh:
        rows = [(i + 2, row) for i, row in enumerate(csv.DictReader(fh))]
    for mode, p in [(m, p) for m in ("adaptive_standard", "adaptive_force_effective_rho_no_inner_refinement")
                    for p in (1, 2, 3)]:
        inner = [(ln, r) for ln, r in rows
                 if r["record_type"] == "inner" and r["mode"] == mode and r["p_int"] == str(p)]
        for k, (ln, r) in enumerate(inner):
            if int(r["outer_iteration"]) != k or r["inner_iteration"] != "0":
                break
            if r["y_true_dofs_before"] != r["y_true_dofs_after"]:
                break
            aout.append({"dim": 1, "p": p, "outer": k, "source": f"algorithm_data_plot/1d/smooth_initial.csv:{ln}",
                         "mode": r["mode"], "x_true_dofs": int(float(r["x_true_dofs"])),
                         "x_true_dofs_after": int(float(r["x_true_dofs_after"])),
                         "y_true_dofs": int(float(r["y_true_dofs_before"])),
                         "eta_squared": float(r["eta_squared"]), "lambda_y_squared": float(r["lambda_y_squared"]),
                         "y_estimator_squared": float(r["y_estimator_squared"]),
                         "y_estimator_threshold_squared": float(r["y_estimator_threshold_squared"])})
    adst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adaptive1d_anchors.json")
    with open(adst, "w") as fh:
        json.dump(aout, fh, indent=1)
    print(f"wrote {len(aout)} adaptive 1+1D anchors to {adst}")
    iout = []
    for p in (1, 2, 3):
        for ln, r in rows:
            if r["record_type"] != "inner" or r["mode"] != "adaptive_standard" or r["p_int"] != str(p):
                continue
            iout.append({"dim": 1, "p": p, "outer": int(r["outer_iteration"]), "inner": int(r["inner_iteration"]),
                         "source": f"algorithm_data_plot/1d/smooth_initial.csv:{ln}", "mode": r["mode"],
                         "x_true_dofs": int(float(r["x_true_dofs"])),
                         "y_true_dofs_before": int(float(r["y_true_dofs_before"])),
                         "y_true_dofs_after": int(float(r["y_true_dofs_after"])),
                         "eta_squared": float(r["eta_squared"]), "lambda_y_squared": float(r["lambda_y_squared"]),
                         "y_estimator_squared": float(r["y_estimator_squared"]),
                         "y_estimator_threshold_squared": float(r["y_estimator_threshold_squared"])})
    for p in (1, 2):
        for ln, r in rows:
            if r["record_type"] != "inner" or r["mode"] != "adaptive_g_effective_rho" or r["p_int"] != str(p):
                continue
            if not r["g_lambda_difference_squared"]:
                continue
            iout.append({"dim": 1, "p": p, "outer": int(r["outer_iteration"]), "inner": int(r["inner_iteration"]),
                         "source": f"algorithm_data_plot/1d/smooth_initial.csv:{ln}", "mode": r["mode"],
                         "x_true_dofs": int(float(r["x_true_dofs"])),
                         "y_true_dofs_before": int(float(r["y_true_dofs_before"])),
                         "y_true_dofs_after": int(float(r["y_true_dofs_after"])),
                         "g_true_dofs": int(float(r["g_true_dofs"])),
                         "g_lambda_difference_squared": float(r["g_lambda_difference_squared"])})
    idst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "adaptive1d_inner_anchors.json")
    with open(idst, "w") as fh:
        json.dump(iout, fh, indent=1)
    print(f"wrote {len(iout)} adaptive 1+1D inner-loop anchors to {idst}")


if __name__ == "__main__":
    main()
