"""Extract the published uniform-run estimator totals used as golden anchors.

Reads the reference's published CSV histories
(algorithm_data_plot/{1d,2d}/smooth_initial.csv, record_type == "outer",
mode == "uniform_standard") and writes tests/golden/anchors.json.  The values
are data (6 significant digits as published); no reference code is run.

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


if __name__ == "__main__":
    main()
