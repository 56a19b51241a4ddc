"""Estimator combination and deterministic Doerfler marking (oracle; test infrastructure only).

Restates, in plain Python:

* combine_time_slab_estimator_squared_errors_deterministic
  (src/finite_element/time_slabs/time_slab_estimator_combination.hpp:88-250):
  estimator^2 = flux^2 + reconstruction^2 per source cell; every squared value
  must be finite and >= -1e-14
  (src/finite_element/time_slabs/detail/time_slab_error_indicator_detail.hpp:63-73);
* CellwiseSquaredError::doerfler_marking_deterministic
  (src/finite_element/time_slabs/time_slab_cellwise_errors.hpp:156-280), with
  neumaier_add_ (:297-308) and deterministic_sum_entries_ (:310-321):
  Neumaier total over ids ascending; entries sorted by value descending, ties
  to the lower id; marked until the compensated cumulative sum reaches
  theta * total;
* compute_y_estimator_threshold_squared (algorithm/detail/indicator_utilities.hpp:475-492).

Pinned by exact rational brute force on small inputs (the marked set is the
shortest prefix of the sorted order whose exact sum reaches theta times the
exact total, when no partial sum is within rounding of the target), by a case
where the compensation changes the marked set, and by special cases of the
threshold (tests/test_marking.py).
"""
from __future__ import annotations

import math


class MarkingError(ValueError):
    pass


def neumaier_add(s, c, v):
    t = s + v
    if abs(s) >= abs(v):
        c += (s - t) + v
    else:
        c += (v - t) + s
    return t, c


def _check_squared(v, what):
    if not math.isfinite(v) or v < -1.0e-14:
        raise MarkingError(f"combine_time_slab_estimator_squared_errors {what}: squared quantity is negative or non-finite.")


def combine_estimator(flux2, recon2, resid2=None):
    for i in range(len(flux2)):
        _check_squared(flux2[i], "flux")
        _check_squared(recon2[i], "reconstruction")
        if resid2 is not None:
            _check_squared(resid2[i], "divergence residual")
    out = [float(flux2[i]) + float(recon2[i]) for i in range(len(flux2))]
    for v in out:
        _check_squared(v, "estimator")
    return out


def doerfler_mark(values, theta, ids=None, near_tie_tolerance=0.0, compensated=True):
    """Returns (marked ids in marking order, diagnostics dict).  `compensated`
    exists only so a test can show that plain summation changes the result."""
    if not (0.0 < theta <= 1.0):
        raise MarkingError("theta must be in (0,1]")
    if not near_tie_tolerance >= 0.0:
        raise MarkingError("near-tie tolerance must be nonnegative")
    n = len(values)
    ids = list(range(n)) if ids is None else [int(i) for i in ids]
    if len(set(ids)) != n:
        raise MarkingError("duplicate cell id")
    entries = [(ids[i], float(values[i])) for i in range(n)]

    def add(s, c, v):
        return neumaier_add(s, c, v) if compensated else (s + v, c)

    s = c = 0.0
    for _, v in sorted(entries, key=lambda e: e[0]):
        s, c = add(s, c, v)
    total = s + c
    diag = dict(total=total, target=theta * total, cutoff_value=0.0, cutoff_margin=0.0, max_indicator=0.0,
                cutoff_index=n, near_tie_count=0)
    if not total > 0.0:
        return [], diag
    order = sorted(entries, key=lambda e: (-e[1], e[0]))
    diag["max_indicator"] = order[0][1] if order else 0.0
    marked = []
    cum = cc = 0.0
    for i, (cid, v) in enumerate(order):
        marked.append(cid)
        cum, cc = add(cum, cc, v)
        if cum + cc >= diag["target"]:
            diag["cutoff_index"] = i
            diag["cutoff_value"] = v
            if i + 1 < n:
                diag["cutoff_margin"] = v - order[i + 1][1]
            break
    if diag["cutoff_index"] < n:
        cut = diag["cutoff_value"]
        diag["near_tie_count"] = sum(1 for i, (_, v) in enumerate(order)
                                     if i != diag["cutoff_index"] and abs(v - cut) <= near_tie_tolerance)
    return marked, diag


def y_estimator_threshold(rho, lambda_y2, trace2):
    if not rho > 0.0:
        raise MarkingError("rho must be positive")
    shifted = rho + 0.5
    return rho * rho * (lambda_y2 + (1.0 + 1.0 / (shifted * shifted)) * trace2)
