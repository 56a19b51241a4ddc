"""Estimator combination and deterministic Doerfler marking: the oracle
pinned by exact rational brute force and a compensation-sensitive case, then
the host C ABI (no GPU) bit-exact against the oracle."""
import math
from fractions import Fraction

import numpy as np
import pytest

import doubleadapparabolicfem_amd as apf
from oracle import marking as OM


def exact_prefix(values, ids, theta):
    """Shortest prefix of the (value desc, id asc) order whose exact sum reaches
    theta * exact total; None when a partial sum is within 1e-9 of the target
    (then rounding may decide and the brute force is not a referee)."""
    total = sum(Fraction(v) for v in values)
    if total <= 0:
        return []
    target = Fraction(theta) * total
    order = sorted(zip(ids, values), key=lambda e: (-e[1], e[0]))
    s = Fraction(0)
    for i, (cid, v) in enumerate(order):
        s += Fraction(v)
        if abs(s - target) <= Fraction(1, 10 ** 9) * total:
            return None
        if s >= target:
            return [c for c, _ in order[: i + 1]]
    return [c for c, _ in order]


def test_oracle_matches_exact_brute_force():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(400):
        n = int(rng.integers(1, 13))
        values = [float(x) for x in np.round(rng.exponential(1.0, n), 3)]
        ids = [int(x) for x in rng.permutation(100)[:n]]
        theta = float(rng.choice([0.1, 0.25, 0.4, 0.5, 0.75, 0.9, 1.0]))
        ref = exact_prefix(values, ids, theta)
        if ref is None:
            continue
        got, _ = OM.doerfler_mark(values, theta, ids)
        assert got == ref
        checked += 1
    assert checked > 300


def test_oracle_ties_go_to_lower_id_and_zero_total_marks_nothing():
    marked, d = OM.doerfler_mark([1.0, 2.0, 2.0, 1.0], 0.5, ids=[7, 9, 3, 1])
    assert marked == [3, 9]            # 2.0 (id 3) first, then 2.0 (id 9) reaches 3.0 = 0.5 * 6
    assert d["cutoff_index"] == 1 and d["cutoff_value"] == 2.0 and d["cutoff_margin"] == 1.0
    marked, d = OM.doerfler_mark([0.0, 0.0], 0.5)
    assert marked == [] and d["cutoff_index"] == 2


def test_compensation_changes_the_marking():
    vals = [1.0] + [1e-16] * 20
    comp, _ = OM.doerfler_mark(vals, 1.0)
    plain, _ = OM.doerfler_mark(vals, 1.0, compensated=False)
    assert len(plain) == 1 and len(comp) == 20      # plain summation loses every small term
    marked, _ = apf.doerfler_mark(vals, 1.0)
    assert list(marked) == comp


def test_neumaier_total_is_within_an_ulp_of_exact():
    rng = np.random.default_rng(5)
    v = list(rng.exponential(1.0, 5000) * 10.0 ** rng.integers(-12, 12, 5000))
    _, d = OM.doerfler_mark(v, 0.5)
    assert abs(d["total"] - math.fsum(v)) <= 2 * math.ulp(math.fsum(v))


@pytest.mark.parametrize("seed", range(6))
def test_abi_matches_oracle_bitwise(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3000))
    vals = rng.exponential(1.0, n) * 10.0 ** rng.integers(-6, 3, n)
    vals[rng.random(n) < 0.2] = 0.0
    vals = np.round(vals, int(rng.integers(2, 12)))      # plenty of exact ties
    ids = rng.permutation(10 * n)[:n].astype(np.int64)
    theta = float(rng.uniform(0.05, 1.0))
    tol = float(rng.choice([0.0, 1e-3]))
    m_abi, d_abi = apf.doerfler_mark(vals, theta, ids, tol)
    m_or, d_or = OM.doerfler_mark(list(vals), theta, list(ids), tol)
    assert list(m_abi) == m_or
    for k in d_or:
        assert d_abi[k] == d_or[k], k


def test_combine_and_threshold():
    rng = np.random.default_rng(3)
    f, r = rng.random(100), rng.random(100)
    out = apf.combine_estimator(f, r, rng.random(100))
    assert np.array_equal(out, np.array(OM.combine_estimator(list(f), list(r))))
    with pytest.raises(apf.APFError, match="flux: squared quantity is negative"):
        apf.combine_estimator([-1e-13], [0.0])
    with pytest.raises(apf.APFError, match="non-finite"):
        apf.combine_estimator([0.0], [np.nan])
    apf.combine_estimator([-1e-15], [0.0])              # within the -1e-14 slack
    # threshold: rho = 1/2 gives factor 2; trace2 = 0 leaves rho^2 lambda_Y^2
    assert apf.y_estimator_threshold(0.5, 3.0, 5.0) == 0.25 * (3.0 + 2.0 * 5.0)
    assert apf.y_estimator_threshold(2.0, 3.0, 0.0) == 12.0
    with pytest.raises(apf.APFError, match="rho must be positive"):
        apf.y_estimator_threshold(0.0, 1.0, 1.0)


def test_marking_errors():
    with pytest.raises(apf.APFError, match="theta must be in"):
        apf.doerfler_mark([1.0], 0.0)
    with pytest.raises(apf.APFError, match="theta must be in"):
        apf.doerfler_mark([1.0], 1.5)
    with pytest.raises(apf.APFError, match="near-tie tolerance"):
        apf.doerfler_mark([1.0], 0.5, near_tie_tolerance=-1.0)
    with pytest.raises(apf.APFError, match="duplicate cell id"):
        apf.doerfler_mark([1.0, 2.0], 0.5, ids=[4, 4])
    marked, d = apf.doerfler_mark([], 0.5)
    assert marked.size == 0 and d["total"] == 0.0
