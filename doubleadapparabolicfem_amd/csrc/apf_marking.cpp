// Host-side estimator combination and deterministic Doerfler marking.
//
// Restates, for per-cell arrays indexed by cell id:
//  * combine_time_slab_estimator_squared_errors_deterministic
//    (src/finite_element/time_slabs/time_slab_estimator_combination.hpp:88-250):
//    estimator^2 = flux^2 + reconstruction^2 per source cell, every squared
//    quantity checked finite and >= -1e-14
//    (time_slabs/detail/time_slab_error_indicator_detail.hpp:63-73, 124-133);
//  * CellwiseSquaredError::doerfler_marking_deterministic
//    (src/finite_element/time_slabs/time_slab_cellwise_errors.hpp:156-280):
//    Neumaier-compensated total over ids ascending, entries sorted by value
//    descending with ties to the lower id, marked until the compensated
//    cumulative sum reaches theta * total (:297-321 for the compensation);
//  * compute_y_estimator_threshold_squared
//    (algorithm/detail/indicator_utilities.hpp:475-492).
// The decisions are integer results and are bit-exact functions of the inputs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/apf.h"

extern thread_local std::string g_last_error;

namespace {

int bad(const std::string& m)
{
    g_last_error = m;
    return APF_BAD_ARG;
}

bool squared_ok(double v) { return std::isfinite(v) && v >= -1.0e-14; }

void neumaier_add(double& sum, double& comp, double value)
{
    const double t = sum + value;
    if (std::abs(sum) >= std::abs(value))
        comp += (sum - t) + value;
    else
        comp += (value - t) + sum;
    sum = t;
}

}  // namespace

extern "C" {

int apf_combine_estimator(int64_t n, const double* flux2, const double* recon2, const double* resid2, double* est2)
{
    if (n < 0 || (n > 0 && (!flux2 || !recon2 || !est2))) return bad("apf_combine_estimator: null argument");
    const char* ctx = "combine_time_slab_estimator_squared_errors";
    for (int64_t i = 0; i < n; ++i) {
        if (!squared_ok(flux2[i])) return bad(std::string(ctx) + " flux: squared quantity is negative or non-finite.");
        if (!squared_ok(recon2[i]))
            return bad(std::string(ctx) + " reconstruction: squared quantity is negative or non-finite.");
        if (resid2 && !squared_ok(resid2[i]))
            return bad(std::string(ctx) + " divergence residual: squared quantity is negative or non-finite.");
    }
    for (int64_t i = 0; i < n; ++i) {
        est2[i] = flux2[i] + recon2[i];
        if (!squared_ok(est2[i]))
            return bad(std::string(ctx) + " estimator: squared quantity is negative or non-finite.");
    }
    return APF_OK;
}

int apf_doerfler_mark(int64_t n, const int64_t* ids, const double* values, double theta, double near_tie_tolerance,
                      int64_t* marked, int64_t* n_marked, apf_doerfler_diag* diag)
{
    if (n < 0 || (n > 0 && (!values || !marked)) || !n_marked) return bad("apf_doerfler_mark: null argument");
    if (!(theta > 0.0 && theta <= 1.0))
        return bad("CellwiseSquaredError::doerfler_marking_deterministic: theta must be in (0,1].");
    if (!(near_tie_tolerance >= 0.0))
        return bad("CellwiseSquaredError::doerfler_marking_deterministic: near-tie tolerance must be nonnegative.");
    try {
        std::vector<int64_t> id(n);
        if (ids)
            std::copy(ids, ids + n, id.begin());
        else
            std::iota(id.begin(), id.end(), int64_t{0});
        // entries by id ascending (ids are keys of a map: they must be unique)
        std::vector<int64_t> by_id(n);
        std::iota(by_id.begin(), by_id.end(), int64_t{0});
        std::sort(by_id.begin(), by_id.end(), [&](int64_t a, int64_t b) { return id[a] < id[b]; });
        for (int64_t i = 1; i < n; ++i)
            if (id[by_id[i]] == id[by_id[i - 1]]) return bad("apf_doerfler_mark: duplicate cell id");
        double sum = 0.0, comp = 0.0;
        for (int64_t i = 0; i < n; ++i) neumaier_add(sum, comp, values[by_id[i]]);
        apf_doerfler_diag d{};
        d.total = sum + comp;
        d.target = theta * d.total;
        d.cutoff_index = n;
        *n_marked = 0;
        if (!(d.total > 0.0)) {
            if (diag) *diag = d;
            return APF_OK;
        }
        std::vector<int64_t> order(n);
        std::iota(order.begin(), order.end(), int64_t{0});
        std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
            if (values[a] != values[b]) return values[a] > values[b];
            return id[a] < id[b];
        });
        d.max_indicator = n ? values[order[0]] : 0.0;
        double cum = 0.0, cc = 0.0;
        int64_t m = 0;
        for (int64_t i = 0; i < n; ++i) {
            marked[m++] = id[order[i]];
            neumaier_add(cum, cc, values[order[i]]);
            if (cum + cc >= d.target) {
                d.cutoff_index = i;
                d.cutoff_value = values[order[i]];
                if (i + 1 < n) d.cutoff_margin = values[order[i]] - values[order[i + 1]];
                break;
            }
        }
        *n_marked = m;
        if (d.cutoff_index < n) {
            int64_t ties = 0;
            for (int64_t i = 0; i < n; ++i) {
                if (i == d.cutoff_index) continue;
                if (std::abs(values[order[i]] - d.cutoff_value) <= near_tie_tolerance) ++ties;
            }
            d.near_tie_count = ties;
        }
        if (diag) *diag = d;
        return APF_OK;
    } catch (const std::exception& e) {
        return bad(std::string("apf_doerfler_mark: ") + e.what());
    }
}

int apf_y_estimator_threshold(double rho, double lambda_y2, double trace2, double* out)
{
    if (!out) return bad("apf_y_estimator_threshold: null argument");
    if (!(rho > 0.0)) return bad("compute_y_estimator_threshold_squared: rho must be positive.");
    const double shifted = rho + 0.5;
    const double factor = 1.0 + 1.0 / (shifted * shifted);
    *out = rho * rho * (lambda_y2 + factor * trace2);
    return APF_OK;
}

}  // extern "C"
