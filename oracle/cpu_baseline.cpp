// FP64 C++/OpenMP restatement of the hot path's CPU work (test infrastructure: the CPU
// baseline that bench.py times on the host cores; never part of the product).
//
//  * cpu_assemble: element assembly as the reference's OpenMP region does it
//    (src/finite_element/assembly/main_system/dense_local_main_blocks_2d.hpp:176-361):
//    per Y cell the full q-loop over the prism rule, local A_y and B accumulated entry by
//    entry, then scattered into the CSR slots (scatter.hpp:66-115, |a_ij| <= 1e-15
//    skipped).  Tensor meshes: the X cell is the Y cell.
//  * cpu_patch_solves: the per-(slab, vertex) dense solves of the estimator
//    (dense_local_error_blocks_2d.hpp:1060-1143): LU with partial pivoting of the coupled
//    N x N patch system, one solve, and the residual acceptance
//    |K x - b|_inf <= 1e-10 max(1, |b|_inf) (:923-949).
//  * cpu_minres: preconditioned MINRES on K with the Chebyshev-Jacobi block
//    preconditioner of apf_csr_minres_solve (same algorithm as the GPU caller's-CSR
//    path), true relative residual checked every check_interval iterations.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

double now() { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); }

int64_t find_slot(const int64_t* rowptr, const int* col, int64_t r, int c)
{
    int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo <= hi && col[lo] == c) ? lo : -1;
}

inline void atomic_add(double* a, double v)
{
#pragma omp atomic
    *a += v;
}

}  // namespace

extern "C" {

// Symbolic CSR pattern of K = [A_y B^T; B C] from per-cell maps (-1 = Dirichlet):
// call with rowptr only to get nnz (returned), then with col to fill.
int64_t cpu_pattern(int64_t nc, int nl, const int* ymap, const int* xmap, int64_t nY, int64_t nX, int64_t nbot,
                    const int* bottom_cells, int64_t* rowptr, int* col)
{
    const int64_t n = nY + nX;
    std::vector<std::vector<int>> rows(n);
    std::vector<int> sy, sx;
    auto uniq = [&](const int* m, std::vector<int>& out) {
        out.clear();
        for (int i = 0; i < nl; ++i)
            if (m[i] >= 0) out.push_back(m[i]);
        std::sort(out.begin(), out.end());
        out.erase(std::unique(out.begin(), out.end()), out.end());
    };
    for (int64_t c = 0; c < nc; ++c) {
        uniq(ymap + c * nl, sy);
        uniq(xmap + c * nl, sx);
        for (int g : sy) {
            for (int h : sy) rows[g].push_back(h);
            for (int h : sx) rows[g].push_back(static_cast<int>(nY + h));
        }
        for (int g : sx)
            for (int h : sy) rows[nY + g].push_back(h);
    }
    for (int64_t b = 0; b < nbot; ++b) {
        uniq(xmap + static_cast<int64_t>(bottom_cells[b]) * nl, sx);
        for (int g : sx)
            for (int h : sx) rows[nY + g].push_back(static_cast<int>(nY + h));
    }
    rowptr[0] = 0;
    for (int64_t r = 0; r < n; ++r) {
        auto& v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        rowptr[r + 1] = rowptr[r] + static_cast<int64_t>(v.size());
        if (col) std::copy(v.begin(), v.end(), col + rowptr[r]);
    }
    return rowptr[n];
}

// Element assembly of A_y and B (q-loop) + scatter, `reps` times; returns seconds per rep.
// geo [nc][8]: spatial vertices (dim+1 points, dim coords each) then t0, t1.
// M [nc][dim*dim].  Tables on the prism rule: w [nq], phi [nq][nl],
// dphi [nq][nl][dim+1] (reference spatial gradient components, then d/dtau).
double cpu_assemble(int dim, int nthreads, int64_t nc, const double* geo, const double* M, int nl, const int* ymap,
                    const int* xmap, int64_t nY, int nq, const double* w, const double* phi, const double* dphi,
                    const int64_t* rowptr, const int* col, int64_t nnz, double* val, int reps)
{
    omp_set_num_threads(nthreads);
    double best = 1e300;
    for (int rep = 0; rep < reps; ++rep) {
        const double t0 = now();
        // the values start from zero in every assembly
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < nnz; ++e) val[e] = 0.0;
#pragma omp parallel
        {
            std::vector<double> A(static_cast<size_t>(nl) * nl), B(static_cast<size_t>(nl) * nl);
            std::vector<double> gx(static_cast<size_t>(nl) * 2), Mg(static_cast<size_t>(nl) * 2), dt(nl);
#pragma omp for schedule(static)
            for (int64_t c = 0; c < nc; ++c) {
                const double* g = geo + c * 8;
                double Ji[4], det;
                if (dim == 1) {
                    const double J = g[1] - g[0];
                    Ji[0] = 1.0 / J;
                    det = std::fabs(J);
                } else {
                    const double a = g[2] - g[0], b = g[4] - g[0], cc = g[3] - g[1], d = g[5] - g[1];
                    const double dd = a * d - b * cc;
                    Ji[0] = d / dd; Ji[1] = -b / dd; Ji[2] = -cc / dd; Ji[3] = a / dd;
                    det = std::fabs(dd);
                }
                const double h = g[7] - g[6];
                const double* Mc = M + c * dim * dim;
                std::fill(A.begin(), A.end(), 0.0);
                std::fill(B.begin(), B.end(), 0.0);
                for (int q = 0; q < nq; ++q) {
                    const double dmu = w[q] * det * h;
                    const double* gr = dphi + static_cast<int64_t>(q) * nl * (dim + 1);
                    for (int j = 0; j < nl; ++j) {
                        if (dim == 1) {
                            gx[2 * j] = Ji[0] * gr[j * 2];
                            gx[2 * j + 1] = 0.0;
                            Mg[2 * j] = Mc[0] * gx[2 * j];
                            Mg[2 * j + 1] = 0.0;
                            dt[j] = gr[j * 2 + 1] / h;
                        } else {
                            const double r0 = gr[j * 3], r1 = gr[j * 3 + 1];
                            gx[2 * j] = Ji[0] * r0 + Ji[2] * r1;
                            gx[2 * j + 1] = Ji[1] * r0 + Ji[3] * r1;
                            Mg[2 * j] = Mc[0] * gx[2 * j] + Mc[1] * gx[2 * j + 1];
                            Mg[2 * j + 1] = Mc[2] * gx[2 * j] + Mc[3] * gx[2 * j + 1];
                            dt[j] = gr[j * 3 + 2] / h;
                        }
                    }
                    const double* ph = phi + static_cast<int64_t>(q) * nl;
                    for (int i = 0; i < nl; ++i) {
                        const double g0 = gx[2 * i], g1 = gx[2 * i + 1], di = dt[i];
                        double* Ai = &A[static_cast<size_t>(i) * nl];
                        double* Bi = &B[static_cast<size_t>(i) * nl];
                        for (int j = 0; j < nl; ++j) {
                            const double s = (g0 * Mg[2 * j] + g1 * Mg[2 * j + 1]) * dmu;
                            Ai[j] += s;
                            Bi[j] += (di * ph[j]) * dmu + s;
                        }
                    }
                }
                const int* ym = ymap + c * nl;
                const int* xm = xmap + c * nl;
                for (int i = 0; i < nl; ++i)
                    for (int j = 0; j < nl; ++j) {
                        const double a = A[static_cast<size_t>(i) * nl + j];
                        if (ym[i] >= 0 && ym[j] >= 0 && std::fabs(a) > 1e-15) {
                            const int64_t s = find_slot(rowptr, col, ym[i], ym[j]);
                            if (s >= 0) atomic_add(&val[s], a);
                        }
                        const double bv = B[static_cast<size_t>(i) * nl + j];
                        if (xm[i] >= 0 && ym[j] >= 0 && std::fabs(bv) > 1e-15) {
                            const int64_t s1 = find_slot(rowptr, col, nY + xm[i], ym[j]);
                            if (s1 >= 0) atomic_add(&val[s1], bv);
                            const int64_t s2 = find_slot(rowptr, col, ym[j], static_cast<int>(nY + xm[i]));
                            if (s2 >= 0) atomic_add(&val[s2], bv);
                        }
                    }
            }
        }
        best = std::min(best, now() - t0);
    }
    return best;
}

// Dense LU with partial pivoting, one solve and the residual acceptance of `npatch`
// patch systems of orders N[i] (row-major, in place).  The matrices are deterministic
// and diagonally weighted (the LU does the same work for any values); returns seconds
// and the largest acceptance statistic.
double cpu_patch_solves(int nthreads, int64_t npatch, const int* N, double* max_stat)
{
    omp_set_num_threads(nthreads);
    double worst = 0.0;
    const double t0 = now();
#pragma omp parallel reduction(max : worst)
    {
        std::vector<double> K, LU, b, x, r;
        std::vector<int> piv;
#pragma omp for schedule(dynamic, 4)
        for (int64_t pi = 0; pi < npatch; ++pi) {
            const int n = N[pi];
            K.resize(static_cast<size_t>(n) * n);
            uint64_t s = 0x9E3779B97F4A7C15ull * static_cast<uint64_t>(pi + 1);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                    K[static_cast<size_t>(i) * n + j] = (static_cast<double>(s >> 11) * 0x1.0p-53 - 0.5) +
                                                        (i == j ? 0.25 * n : 0.0);
                }
            LU = K;
            piv.resize(n);
            for (int k = 0; k < n; ++k) {
                int p = k;
                double m = std::fabs(LU[static_cast<size_t>(k) * n + k]);
                for (int i = k + 1; i < n; ++i) {
                    const double v = std::fabs(LU[static_cast<size_t>(i) * n + k]);
                    if (v > m) { m = v; p = i; }
                }
                piv[k] = p;
                if (p != k)
                    std::swap_ranges(&LU[static_cast<size_t>(k) * n], &LU[static_cast<size_t>(k) * n] + n,
                                     &LU[static_cast<size_t>(p) * n]);
                const double inv = 1.0 / LU[static_cast<size_t>(k) * n + k];
                const double* rk = &LU[static_cast<size_t>(k) * n];
                for (int i = k + 1; i < n; ++i) {
                    double* ri = &LU[static_cast<size_t>(i) * n];
                    const double l = ri[k] * inv;
                    ri[k] = l;
                    for (int j = k + 1; j < n; ++j) ri[j] -= l * rk[j];
                }
            }
            b.assign(n, 0.0);
            for (int i = 0; i < n; ++i) b[i] = std::sin(0.37 * (i + 1));
            x = b;
            for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < i; ++j) x[i] -= LU[static_cast<size_t>(i) * n + j] * x[j];
            for (int i = n - 1; i >= 0; --i) {
                for (int j = i + 1; j < n; ++j) x[i] -= LU[static_cast<size_t>(i) * n + j] * x[j];
                x[i] /= LU[static_cast<size_t>(i) * n + i];
            }
            double rm = 0.0, bm = 0.0;
            for (int i = 0; i < n; ++i) {
                double acc = b[i];
                for (int j = 0; j < n; ++j) acc -= K[static_cast<size_t>(i) * n + j] * x[j];
                rm = std::max(rm, std::fabs(acc));
                bm = std::max(bm, std::fabs(b[i]));
            }
            worst = std::max(worst, rm / std::max(1.0, bm));
        }
    }
    if (max_stat) *max_stat = worst;
    return now() - t0;
}

// MINRES (Paige-Saunders) on K x = b with blockdiag(q(D^-1 A_y) D^-1, S_diag^-1), q the
// Chebyshev polynomial of `degree` steps on [lmax / degree^2, lmax] (Gershgorin lmax),
// S_diag the Schur-complement diagonal of the X block: the preconditioner of
// apf_csr_minres_solve.  Runs at most max_it iterations, checks the true relative
// residual every check_interval; returns seconds, writes iterations and residual.
double cpu_minres(int nthreads, int64_t n, int64_t nY, const int64_t* rowptr, const int* col, const double* val,
                  const double* b, double* x, double tol, int max_it, int check_interval, int degree, int* iters,
                  double* rel)
{
    omp_set_num_threads(nthreads);
    const double t_start = now();
    auto spmv = [&](const double* xv, double* yv) {
#pragma omp parallel for schedule(static, 1024)
        for (int64_t i = 0; i < n; ++i) {
            double s = 0.0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += val[e] * xv[col[e]];
            yv[i] = s;
        }
    };
    std::vector<double> dinv(n), r1(n), r2(n), t(n), y(n), v(n), w(n), w1(n), w2(n), cx0(n), cx1(n), cd(n);
    double lmax = 0.0;
#pragma omp parallel for reduction(max : lmax)
    for (int64_t i = 0; i < nY; ++i) {
        double d = 0.0, a = 0.0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
            if (col[e] < nY) {
                a += std::fabs(val[e]);
                if (col[e] == i) d += val[e];
            }
        dinv[i] = d > 0.0 ? 1.0 / d : 1.0;
        if (d > 0.0) lmax = std::max(lmax, a / d);
    }
#pragma omp parallel for
    for (int64_t j = nY; j < n; ++j) {
        double s = 0.0;
        for (int64_t e = rowptr[j]; e < rowptr[j + 1]; ++e) {
            if (col[e] < nY) s += val[e] * val[e] * dinv[col[e]];
            else if (col[e] == j) s -= val[e];
        }
        dinv[j] = s > 0.0 ? 1.0 / s : 1.0;
    }
    const double ratio = std::max(4.0, static_cast<double>(degree) * degree);
    const double th = 0.5 * (1.0 + 1.0 / ratio) * lmax, de = 0.5 * (1.0 - 1.0 / ratio) * lmax;
    const double sigma = th / de;
    auto precond = [&](const double* rr, double* z) {
        // Y block: Chebyshev semi-iteration from zero (Saad, Alg. 12.1)
        double rho = 1.0 / sigma;
#pragma omp parallel for
        for (int64_t i = 0; i < nY; ++i) {
            cd[i] = dinv[i] * rr[i] / th;
            cx0[i] = cd[i];
        }
        for (int k = 1; k < degree; ++k) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            const double c1 = rn * rho, c2 = 2.0 * rn / de;
#pragma omp parallel for schedule(static, 1024)
            for (int64_t i = 0; i < nY; ++i) {
                double s = 0.0;
                for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e)
                    if (col[e] < nY) s += val[e] * cx0[col[e]];
                cd[i] = c1 * cd[i] + c2 * dinv[i] * (rr[i] - s);
                cx1[i] = cx0[i] + cd[i];
            }
            std::swap(cx0, cx1);
            rho = rn;
        }
#pragma omp parallel for
        for (int64_t i = 0; i < nY; ++i) z[i] = cx0[i];
#pragma omp parallel for
        for (int64_t j = nY; j < n; ++j) z[j] = dinv[j] * rr[j];
    };
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
        double s = 0.0;
#pragma omp parallel for reduction(+ : s)
        for (int64_t i = 0; i < n; ++i) s += a[i] * c[i];
        return s;
    };
    std::fill(x, x + n, 0.0);
    double bb = 0.0;
    for (int64_t i = 0; i < n; ++i) bb += b[i] * b[i];
    for (int64_t i = 0; i < n; ++i) r1[i] = b[i];
    r2 = r1;
    precond(r1.data(), y.data());
    double beta1 = std::sqrt(std::max(dot(r1, y), 0.0));
    double beta = beta1, oldb = 0.0, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    std::fill(w.begin(), w.end(), 0.0);
    std::fill(w2.begin(), w2.end(), 0.0);
    int it = 0;
    double relr = 1.0;
    while (it < max_it) {
        spmv(y.data(), t.data());
        const double s = 1.0 / beta, c = it >= 1 ? beta / oldb : 0.0;
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) {
            v[i] = s * y[i];
            t[i] = s * t[i] - c * r1[i];
        }
        const double alfa = dot(v, t);
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) t[i] -= (alfa / beta) * r2[i];
        precond(t.data(), y.data());
        oldb = beta;
        beta = std::sqrt(std::max(dot(t, y), 0.0));
        const double oldeps = epsln;
        const double delta = cs * dbar + sn * alfa;
        const double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        const double gamma = std::max(std::hypot(gbar, beta), 1e-300);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        std::swap(r1, r2);
        std::swap(r2, t);
#pragma omp parallel for
        for (int64_t i = 0; i < n; ++i) {
            const double wn = (v[i] - oldeps * w2[i] - delta * w[i]) / gamma;
            w1[i] = wn;
            x[i] += phi * wn;
        }
        std::swap(w1, w2);
        std::swap(w2, w);
        ++it;
        if (it % check_interval == 0 || it == max_it) {
            spmv(x, t.data());
            double rr = 0.0;
            for (int64_t i = 0; i < n; ++i) rr += (b[i] - t[i]) * (b[i] - t[i]);
            relr = std::sqrt(rr / bb);
            if (relr <= tol) break;
        }
    }
    if (iters) *iters = it;
    if (rel) *rel = relr;
    return now() - t_start;
}

}  // extern "C"
