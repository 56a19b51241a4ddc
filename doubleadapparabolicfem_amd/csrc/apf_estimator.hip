// Equilibrated-flux estimator on the device.
//
// Restates the estimator part of compute_adaptive_test_space_iteration
// (src/finite_element/time_slabs/time_slab_adaptive_loop.hpp:833-2026):
//  * one mixed patch problem per (slab, spatial vertex a)
//      [A  B^T 0] [s]   [f]    A = int M^-1 s_i.s_j, B = int q_i div s_j,
//      [B  0   c] [r] = [g]    f_i = -int psi_a grad(theta).s_i,
//      [0  c^T 0] [m]   [0]    g_i = int [psi_a (ell - d_t u) - grad psi_a . M grad(theta)] q_i
//    (time_slab_equilibrated_flux_reconstruction.hpp:2515, 3300-3380;
//     assembly/error_system/{mat_A_2d,mat_B_2d,vec_f_2d,vec_g_2d}.hpp,
//     dense_local_error_blocks_2d.hpp:719-768, 812-879, 1060-1143),
//  * per-cell flux / divergence-residual / reconstruction terms
//    (time_slab_equilibrated_flux_indicators.hpp:34-525, time_slab_error_norms.hpp:371-537)
//    and the X-marking components lambda_Y^2, trace^2
//    (algorithm/detail/indicator_utilities.hpp:339-392, 444-472).
//
// MI355X formulation (exact for piecewise-constant, time-independent M and
// constant ell on tensor meshes):
//  * every element integral is a geometry-free reference table contracted
//    with the cell metric (det, G = det J^-1 M J^-T, H = J^T M^-1 J / det);
//  * the time basis of the patch spaces is shifted Legendre, so the patch
//    system is block diagonal over the p+1 time modes with one shared matrix,
//    and on a tensor mesh that matrix is the same in every slab: it is built
//    and inverted once per spatial vertex and applied to all slabs' right-hand
//    sides as one dense product;
//  * the inverse is formed in FT (double, or float for the mixed path); the
//    mixed path recovers FP64 accuracy by iterative refinement with FP64
//    residuals of the FP64 patch matrix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/apf.h"
#include "apf_internal.hpp"

using namespace apf;

extern thread_local std::string g_last_error;

namespace {

struct EstParams {
    int dim, p, Ns, Nt, Nf, Nq, Nfc, Nfm, NV;
    int ns, m, nf;
    int64_t nY;
    double ell;
    int n_items, per_slab, chunk;
    int64_t scratch_stride;   // doubles per workgroup
    int nmax;
    int refine_steps;
    int n_lds;                // largest patch order whose inverse is kept in LDS
};

struct EstTables {
    const double* RTab;
    const double* Bhat;
    const double* qbar;
    const double* R;
    const double* Pv;
    const double* Q;
    const double* E;
    const double* Phi;
    const double* Khat;
    const double* dbary;
    const double* W;
    const double* Wd;
    const double* Lnorm;
    const double* Mt;
    const double* Igrad;
    const double* bw;
    const double* bphi;
};

struct PatchTopo {
    const int* cell_ptr;
    const int* cells;
    const int* lv;
    const int* fmap;
    const int8_t* fsign;
    const int* nflux;
    const int* nscal;
    const uint8_t* boundary;
    const int* order;
};

__device__ inline double block_max(double v, double* sh)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = sh[0];
        for (int i = 1; i < static_cast<int>(blockDim.x / 64); ++i) r = fmax(r, sh[i]);
        sh[0] = r;
    }
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ inline void atomic_max_pos(double* addr, double v)
{
    // v >= 0: the IEEE order of non-negative doubles matches their bit order
    atomicMax(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// argmax_{r in [k, n)} |A[r][k]| over the workgroup
template <typename FT>
__device__ inline int block_argmax_col(const FT* A, int n, int k, double* shv, int* shi)
{
    double best = -1.0;
    int bi = k;
    for (int r = k + threadIdx.x; r < n; r += blockDim.x) {
        const double v = fabs(static_cast<double>(A[static_cast<int64_t>(r) * n + k]));
        if (v > best) { best = v; bi = r; }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_down(best, off, 64);
        const int oi = __shfl_down(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    __syncthreads();
    if (l == 0) { shv[w] = best; shi[w] = bi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < static_cast<int>(blockDim.x / 64); ++i)
            if (shv[i] > shv[0] || (shv[i] == shv[0] && shi[i] < shi[0])) { shv[0] = shv[i]; shi[0] = shi[i]; }
    }
    __syncthreads();
    const int r = shi[0];
    __syncthreads();
    return r;
}

// In-place Gauss-Jordan inversion with partial (row) pivoting; returns the
// smallest pivot magnitude seen.
template <typename FT>
__device__ double gauss_jordan_inverse(FT* A, int n, int* perm, FT* fac, double* shv, int* shi)
{
    double minpiv = 1e300;
    for (int k = 0; k < n; ++k) {
        const int pr = block_argmax_col(A, n, k, shv, shi);
        if (threadIdx.x == 0) perm[k] = pr;
        if (pr != k)
            for (int j = threadIdx.x; j < n; j += blockDim.x) {
                const FT t = A[static_cast<int64_t>(k) * n + j];
                A[static_cast<int64_t>(k) * n + j] = A[static_cast<int64_t>(pr) * n + j];
                A[static_cast<int64_t>(pr) * n + j] = t;
            }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += blockDim.x) fac[i] = A[static_cast<int64_t>(i) * n + k];
        __syncthreads();
        const FT d = fac[k];
        minpiv = fmin(minpiv, fabs(static_cast<double>(d)));
        const FT inv = static_cast<FT>(1) / d;
        for (int j = threadIdx.x; j < n; j += blockDim.x) {
            const FT a = (j == k) ? static_cast<FT>(1) : A[static_cast<int64_t>(k) * n + j];
            A[static_cast<int64_t>(k) * n + j] = a * inv;
        }
        __syncthreads();
        const int64_t nn = static_cast<int64_t>(n) * n;
        for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) {
            const int i = static_cast<int>(e / n), j = static_cast<int>(e % n);
            if (i == k) continue;
            const FT a = (j == k) ? static_cast<FT>(0) : A[e];
            A[e] = a - fac[i] * A[static_cast<int64_t>(k) * n + j];
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const int pr = perm[k];
        if (pr != k)
            for (int i = threadIdx.x; i < n; i += blockDim.x) {
                const FT t = A[static_cast<int64_t>(i) * n + k];
                A[static_cast<int64_t>(i) * n + k] = A[static_cast<int64_t>(i) * n + pr];
                A[static_cast<int64_t>(i) * n + pr] = t;
            }
        __syncthreads();
    }
    return minpiv;
}

struct Metric2 {
    double det, G[4], H[4];
};

__device__ inline Metric2 metric_of(int dim, int c, const double* verts, const int* cells, const double* Mcell)
{
    const CellMetric g = cell_metric(dim, c, verts, cells, Mcell);
    Metric2 r;
    r.det = g.det;
    for (int i = 0; i < 4; ++i) { r.G[i] = g.G[i]; r.H[i] = g.H[i]; }
    return r;
}

// theta = lambda_tilde + u and u local coefficients of space-time cell (k, c):
// [Ns][Nt], Lagrange in time.
__device__ inline void gather_cell(const EstParams& P, int k, int c, const int* cell_free, const double* lt,
                                   const double* x, int m_, int l, double& th, double& uu)
{
    const int s = cell_free[c * P.Ns + m_];
    if (s < 0) { th = 0.0; uu = 0.0; return; }
    const double lam = lt[(static_cast<int64_t>(k) * P.Nt + l) * P.nf + s];
    const double u = x[P.nY + (static_cast<int64_t>(k) * P.p + l) * P.nf + s];
    th = lam + u;
    uu = u;
}

constexpr int PATCH_THREADS = 1024;

template <typename FT>
__global__ void __launch_bounds__(PATCH_THREADS) k_patch(EstParams P, EstTables T, PatchTopo topo, const double* verts,
                                               const int* cells, const double* Mcell, const int* cell_free,
                                               const double* times, const double* lt, const double* x,
                                               double* scratch, double* sig, double* patch_res)
{
    extern __shared__ double lds[];
    __shared__ double shv[PATCH_THREADS / 64];
    __shared__ int shi[PATCH_THREADS / 64];
    const int Nf = P.Nf, Nq = P.Nq, Nt = P.Nt, Ns = P.Ns, dim = P.dim;
    double* base = scratch + static_cast<int64_t>(blockIdx.x) * P.scratch_stride;
    const int nmax = P.nmax;
    const int C = P.chunk * Nt;
    double* Kd = base;                                            // nmax^2
    FT* Ki_g = reinterpret_cast<FT*>(base + static_cast<int64_t>(nmax) * nmax);  // nmax^2
    double* RHS = base + 2 * static_cast<int64_t>(nmax) * nmax;   // C * nmax (column major)
    double* X = RHS + static_cast<int64_t>(C) * nmax;
    double* Rs = X + static_cast<int64_t>(C) * nmax;
    int* perm = reinterpret_cast<int*>(Rs + static_cast<int64_t>(C) * nmax);        // nmax ints
    FT* fac = reinterpret_cast<FT*>(Rs + static_cast<int64_t>(C) * nmax + nmax);    // nmax FT
    double* stage = lds;  // [chunk][Ns][Nt][2]
    // the inverse lives in LDS when it fits (n <= P.n_lds), else in the scratch
    FT* Ki_l = reinterpret_cast<FT*>(lds + static_cast<int64_t>(P.chunk) * Ns * Nt * 2);

    for (int item = blockIdx.x; item < P.n_items; item += gridDim.x) {
        const int a = topo.order[P.per_slab ? item / P.m : item];
        const int k_begin = P.per_slab ? item % P.m : 0;
        const int k_end = P.per_slab ? k_begin + 1 : P.m;
        const int nF = topo.nflux[a], nS = topo.nscal[a];
        const bool interior = topo.boundary[a] == 0;
        const int n = nF + nS + (interior ? 1 : 0);
        const int cb = topo.cell_ptr[a], ce = topo.cell_ptr[a + 1];
        // ---- build the FP64 patch matrix (one time mode; shared by all modes and slabs)
        for (int64_t e = threadIdx.x; e < static_cast<int64_t>(n) * n; e += blockDim.x) Kd[e] = 0.0;
        __syncthreads();
        for (int ci = cb; ci < ce; ++ci) {
            const int c = topo.cells[ci];
            const Metric2 g = metric_of(dim, c, verts, cells, Mcell);
            const int* fm = topo.fmap + static_cast<int64_t>(ci) * Nf;
            const int8_t* fs = topo.fsign + static_cast<int64_t>(ci) * Nf;
            for (int e = threadIdx.x; e < Nf * Nf; e += blockDim.x) {
                const int i = e / Nf, j = e % Nf;
                const int fi = fm[i], fj = fm[j];
                if (fi < 0 || fj < 0) continue;
                double v = 0.0;
                for (int ab = 0; ab < dim * dim; ++ab)
                    v += g.H[(ab / dim) * 2 + (ab % dim)] * T.RTab[(static_cast<int64_t>(ab) * Nf + i) * Nf + j];
                Kd[static_cast<int64_t>(fi) * n + fj] += static_cast<double>(fs[i] * fs[j]) * v;
            }
            const int r0 = nF + (ci - cb) * Nq;
            for (int e = threadIdx.x; e < Nq * Nf; e += blockDim.x) {
                const int q = e / Nf, j = e % Nf;
                const int fj = fm[j];
                if (fj < 0) continue;
                const double v = fs[j] * T.Bhat[q * Nf + j];
                Kd[static_cast<int64_t>(r0 + q) * n + fj] = v;
                Kd[static_cast<int64_t>(fj) * n + r0 + q] = v;
            }
            if (interior)
                for (int q = threadIdx.x; q < Nq; q += blockDim.x) {
                    const double v = g.det * T.qbar[q];
                    Kd[static_cast<int64_t>(r0 + q) * n + (n - 1)] = v;
                    Kd[static_cast<int64_t>(n - 1) * n + r0 + q] = v;
                }
            __syncthreads();
        }
        // ---- invert in FT
        FT* Ki = (n <= P.n_lds) ? Ki_l : Ki_g;
        for (int64_t e = threadIdx.x; e < static_cast<int64_t>(n) * n; e += blockDim.x) Ki[e] = static_cast<FT>(Kd[e]);
        __syncthreads();
        gauss_jordan_inverse<FT>(Ki, n, perm, fac, shv, shi);
        // ---- right-hand sides, solve, refine, write contributions, chunk of slabs at a time
        double worst = 0.0;
        for (int k0 = k_begin; k0 < k_end; k0 += P.chunk) {
            const int nsl = min(P.chunk, k_end - k0);
            const int ncol = nsl * Nt;
            for (int64_t e = threadIdx.x; e < static_cast<int64_t>(ncol) * n; e += blockDim.x) RHS[e] = 0.0;
            __syncthreads();
            for (int ci = cb; ci < ce; ++ci) {
                const int c = topo.cells[ci];
                const int v = topo.lv[ci];
                const Metric2 g = metric_of(dim, c, verts, cells, Mcell);
                const int* fm = topo.fmap + static_cast<int64_t>(ci) * Nf;
                const int8_t* fs = topo.fsign + static_cast<int64_t>(ci) * Nf;
                // stage theta and u of this cell for every slab of the chunk
                for (int e = threadIdx.x; e < nsl * Ns * Nt; e += blockDim.x) {
                    const int sl = e / (Ns * Nt), r = e % (Ns * Nt);
                    const int mm = r / Nt, l = r % Nt;
                    double th, uu;
                    gather_cell(P, k0 + sl, c, cell_free, lt, x, mm, l, th, uu);
                    stage[e * 2] = th;
                    stage[e * 2 + 1] = uu;
                }
                __syncthreads();
                // grad psi_a pulled through the stiffness metric: gv = G grad^lam_v
                double gv[2] = {0.0, 0.0};
                for (int a2 = 0; a2 < dim; ++a2)
                    for (int b = 0; b < dim; ++b) gv[a2] += g.G[a2 * 2 + b] * T.dbary[v * dim + b];
                const int r0 = nF + (ci - cb) * Nq;
                const int nwork = nsl * (Nf + Nq) * Nt;
                for (int e = threadIdx.x; e < nwork; e += blockDim.x) {
                    const int sl = e / ((Nf + Nq) * Nt);
                    const int r = e % ((Nf + Nq) * Nt);
                    const int row = r / Nt, kk = r % Nt;
                    const int k = k0 + sl;
                    const double h = times[k + 1] - times[k];
                    const double scale = 1.0 / (h * T.Lnorm[kk]);
                    const double* st = stage + static_cast<int64_t>(sl) * Ns * Nt * 2;
                    if (row < Nf) {
                        const int fi = fm[row];
                        if (fi < 0) continue;
                        // F = -h R_v Theta W
                        double acc = 0.0;
                        for (int mm = 0; mm < Ns; ++mm) {
                            double tw = 0.0;
                            for (int l = 0; l < Nt; ++l) tw += st[(mm * Nt + l) * 2] * T.W[l * Nt + kk];
                            acc += T.R[(static_cast<int64_t>(v) * Nf + row) * Ns + mm] * tw;
                        }
                        RHS[static_cast<int64_t>(sl * Nt + kk) * n + fi] += -h * fs[row] * acc * scale;
                    } else {
                        const int q = row - Nf;
                        double acc = (kk == 0) ? P.ell * g.det * h * T.Pv[v * Nq + q] : 0.0;
                        double du = 0.0, dth = 0.0;
                        for (int mm = 0; mm < Ns; ++mm) {
                            double uw = 0.0, tw = 0.0;
                            for (int l = 0; l < Nt; ++l) {
                                uw += st[(mm * Nt + l) * 2 + 1] * T.Wd[l * Nt + kk];
                                tw += st[(mm * Nt + l) * 2] * T.W[l * Nt + kk];
                            }
                            du += T.Q[(static_cast<int64_t>(v) * Nq + q) * Ns + mm] * uw;
                            double eb = 0.0;
                            for (int b = 0; b < dim; ++b) eb += gv[b] * T.E[(static_cast<int64_t>(b) * Nq + q) * Ns + mm];
                            dth += eb * tw;
                        }
                        acc += -g.det * du - h * dth;
                        RHS[static_cast<int64_t>(sl * Nt + kk) * n + r0 + q] = acc * scale;
                    }
                }
                __syncthreads();
            }
            // ---- X = Ki RHS, residual Rs = RHS - Kd X (FP64), FT corrections.
            // Both matrices are symmetric: reading column i as row i keeps the
            // loads coalesced; each thread owns one row i and 4 columns.
            const int ncg = (ncol + 3) / 4;
            for (int e = threadIdx.x; e < ncg * n; e += blockDim.x) {
                const int cg = e / n, i = e % n;
                const int c0 = cg * 4;
                FT acc[4] = {0, 0, 0, 0};
                for (int j = 0; j < n; ++j) {
                    const FT kij = Ki[static_cast<int64_t>(j) * n + i];
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (c0 + q < ncol) acc[q] += kij * static_cast<FT>(RHS[static_cast<int64_t>(c0 + q) * n + j]);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (c0 + q < ncol) X[static_cast<int64_t>(c0 + q) * n + i] = static_cast<double>(acc[q]);
            }
            __syncthreads();
            const int steps = P.refine_steps;
            for (int it = 0; it <= steps; ++it) {
                // residual in FP64
                for (int e = threadIdx.x; e < ncg * n; e += blockDim.x) {
                    const int cg = e / n, i = e % n;
                    const int c0 = cg * 4;
                    double acc[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) acc[q] = (c0 + q < ncol) ? RHS[static_cast<int64_t>(c0 + q) * n + i] : 0.0;
                    for (int j = 0; j < n; ++j) {
                        const double kij = Kd[static_cast<int64_t>(j) * n + i];
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (c0 + q < ncol) acc[q] -= kij * X[static_cast<int64_t>(c0 + q) * n + j];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (c0 + q < ncol) Rs[static_cast<int64_t>(c0 + q) * n + i] = acc[q];
                }
                __syncthreads();
                if (it == steps) break;
                // correction in FT
                for (int e = threadIdx.x; e < ncg * n; e += blockDim.x) {
                    const int cg = e / n, i = e % n;
                    const int c0 = cg * 4;
                    FT acc[4] = {0, 0, 0, 0};
                    for (int j = 0; j < n; ++j) {
                        const FT kij = Ki[static_cast<int64_t>(j) * n + i];
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (c0 + q < ncol) acc[q] += kij * static_cast<FT>(Rs[static_cast<int64_t>(c0 + q) * n + j]);
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (c0 + q < ncol) X[static_cast<int64_t>(c0 + q) * n + i] += static_cast<double>(acc[q]);
                }
                __syncthreads();
            }
            // acceptance statistic: |K x - b|_inf / max(1, |b|_inf)
            // (dense_local_error_blocks_2d.hpp:923-949)
            double rmax = 0.0, bmax = 0.0;
            for (int64_t e = threadIdx.x; e < static_cast<int64_t>(ncol) * n; e += blockDim.x) {
                rmax = fmax(rmax, fabs(Rs[e]));
                bmax = fmax(bmax, fabs(RHS[e]));
            }
            rmax = block_max(rmax, shv);
            bmax = block_max(bmax, shv);
            worst = fmax(worst, rmax / fmax(1.0, bmax));
            // ---- write per (space-time cell, local vertex) flux coefficients
            for (int ci = cb; ci < ce; ++ci) {
                const int c = topo.cells[ci];
                const int v = topo.lv[ci];
                const int* fm = topo.fmap + static_cast<int64_t>(ci) * Nf;
                const int8_t* fs = topo.fsign + static_cast<int64_t>(ci) * Nf;
                for (int e = threadIdx.x; e < nsl * Nf * Nt; e += blockDim.x) {
                    const int sl = e / (Nf * Nt), r = e % (Nf * Nt);
                    const int i = r / Nt, kk = r % Nt;
                    const int fi = fm[i];
                    const double val = (fi >= 0) ? fs[i] * X[static_cast<int64_t>(sl * Nt + kk) * n + fi] : 0.0;
                    const int64_t cell = static_cast<int64_t>(k0 + sl) * P.ns + c;
                    sig[((cell * P.NV + v) * Nf + i) * Nt + kk] = val;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) atomic_max_pos(&patch_res[a], worst);
        __syncthreads();
    }
}

// Per space-time cell terms, in exact coefficient form:
//   flux^2  = sum_k h |L_k|^2 d_k^T A_c d_k,  d = s + coef(M grad theta) (mismatch in the flux basis)
//   resid^2 = sum_k det h |L_k|^2 |rho_k|^2,   rho = projection of ell - d_t u - div s
//   recon^2 = h sum_{l,l'} Mt_{ll'} e_l^T K_c e_l',  e = lambda - lambda_tilde
//   lam^2   = same with lambda,  trace^2 = int_{t=0} (u0 - u)^2 on the first layer.
__global__ void k_diagnostics(EstParams P, EstTables T, const double* verts, const int* cells, const double* Mcell,
                              const int* cell_free, const double* times, const double* lam, const double* lt,
                              const double* x, const double* sig, const double* u0b, int nb, double* flux,
                              double* resid, double* recon, double* lam2, double* trace)
{
    const int64_t cell = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (cell >= static_cast<int64_t>(P.m) * P.ns) return;
    const int c = static_cast<int>(cell % P.ns), k = static_cast<int>(cell / P.ns);
    const int Nf = P.Nf, Nq = P.Nq, Nt = P.Nt, Ns = P.Ns, dim = P.dim;
    const double h = times[k + 1] - times[k];
    const Metric2 g = metric_of(dim, c, verts, cells, Mcell);
    constexpr int MAXNS = 15, MAXNT = 5, MAXNF = 24, MAXNQ = 15;
    double th[MAXNS][MAXNT], uu[MAXNS][MAXNT], la[MAXNS][MAXNT], el[MAXNS][MAXNT];
    for (int mm = 0; mm < Ns; ++mm) {
        const int s = cell_free[c * Ns + mm];
        for (int l = 0; l < Nt; ++l) {
            if (s < 0) { th[mm][l] = uu[mm][l] = la[mm][l] = el[mm][l] = 0.0; continue; }
            const int64_t yi = (static_cast<int64_t>(k) * Nt + l) * P.nf + s;
            const double u = x[P.nY + (static_cast<int64_t>(k) * P.p + l) * P.nf + s];
            th[mm][l] = lt[yi] + u;
            uu[mm][l] = u;
            la[mm][l] = lam[yi];
            el[mm][l] = lam[yi] - lt[yi];
        }
    }
    // Legendre modes of theta
    double thL[MAXNS][MAXNT];
    for (int mm = 0; mm < Ns; ++mm)
        for (int kk = 0; kk < Nt; ++kk) {
            double a = 0.0;
            for (int l = 0; l < Nt; ++l) a += th[mm][l] * T.W[l * Nt + kk];
            thL[mm][kk] = a / T.Lnorm[kk];
        }
    double fl = 0.0, rs = 0.0;
    for (int kk = 0; kk < Nt; ++kk) {
        double d[MAXNF];
        for (int i = 0; i < Nf; ++i) {
            double s = 0.0;
            for (int v = 0; v < P.NV; ++v) s += sig[((cell * P.NV + v) * Nf + i) * Nt + kk];
            double cg = 0.0;
            for (int ab = 0; ab < dim * dim; ++ab) {
                const double gab = g.G[(ab / dim) * 2 + (ab % dim)];
                double t = 0.0;
                for (int mm = 0; mm < Ns; ++mm) t += T.Igrad[(static_cast<int64_t>(i) * dim * dim + ab) * Ns + mm] * thL[mm][kk];
                cg += gab * t;
            }
            d[i] = s + cg;
        }
        double q2 = 0.0;
        for (int i = 0; i < Nf; ++i) {
            double row = 0.0;
            for (int j = 0; j < Nf; ++j) {
                double a = 0.0;
                for (int ab = 0; ab < dim * dim; ++ab)
                    a += g.H[(ab / dim) * 2 + (ab % dim)] * T.RTab[(static_cast<int64_t>(ab) * Nf + i) * Nf + j];
                row += a * d[j];
            }
            q2 += d[i] * row;
        }
        fl += h * T.Lnorm[kk] * q2;
        // divergence residual coefficients
        double r2 = 0.0;
        for (int q = 0; q < Nq; ++q) {
            double rho = (kk == 0) ? P.ell * T.qbar[q] : 0.0;
            double dtu = 0.0;
            for (int mm = 0; mm < Ns; ++mm) {
                double uw = 0.0;
                for (int l = 0; l < Nt; ++l) uw += uu[mm][l] * T.Wd[l * Nt + kk];
                dtu += T.Phi[q * Ns + mm] * uw;
            }
            rho -= dtu / (T.Lnorm[kk] * h);
            double dv = 0.0;
            for (int j = 0; j < Nf; ++j) {
                double s = 0.0;
                for (int v = 0; v < P.NV; ++v) s += sig[((cell * P.NV + v) * Nf + j) * Nt + kk];
                dv += T.Bhat[q * Nf + j] * s;
            }
            rho -= dv / g.det;
            r2 += rho * rho;
        }
        rs += g.det * h * T.Lnorm[kk] * r2;
    }
    (void)MAXNQ;
    // Y-norm terms with the spatial stiffness K_c = sum_ab G_ab Khat_ab
    double lq = 0.0, eq = 0.0;
    for (int i = 0; i < Ns; ++i)
        for (int j = 0; j < Ns; ++j) {
            double kij = 0.0;
            for (int ab = 0; ab < dim * dim; ++ab)
                kij += g.G[(ab / dim) * 2 + (ab % dim)] * T.Khat[(static_cast<int64_t>(ab) * Ns + i) * Ns + j];
            double tl = 0.0, te = 0.0;
            for (int l = 0; l < Nt; ++l)
                for (int l2 = 0; l2 < Nt; ++l2) {
                    tl += la[i][l] * T.Mt[l * Nt + l2] * la[j][l2];
                    te += el[i][l] * T.Mt[l * Nt + l2] * el[j][l2];
                }
            lq += kij * tl;
            eq += kij * te;
        }
    double tr = 0.0;
    if (k == 0)
        for (int qb = 0; qb < nb; ++qb) {
            double uv = 0.0;
            for (int mm = 0; mm < Ns; ++mm) uv += uu[mm][0] * T.bphi[qb * Ns + mm];
            const double dd = u0b[static_cast<int64_t>(c) * nb + qb] - uv;
            tr += T.bw[qb] * g.det * dd * dd;
        }
    flux[cell] = fl;
    resid[cell] = rs;
    recon[cell] = h * eq;
    lam2[cell] = h * lq;
    trace[cell] = tr;
}

__global__ void k_sum5(int64_t n, const double* a, const double* b, const double* c, const double* d,
                       const double* e, double* part)
{
    __shared__ double sh[5][4];
    double s[5] = {0, 0, 0, 0, 0};
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        s[0] += a[i]; s[1] += b[i]; s[2] += c[i]; s[3] += d[i]; s[4] += e[i];
    }
    for (int q = 0; q < 5; ++q) {
        double v = s[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (threadIdx.x % 64 == 0) sh[q][threadIdx.x / 64] = v;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double v = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) v += sh[threadIdx.x][w];
        part[static_cast<int64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = v;
    }
}

__global__ void k_sum5_final(const double* part, int nparts, double* out)
{
    if (threadIdx.x < 5) {
        double v = 0.0;
        for (int i = 0; i < nparts; ++i) v += part[static_cast<int64_t>(threadIdx.x) * nparts + i];
        out[threadIdx.x] = v;
    }
}

__global__ void k_max_reduce(int n, const double* a, double* out)
{
    __shared__ double sh[8];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, a[i]);
    v = block_max(v, sh);
    if (threadIdx.x == 0) out[0] = v;
}

}  // namespace

namespace apf {

void estimator_alloc(System* sys)
{
    const Space& S = *sys->space;
    const RefTables& T = *sys->tables;
    EstimatorBuffers& B = sys->est;
    auto up = [&](const std::vector<double>& v) {
        double* p = dalloc<double>(sys, v.size());
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
        return p;
    };
    B.RTab = up(T.RTab);
    B.Bhat = up(T.Bhat);
    B.qbar = up(T.qbar);
    B.R = up(T.R);
    B.Pv = up(T.Pv);
    B.Q = up(T.Q);
    B.E = up(T.E);
    B.Phi = up(T.Phi);
    B.Khat = sys->d_Khat;
    B.Mhat = sys->d_Mhat;
    B.dbary = up(T.dbary);
    B.Wt = up(T.W);
    B.Wdt = up(T.Wd);
    B.Mdd = up(T.Mdd);
    B.Lnorm_ = up(T.Lnorm);
    B.Mt = up(T.Mt);
    B.Igrad = up(T.Igrad);
    const PatchTopology& P = S.patches;
    auto upi = [&](const std::vector<int>& v) {
        int* p = dalloc<int>(sys, v.size());
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice));
        return p;
    };
    B.p_cell_ptr = upi(P.cell_ptr);
    B.p_cells = upi(P.cells);
    B.p_lv = upi(P.local_vertex);
    B.p_fmap = upi(P.fmap);
    B.p_fsign = dalloc<int8_t>(sys, P.fsign.size());
    HIP_TRY(hipMemcpy(B.p_fsign, P.fsign.data(), P.fsign.size(), hipMemcpyHostToDevice));
    B.p_nflux = upi(P.n_flux);
    B.p_nscal = upi(P.n_scalar);
    B.p_boundary = dalloc<uint8_t>(sys, P.boundary.size());
    HIP_TRY(hipMemcpy(B.p_boundary, P.boundary.data(), P.boundary.size(), hipMemcpyHostToDevice));
    // largest patches first for load balance
    std::vector<int> order(P.n_patches);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return P.n_flux[a] + P.n_scalar[a] > P.n_flux[b] + P.n_scalar[b];
    });
    B.order = upi(order);
    int nmax = 0;
    for (int a = 0; a < P.n_patches; ++a) nmax = std::max(nmax, P.n_flux[a] + P.n_scalar[a] + 1);
    B.nmax = nmax;
    const int64_t ncell = static_cast<int64_t>(S.mesh.m()) * S.mesh.ns();
    B.sig = dalloc<double>(sys, static_cast<size_t>(ncell) * (S.dim + 1) * T.Nf * T.Nt);
    B.flux = dalloc<double>(sys, ncell);
    B.resid = dalloc<double>(sys, ncell);
    B.recon = dalloc<double>(sys, ncell);
    B.lam2 = dalloc<double>(sys, ncell);
    B.trace = dalloc<double>(sys, ncell);
    B.patch_res = dalloc<double>(sys, P.n_patches);
    B.lt = dalloc<double>(sys, S.nY);
    B.sums = dalloc<double>(sys, 8);
    B.chunk = std::min(S.mesh.m(), 32);
    const int C = B.chunk * T.Nt;
    B.scratch_per_wg = 2 * static_cast<size_t>(nmax) * nmax + 3 * static_cast<size_t>(C) * nmax + 2 * nmax + 8;
    B.n_wg = 0;  // allocated on first use
}

}  // namespace apf

extern "C" int apf_estimate(apf_system* sys, int prec, int refine_steps, apf_estimate_out* out, double* flux_by_cell,
                            double* resid_by_cell, double* recon_by_cell)
{
    if (!sys || !out) {
        g_last_error = "apf_estimate: null argument";
        return APF_BAD_ARG;
    }
    auto fail = [&](int code, const std::string& m) {
        g_last_error = m;
        sys->ctx->error = m;
        return code;
    };
    if (!sys->assembled) return fail(APF_BAD_ARG, "apf_estimate: assemble and solve first");
    if (prec != APF_FP64 && prec != APF_FP32_REFINE) return fail(APF_BAD_ARG, "apf_estimate: unknown precision");
    if (refine_steps < 0 || refine_steps > 8) return fail(APF_BAD_ARG, "apf_estimate: refine_steps must be 0..8");
    const Space& S = *sys->space;
    const RefTables& T = *sys->tables;
    if (S.p > 3) return fail(APF_BAD_ARG, "apf_estimate: the estimator supports p <= 3");
    EstimatorBuffers& B = sys->est;
    try {
        HIP_TRY(hipSetDevice(sys->ctx->device));
        hipStream_t st = sys->stream;
        if (B.n_wg == 0) {
            int dev = sys->ctx->device, cus = 0;
            HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            const int64_t items = S.patches.n_patches;
            B.n_wg = static_cast<int>(std::min<int64_t>(items, 4LL * cus));
            const size_t bytes = B.scratch_per_wg * sizeof(double) * B.n_wg;
            B.scratch = dalloc<double>(sys, bytes / sizeof(double));
        }
        EstParams P{};
        P.dim = S.dim;
        P.p = S.p;
        P.Ns = S.Ns;
        P.Nt = S.Nt;
        P.Nf = T.Nf;
        P.Nq = T.Nq;
        P.Nfc = T.Nfc;
        P.Nfm = T.Nfm;
        P.NV = S.dim + 1;
        P.ns = S.mesh.ns();
        P.m = S.mesh.m();
        P.nf = S.nf;
        P.nY = S.nY;
        P.ell = sys->ell;
        P.per_slab = 0;
        P.n_items = S.patches.n_patches;
        P.chunk = B.chunk;
        P.scratch_stride = static_cast<int64_t>(B.scratch_per_wg);
        P.nmax = B.nmax;
        P.refine_steps = (prec == APF_FP32_REFINE) ? refine_steps : 0;
        EstTables Tb{B.RTab, B.Bhat, B.qbar, B.R, B.Pv, B.Q, B.E, B.Phi, B.Khat, B.dbary, B.Wt, B.Wdt, B.Lnorm_, B.Mt,
                     B.Igrad, sys->d_bw, sys->d_bphi};
        PatchTopo topo{B.p_cell_ptr, B.p_cells, B.p_lv, B.p_fmap, B.p_fsign, B.p_nflux, B.p_nscal, B.p_boundary, B.order};
        // slab reconstruction: on a tensor mesh the slab space is Y restricted to
        // the slab, so lambda_tilde solves the lambda rows of the main system
        // slab by slab; with the main solve converged it equals lambda.
        HIP_TRY(hipMemcpyAsync(B.lt, sys->d_x, S.nY * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemsetAsync(B.patch_res, 0, S.patches.n_patches * sizeof(double), st));
        const size_t stage_bytes = static_cast<size_t>(B.chunk) * S.Ns * S.Nt * 2 * sizeof(double);
        const size_t ft_bytes = (prec == APF_FP64) ? sizeof(double) : sizeof(float);
        const size_t lds_cap = 150 * 1024;
        int n_lds = static_cast<int>(std::sqrt(static_cast<double>(lds_cap - stage_bytes) / ft_bytes));
        n_lds = std::min(n_lds, B.nmax);
        P.n_lds = n_lds;
        const size_t lds = stage_bytes + static_cast<size_t>(n_lds) * n_lds * ft_bytes;
        if (prec == APF_FP64)
            HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_patch<double>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        else
            HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_patch<float>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        HIP_TRY(hipEventRecord(sys->ev0, st));
        if (prec == APF_FP64)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch<double>), dim3(B.n_wg), dim3(PATCH_THREADS), lds, st, P, Tb, topo,
                               sys->d_verts, sys->d_cells, sys->d_Mcell, sys->d_cell_free, sys->d_times, B.lt,
                               sys->d_x, B.scratch, B.sig, B.patch_res);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch<float>), dim3(B.n_wg), dim3(PATCH_THREADS), lds, st, P, Tb, topo,
                               sys->d_verts, sys->d_cells, sys->d_Mcell, sys->d_cell_free, sys->d_times, B.lt,
                               sys->d_x, B.scratch, B.sig, B.patch_res);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(sys->ev1, st));
        const int64_t ncell = static_cast<int64_t>(P.m) * P.ns;
        hipEvent_t e2, e3;
        HIP_TRY(hipEventCreate(&e2));
        HIP_TRY(hipEventCreate(&e3));
        HIP_TRY(hipEventRecord(e2, st));
        hipLaunchKernelGGL(k_diagnostics, dim3(static_cast<unsigned>((ncell + 127) / 128)), dim3(128), 0, st, P, Tb,
                           sys->d_verts, sys->d_cells, sys->d_Mcell, sys->d_cell_free, sys->d_times, sys->d_x, B.lt,
                           sys->d_x, B.sig, sys->d_u0b, T.nb, B.flux, B.resid, B.recon, B.lam2, B.trace);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(e3, st));
        const int nparts = 256;
        hipLaunchKernelGGL(k_sum5, dim3(nparts), dim3(256), 0, st, ncell, B.flux, B.resid, B.recon, B.lam2, B.trace,
                           sys->d_part);
        hipLaunchKernelGGL(k_sum5_final, dim3(1), dim3(64), 0, st, sys->d_part, nparts, B.sums);
        hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, st, S.patches.n_patches, B.patch_res, B.sums + 5);
        HIP_TRY(hipGetLastError());
        double h[6];
        HIP_TRY(hipMemcpyAsync(h, B.sums, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float ms1 = 0.f, ms2 = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms1, sys->ev0, sys->ev1));
        HIP_TRY(hipEventElapsedTime(&ms2, e2, e3));
        (void)hipEventDestroy(e2);
        (void)hipEventDestroy(e3);
        sys->t_patch_ms = ms1;
        sys->t_diag_ms = ms2;
        out->flux2 = h[0];
        out->div_residual_l2 = std::sqrt(h[1]);
        out->recon2 = h[2];
        out->lambda_y2 = h[3];
        out->trace2 = h[4];
        out->eta2 = h[3] + h[4];
        out->y_est2 = h[0] + h[2];
        out->max_patch_residual = h[5];
        out->seconds_patches = ms1 * 1e-3;
        out->seconds_diagnostics = ms2 * 1e-3;
        if (flux_by_cell) HIP_TRY(hipMemcpy(flux_by_cell, B.flux, ncell * sizeof(double), hipMemcpyDeviceToHost));
        if (resid_by_cell) HIP_TRY(hipMemcpy(resid_by_cell, B.resid, ncell * sizeof(double), hipMemcpyDeviceToHost));
        if (recon_by_cell) HIP_TRY(hipMemcpy(recon_by_cell, B.recon, ncell * sizeof(double), hipMemcpyDeviceToHost));
        if (!(h[5] <= 1e-10))
            return fail(APF_RESIDUAL_TOO_LARGE,
                        "apf_estimate: patch residual " + std::to_string(h[5]) + " above the 1e-10 acceptance");
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(APF_HIP_ERROR, std::string("apf_estimate: ") + e.what());
    }
}
