// MINRES on a symmetric saddle-point matrix assembled by the caller.
//
// This is the boundary for a K that did not come from this library's own
// assembly: the compressed arrays that SaddlePreconditionedSolveDispatcher<Backend>::solve
// receives (src/linear_algebra/eigen_backend/preconditioners/solve_saddle_system.hpp:32-121;
// Eigen compressed storage, eigen_backend/sparse_matrix.hpp:29-37), the size of the
// lambda block (blocks.A.rows()) and optionally the graph-norm matrix Hhat of the
// preconditioner context (schur_approximations.hpp:24-28, :233-287).
//
// Preconditioner.  The reference applies P^-1 = blockdiag(A_y^-1, Hhat^-1) with sparse
// Cholesky factors (parabolic_graph_norm_minres.hpp:46-181).  Triangular solves are
// sequential and do not map to a GPU, so each block is replaced by a fixed Chebyshev
// polynomial in the Jacobi-scaled block:
//     P_B^-1 = q(D^-1 B) D^-1,   B = A_y (the leading n_lambda rows and columns of K)
//                                    or Hhat,
// the Chebyshev semi-iteration (Saad, Iterative Methods for Sparse Linear Systems,
// 2nd ed., Alg. 12.1) of `degree` steps from a zero start on the interval
// [lambda_max / degree^2, lambda_max], with lambda_max the Gershgorin bound of D^-1 B
// (an upper bound, so q(D^-1 B) D^-1 is symmetric positive definite).  Without Hhat
// the X block is the same polynomial in D_S^-1 S for the Schur complement
// S = -C + B diag(A_y)^-1 B^T, applied from K's blocks without forming S (D_S = diag(S),
// lambda_max from a Gershgorin bound of |B| diag(A_y)^-1 |B^T| + |C|).
// The polynomial is applied in FP32 (APF_FP32_REFINE) or FP64; the SpMV, the Lanczos
// recurrences and the inner products stay FP64.  The stopping rule is the reference's:
// the true relative residual |b - K x| / |b| is checked every `check_interval` iterations
// (iterative_residual_acceptance.hpp:117-216).
//
// Two entry points share one driver and the same kernels, so they give bit-identical
// results: apf_csr_minres_solve (host vectors; work buffers allocated once per matrix,
// on its first solve) and apf_csr_minres_solve_noalloc (device vectors; work buffers
// carved from the file-scope device array `scratch`, so the call allocates nothing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/apf.h"
#include "apf_internal.hpp"
#include "apf_solver_dev.hpp"

extern thread_local std::string g_last_error;

using apf::HipFailure;
using apf::MinresScalars;

namespace {

constexpr int kRowGrid = 1024;   // blocks of the grid-stride row kernels (= dot partials per block of rows)
constexpr int kRowLanes = 16;    // lanes per matrix row
constexpr size_t kPartDoubles = 4 * kRowGrid;

// Scratch of the allocation-free entry point: one array per GPU, held until the process
// ends (13 vectors of APF_NOALLOC_MAX_N doubles and the reduction partials, 27.3 MB).
// Its layout for n unknowns is work_layout(n).  Host code never uses the name `scratch`
// as an address: apf_internal_noalloc_init resolves the device address per device.
constexpr size_t kScratchBytes = 13 * sizeof(double) * static_cast<size_t>(APF_NOALLOC_MAX_N) +
                                 kPartDoubles * sizeof(double) + 4096;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[64] = {};   // device address of `scratch` per device (hipGetSymbolAddress)

int fail(apf_ctx* ctx, int code, const std::string& m)
{
    g_last_error = m;
    if (ctx) ctx->error = m;
    return code;
}

// Work buffers for n unknowns inside one allocation (offsets in bytes, 256-aligned).
struct WorkLayout {
    size_t r1, r2, t, y, v, w, w1, w2, dinv, cx0, cx1, cd, cw, part, scal, res, lam, total;
};

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

WorkLayout work_layout(int64_t n)
{
    WorkLayout L{};
    size_t o = 0;
    const size_t vec = align256(static_cast<size_t>(n) * sizeof(double));
    L.r1 = o; o += vec;
    L.r2 = o; o += vec;
    L.t = o; o += vec;
    L.y = o; o += vec;
    L.v = o; o += vec;
    L.w = o; o += vec;
    L.w1 = o; o += vec;
    L.w2 = o; o += vec;
    L.dinv = o; o += vec;
    L.cx0 = o; o += vec;  // Chebyshev iterates and direction (FP32 uses half of each)
    L.cx1 = o; o += vec;
    L.cd = o; o += vec;
    L.cw = o; o += vec;   // D_Y^-1 B^T v inside a Schur-complement product
    L.part = o; o += align256(kPartDoubles * sizeof(double));
    L.scal = o; o += align256(sizeof(MinresScalars));
    L.res = o; o += align256(4 * sizeof(double));
    L.lam = o; o += align256(4 * sizeof(double));
    L.total = o;
    return L;
}

// ------------------------------------------------------------------ kernels
__global__ void k_fill(int64_t n, double v, double* __restrict__ a)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        a[i] = v;
}

__global__ void k_copy(int64_t n, const double* __restrict__ a, double* __restrict__ b)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        b[i] = a[i];
}

__device__ inline double group_sum(double v)
{
    for (int off = kRowLanes / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kRowLanes);
    return v;
}

__device__ inline double block_max_nn(double v)
{
    // max of non-negative values over the block (a NaN counts as +inf)
    __shared__ double sh[16];
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_down(v, off, 64);
        v = (v != v || o != o) ? INFINITY : fmax(v, o);
    }
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < static_cast<int>(blockDim.x / 64); ++i) r = (sh[i] != sh[i]) ? INFINITY : fmax(r, sh[i]);
    __syncthreads();
    return r;
}

// Jacobi data of the block rows [0, rows) x columns [0, rows) of a CSR matrix:
// dinv_i = 1 / B_ii, and the Gershgorin bound max_i sum_j |B_ij| / B_ii of D^-1 B
// (block partial maxima).  A non-positive diagonal is reported as a partial of -1.
__global__ void __launch_bounds__(256) k_block_jacobi(int64_t rows, const int64_t* __restrict__ rowptr,
                                                      const int* __restrict__ col, const double* __restrict__ val,
                                                      double* __restrict__ dinv, double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    double gmax = 0.0;
    bool bad = false;
    for (int64_t i = g0; i < rows; i += gs) {
        double diag = 0.0, asum = 0.0;
        for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kRowLanes) {
            const int c = col[e];
            if (c < rows) {
                const double a = val[e];
                asum += fabs(a);
                if (c == i) diag += a;
            }
        }
        diag = group_sum(diag);
        asum = group_sum(asum);
        if (lane == 0) {
            if (!(diag > 0.0)) {
                bad = true;
                dinv[i] = 1.0;
            } else {
                dinv[i] = 1.0 / diag;
                gmax = fmax(gmax, asum / diag);
            }
        }
    }
    double m = block_max_nn(gmax);
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (threadIdx.x == 0) part[blockIdx.x] = anybad ? -1.0 : m;
}

// X-block diagonal of the Schur complement S = -C + B diag(A_y)^-1 B^T from the X rows
// of K = [A_y B^T; B C]:  S_jj = sum_{c < nY} K_jc^2 / K_cc - K_jj.
__global__ void __launch_bounds__(256) k_schur_diag(int64_t nY, int64_t n, const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ dinvY, double* __restrict__ dinvX,
                                                    double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    bool bad = false;
    for (int64_t j = nY + g0; j < n; j += gs) {
        double s = 0.0;
        for (int64_t e = rowptr[j] + lane; e < rowptr[j + 1]; e += kRowLanes) {
            const int c = col[e];
            const double a = val[e];
            if (c < nY) s += a * a * dinvY[c];
            else if (c == j) s -= a;
        }
        s = group_sum(s);
        if (lane == 0) {
            if (!(s > 0.0)) {
                bad = true;
                dinvX[j - nY] = 1.0;
            } else {
                dinvX[j - nY] = 1.0 / s;
            }
        }
    }
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (threadIdx.x == 0) part[blockIdx.x] = anybad ? -1.0 : 0.0;
}

// lam[slot] = max of the partials (a negative partial marks a non-positive diagonal)
__global__ void k_max_partials(const double* __restrict__ part, int nparts, double* __restrict__ lam, int slot)
{
    double v = 0.0;
    bool bad = false;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) {
        const double p = part[i];
        if (p < 0.0) bad = true;
        else v = (p != p) ? INFINITY : fmax(v, p);
    }
    v = block_max_nn(v);
    const int anybad = __syncthreads_or(bad ? 1 : 0);
    if (threadIdx.x == 0) lam[slot] = anybad ? -1.0 : v;
}

// One step of the Chebyshev semi-iteration for B y = r on the block rows [0, rows)
// (columns >= rows are not part of the block and are skipped), D = diag(B):
//   step 0:  d = (1/theta) D^-1 r,                          y1 = d
//   step k:  d = rho_k rho_{k-1} d + (2 rho_k / delta) D^-1 (r - B y_k),  y_{k+1} = y_k + d
// theta and delta scale with lambda_max (lam[slot]), read on the device; the host passes
// the lambda_max-free coefficients.  The last step writes z = y (FP64) and the block
// partials of z . r for the MINRES scalars.
template <typename CT>
__global__ void __launch_bounds__(256) k_cheb_step(int64_t rows, const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col, const double* __restrict__ val,
                                                   const double* __restrict__ dinv, const double* __restrict__ r,
                                                   const CT* __restrict__ yo, CT* __restrict__ yn,
                                                   CT* __restrict__ d, const double* __restrict__ lam, int slot,
                                                   int step, double c_dir, double c_res, int last,
                                                   double* __restrict__ z, double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    const CT cres = static_cast<CT>(c_res / lam[slot]);
    const CT cdir = static_cast<CT>(c_dir);
    double dot = 0.0;
    for (int64_t i = g0; i < rows; i += gs) {
        CT s = 0;
        if (step > 0) {
            for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kRowLanes) {
                const int c = col[e];
                if (c < rows) s += static_cast<CT>(val[e]) * yo[c];
            }
            for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kRowLanes);
        }
        if (lane == 0) {
            const CT ri = static_cast<CT>(r[i]);
            const CT di = static_cast<CT>(dinv[i]);
            CT dn, yv;
            if (step == 0) {
                dn = cres * di * ri;
                yv = dn;
            } else {
                dn = cdir * d[i] + cres * di * (ri - s);
                yv = yo[i] + dn;
            }
            if (last) {
                const double zi = static_cast<double>(yv);
                z[i] = zi;
                dot += zi * r[i];
            } else {
                d[i] = dn;
                yn[i] = yv;
            }
        }
    }
    const double bs = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// r_i = sum over the B^T entries (columns >= nY) of |K_ik|, Y rows
__global__ void __launch_bounds__(256) k_bt_absrow(int64_t nY, const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col, const double* __restrict__ val,
                                                   double* __restrict__ out)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    for (int64_t i = g0; i < nY; i += gs) {
        double a = 0.0;
        for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kRowLanes)
            if (col[e] >= nY) a += fabs(val[e]);
        a = group_sum(a);
        if (lane == 0) out[i] = a;
    }
}

// Gershgorin bound of D_S^-1 S on the X rows, S = -K_XX + K_XY D_Y^-1 K_YX:
//   sum_k |S_jk| <= sum_{i < nY} |K_ji| r_i / K_ii + sum_{k >= nY} |K_jk|   (block partial maxima)
__global__ void __launch_bounds__(256) k_schur_gersh(int64_t nY, int64_t n, const int64_t* __restrict__ rowptr,
                                                     const int* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ dinvY, const double* __restrict__ r,
                                                     const double* __restrict__ dinvX, double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    double gmax = 0.0;
    for (int64_t j = nY + g0; j < n; j += gs) {
        double a = 0.0;
        for (int64_t e = rowptr[j] + lane; e < rowptr[j + 1]; e += kRowLanes) {
            const int c = col[e];
            a += (c < nY) ? fabs(val[e]) * r[c] * dinvY[c] : fabs(val[e]);
        }
        a = group_sum(a);
        if (lane == 0) gmax = fmax(gmax, a * dinvX[j - nY]);
    }
    const double m = block_max_nn(gmax);
    if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// w_i = D_Y^-1 (K_YX v)_i on the Y rows; v holds X values at the column indices [nY, n)
template <typename CT>
__global__ void __launch_bounds__(256) k_schur_w(int64_t nY, const int64_t* __restrict__ rowptr,
                                                 const int* __restrict__ col, const double* __restrict__ val,
                                                 const double* __restrict__ dinvY, const CT* __restrict__ v,
                                                 CT* __restrict__ w)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    for (int64_t i = g0; i < nY; i += gs) {
        CT s = 0;
        for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kRowLanes) {
            const int c = col[e];
            if (c >= nY) s += static_cast<CT>(val[e]) * v[c];
        }
        for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kRowLanes);
        if (lane == 0) w[i] = s * static_cast<CT>(dinvY[i]);
    }
}

// One Chebyshev step on the X rows for S y = r with S y = K_XY w - K_XX y, w = D_Y^-1 K_YX y
// (k_schur_w of the same iterate); otherwise as k_cheb_step.  Vectors are indexed by K's
// row / column index (X part at [nY, n)).
template <typename CT>
__global__ void __launch_bounds__(256) k_cheb_x_step(int64_t nY, int64_t n, const int64_t* __restrict__ rowptr,
                                                     const int* __restrict__ col, const double* __restrict__ val,
                                                     const double* __restrict__ dinvX, const double* __restrict__ r,
                                                     const CT* __restrict__ w, const CT* __restrict__ yo,
                                                     CT* __restrict__ yn, CT* __restrict__ d,
                                                     const double* __restrict__ lam, int slot, int step, double c_dir,
                                                     double c_res, int last, double* __restrict__ z,
                                                     double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    const CT cres = static_cast<CT>(c_res / lam[slot]);
    const CT cdir = static_cast<CT>(c_dir);
    double dot = 0.0;
    for (int64_t j = nY + g0; j < n; j += gs) {
        CT s = 0;
        if (step > 0) {
            for (int64_t e = rowptr[j] + lane; e < rowptr[j + 1]; e += kRowLanes) {
                const int c = col[e];
                const CT a = static_cast<CT>(val[e]);
                s += (c < nY) ? a * w[c] : -a * yo[c];
            }
            for (int off = kRowLanes / 2; off > 0; off >>= 1) s += __shfl_down(s, off, kRowLanes);
        }
        if (lane == 0) {
            const CT rj = static_cast<CT>(r[j]);
            const CT dj = static_cast<CT>(dinvX[j - nY]);
            CT dn, yv;
            if (step == 0) {
                dn = cres * dj * rj;
                yv = dn;
            } else {
                dn = cdir * d[j] + cres * dj * (rj - s);
                yv = yo[j] + dn;
            }
            if (last) {
                const double zj = static_cast<double>(yv);
                z[j] = zj;
                dot += zj * r[j];
            } else {
                d[j] = dn;
                yn[j] = yv;
            }
        }
    }
    const double bs = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}


// ------------------------------------------------------------------ compact lambda block
// entries of the leading block (columns < rows) of each of the first `rows` rows
__global__ void k_blk_count(int64_t rows, const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                            int* __restrict__ cnt)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int n = 0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) n += (col[e] < rows) ? 1 : 0;
        cnt[i] = n;
    }
}

// copy those entries, in their order, into the compact CSR (bptr from the counts' prefix sums)
__global__ void k_blk_fill(int64_t rows, const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                           const double* __restrict__ val, const int64_t* __restrict__ bptr, int* __restrict__ bcol,
                           double* __restrict__ bval)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < rows;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int64_t o = bptr[i];
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            const int c = col[e];
            if (c < rows) {
                bcol[o] = c;
                bval[o] = val[e];
                ++o;
            }
        }
    }
}

// ------------------------------------------------------------------ block SPD solve
// y = B x and the block partials of x . y for the leading block rows [0, rows) x
// columns [0, rows) of a CSR matrix (columns >= rows are skipped).
__global__ void __launch_bounds__(256) k_block_spmv(int64_t rows, const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ col, const double* __restrict__ val,
                                                    const double* __restrict__ x, double* __restrict__ y,
                                                    double* __restrict__ part)
{
    const int lane = threadIdx.x % kRowLanes;
    const int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kRowLanes;
    const int64_t gs = static_cast<int64_t>(gridDim.x) * blockDim.x / kRowLanes;
    double dot = 0.0;
    for (int64_t i = g0; i < rows; i += gs) {
        double s = 0.0;
        for (int64_t e = rowptr[i] + lane; e < rowptr[i + 1]; e += kRowLanes) {
            const int c = col[e];
            if (c < rows) s += val[e] * x[c];
        }
        s = group_sum(s);
        if (lane == 0) {
            y[i] = s;
            dot += x[i] * s;
        }
    }
    const double bs = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// x += alpha p, r -= alpha q (alpha = rz / p.Ap from the device scalars), partials of r . r
__global__ void k_pcg_update(int64_t n, const double* __restrict__ sc, const double* __restrict__ p,
                             const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r,
                             double* __restrict__ part)
{
    // a cycle that has reached round-off can meet p.Ap <= 0 (or 0 / 0): it stops moving, and
    // the true residual after the batch decides
    const double alpha = (sc[1] > 0.0 && sc[0] > 0.0) ? sc[0] / sc[1] : 0.0;
    double rr = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        rr += ri * ri;
    }
    const double bs = block_sum(rr);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// p = z + beta p, beta = rz_new / rz_old; the first step writes p = z without reading p
// (fresh memory may hold any bit pattern, NaN included, and 0 * NaN is NaN)
__global__ void k_pcg_dir(int64_t n, const double* __restrict__ sc, int first, const double* __restrict__ z,
                          double* __restrict__ p)
{
    const double beta = (first || !(sc[0] > 0.0)) ? 0.0 : sc[2] / sc[0];
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = first ? z[i] : z[i] + beta * p[i];
}

// sc[slot] = sum of the partials (fixed order); slot 2 (new rz) moves to slot 0 after use by
// the direction kernel when `shift` is set
__global__ void k_scalar_from_partials(const double* __restrict__ part, int nparts, double* __restrict__ sc,
                                       int slot)
{
    const double v = reduce_partials(part, nparts);
    if (threadIdx.x == 0) sc[slot] = v;
}

__global__ void k_scalar_copy(double* __restrict__ sc, int from, int to)
{
    if (threadIdx.x == 0) sc[to] = sc[from];
}

// r = b - y and the partials of r . r and b . b (the true residual of the block solve)
__global__ void k_block_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ y,
                                 double* __restrict__ r, double* __restrict__ part)
{
    double rr = 0.0, bb = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double ri = b[i] - y[i];
        r[i] = ri;
        rr += ri * ri;
        bb += b[i] * b[i];
    }
    const double a = block_sum(rr);
    const double c = block_sum(bb);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = c;
    }
}

// ------------------------------------------------------------------ driver
struct Dev {
    int64_t n;
    const int64_t* rowptr;
    const int* col;
    const double* val;
};

// The Chebyshev coefficients of Alg. 12.1 on [lmax / ratio, lmax] divided by lmax:
// step 0: 1/theta; step k: rho_k rho_{k-1} and 2 rho_k / delta.
struct Cheb {
    std::vector<double> cdir, cres;
};

Cheb chebyshev_coefficients(int degree)
{
    const double ratio = std::max(4.0, static_cast<double>(degree) * degree);
    const double th = 0.5 * (1.0 + 1.0 / ratio), de = 0.5 * (1.0 - 1.0 / ratio);
    const double sigma = th / de;
    Cheb c;
    double rho = 1.0 / sigma;
    c.cdir.push_back(0.0);
    c.cres.push_back(1.0 / th);
    for (int k = 1; k < degree; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        c.cdir.push_back(rn * rho);
        c.cres.push_back(2.0 * rn / de);
        rho = rn;
    }
    return c;
}

template <typename T>
T* at(unsigned char* base, size_t off)
{
    return reinterpret_cast<T*>(base + off);
}

struct Outcome {
    int iterations = 0;
    bool converged = false;
    double rel = 0.0;
    double estimate = 0.0;
};

// d_x holds the initial guess when use_x0, else it is zeroed; on return it holds x.
// launch_opt: replay the iterations from a hipGraph and run the two preconditioner blocks
// on st and st2 (the default; APF_CSR_GRAPH=0 / APF_CSR_FORK=0 switch either off).  The
// allocation-free entry passes false: graphs, streams and events are resources it must not
// create.
// Ay (may be NULL): K's leading lambda block as a compact CSR, read by the lambda block's
// Chebyshev steps in place of K's lambda rows (which also hold B^T).
Outcome run_minres(hipStream_t st, hipStream_t st2, const Dev& K, int64_t nY, const Dev* H, unsigned char* work,
                   const double* d_b, double* d_x, bool use_x0, double tol, int max_it, int check_interval,
                   int degree, bool fp32, bool launch_opt, const Dev* Ay = nullptr)
{
    const int64_t n = K.n, nX = n - nY;
    const WorkLayout L = work_layout(n);
    double* r1 = at<double>(work, L.r1);
    double* r2 = at<double>(work, L.r2);
    double* t = at<double>(work, L.t);
    double* y = at<double>(work, L.y);
    double* v = at<double>(work, L.v);
    double* w = at<double>(work, L.w);
    double* w1 = at<double>(work, L.w1);
    double* w2 = at<double>(work, L.w2);
    double* dinv = at<double>(work, L.dinv);
    double* part = at<double>(work, L.part);
    MinresScalars* scal = at<MinresScalars>(work, L.scal);
    double* res = at<double>(work, L.res);
    double* lam = at<double>(work, L.lam);
    const dim3 grid(kRowGrid), block(256);
    const int G = kRowGrid;

    // ---- preconditioner set-up: Jacobi data and Gershgorin bounds of both blocks
    hipLaunchKernelGGL(k_block_jacobi, grid, block, 0, st, nY, K.rowptr, K.col, K.val, dinv, part);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, st, part, G, lam, 0);
    if (H) {
        hipLaunchKernelGGL(k_block_jacobi, grid, block, 0, st, nX, H->rowptr, H->col, H->val, dinv + nY, part);
        hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, st, part, G, lam, 1);
    } else {
        hipLaunchKernelGGL(k_schur_diag, grid, block, 0, st, nY, n, K.rowptr, K.col, K.val, dinv, dinv + nY, part);
        hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, st, part, G, lam, 1);  // flags a bad diagonal
        hipLaunchKernelGGL(k_bt_absrow, grid, block, 0, st, nY, K.rowptr, K.col, K.val, t);
        hipLaunchKernelGGL(k_schur_gersh, grid, block, 0, st, nY, n, K.rowptr, K.col, K.val, dinv, t, dinv + nY,
                           part + G);
        hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, st, part + G, G, lam, 2);
    }
    double hl[3] = {0.0, 0.0, 0.0};
    HIP_TRY(hipMemcpyAsync(hl, lam, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (!(hl[0] > 0.0) || !std::isfinite(hl[0]))
        throw std::runtime_error("the lambda block A_y is not positive definite (non-positive diagonal)");
    if (!(hl[1] >= 0.0) || !std::isfinite(hl[1]))
        throw std::runtime_error(H ? "Hhat is not positive definite (non-positive diagonal)"
                                   : "the Schur complement diagonal of the X block is not positive");
    if (!H && !(hl[2] > 0.0 && std::isfinite(hl[2])))
        throw std::runtime_error("the Schur complement of the X block has no positive Gershgorin bound");

    const Cheb cc = chebyshev_coefficients(degree);
    // z = P^-1 r with the partials of z . r in part[0, 2G)
    // (the work arrays from element `off` on, on stream `sb`)
    auto cheb = [&](int64_t rows, const Dev& B, const double* dv, const double* r, double* z, double* pp, int slot,
                    int64_t off, hipStream_t sb) {
        auto run = [&](auto tag) {
            using CT = decltype(tag);
            CT* cx0 = at<CT>(work, L.cx0) + off;
            CT* cx1 = at<CT>(work, L.cx1) + off;
            CT* cd = at<CT>(work, L.cd) + off;
            for (int k = 0; k < degree; ++k) {
                CT* yo = (k % 2 == 0) ? cx1 : cx0;
                CT* yn = (k % 2 == 0) ? cx0 : cx1;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cheb_step<CT>), grid, block, 0, sb, rows, B.rowptr, B.col, B.val,
                                   dv, r, yo, yn, cd, lam, slot, k, cc.cdir[k], cc.cres[k], k == degree - 1 ? 1 : 0,
                                   z, pp);
            }
        };
        if (fp32) run(float{});
        else run(double{});
    };
    // X block without Hhat: the Chebyshev polynomial in D_S^-1 S, S applied from K's blocks
    auto cheb_schur = [&](const double* r, double* z, double* pp, hipStream_t sb) {
        auto run = [&](auto tag) {
            using CT = decltype(tag);
            CT* cx0 = at<CT>(work, L.cx0);
            CT* cx1 = at<CT>(work, L.cx1);
            CT* cd = at<CT>(work, L.cd);
            CT* cw = at<CT>(work, L.cw);
            for (int k = 0; k < degree; ++k) {
                CT* yo = (k % 2 == 0) ? cx1 : cx0;
                CT* yn = (k % 2 == 0) ? cx0 : cx1;
                if (k > 0)
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_schur_w<CT>), grid, block, 0, sb, nY, K.rowptr, K.col, K.val,
                                       dinv, yo, cw);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cheb_x_step<CT>), grid, block, 0, sb, nY, n, K.rowptr, K.col,
                                   K.val, dinv + nY, r, cw, yo, yn, cd, lam, 2, k, cc.cdir[k], cc.cres[k],
                                   k == degree - 1 ? 1 : 0, z, pp);
            }
        };
        if (fp32) run(float{});
        else run(double{});
    };
    // The two blocks are independent: with a second stream the X block's chain runs beside
    // the lambda block's (its work arrays at elements [nY, n), its partials at [G, 2G)).
    struct Fork {
        hipStream_t s2 = nullptr;
        hipEvent_t a = nullptr, b = nullptr;
        ~Fork()
        {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } fk;
    bool fork = launch_opt && st2 != nullptr;
    if (const char* e = std::getenv("APF_CSR_FORK")) fork = fork && std::atoi(e) != 0;
    if (fork) {
        fk.s2 = st2;
        HIP_TRY(hipEventCreateWithFlags(&fk.a, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&fk.b, hipEventDisableTiming));
    }
    auto precond = [&](double* tt, const double* rr2, double* z) {
        if (rr2) hipLaunchKernelGGL(k_minres_upd, grid, block, 0, st, n, scal, tt, rr2);
        hipStream_t sx = st;
        if (fork) {
            HIP_TRY(hipEventRecord(fk.a, st));
            HIP_TRY(hipStreamWaitEvent(fk.s2, fk.a, 0));
            sx = fk.s2;
        }
        if (H) cheb(nX, *H, dinv + nY, tt + nY, z + nY, part + G, 1, nY, sx);
        else cheb_schur(tt, z, part + G, sx);
        cheb(nY, Ay ? *Ay : K, dinv, tt, z, part, 0, 0, st);
        if (fork) {
            HIP_TRY(hipEventRecord(fk.b, fk.s2));
            HIP_TRY(hipStreamWaitEvent(st, fk.b, 0));
        }
    };
    auto spmv = [&](const double* xv, double* yv) {
        const int64_t rows_per_block = (256 / 16) * 2;
        const int64_t g = (n + rows_per_block - 1) / rows_per_block;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spmv<16, 4, 2>), dim3(static_cast<unsigned>(g)), block, 0, st, n,
                           K.rowptr, K.col, K.val, xv, yv);
    };

    if (!use_x0) hipLaunchKernelGGL(k_fill, grid, block, 0, st, n, 0.0, d_x);
    spmv(d_x, r1);
    hipLaunchKernelGGL(k_axpby, grid, block, 0, st, n, 1.0, d_b, -1.0, r1);
    hipLaunchKernelGGL(k_copy, grid, block, 0, st, n, r1, r2);
    hipLaunchKernelGGL(k_fill, grid, block, 0, st, n, 0.0, w);
    hipLaunchKernelGGL(k_fill, grid, block, 0, st, n, 0.0, w2);
    hipLaunchKernelGGL(k_copy, grid, block, 0, st, n, r1, t);
    precond(t, nullptr, y);
    hipLaunchKernelGGL(k_minres_init, dim3(1), dim3(256), 0, st, scal, part, 2 * G);
    auto true_residual = [&]() {
        spmv(d_x, t);
        hipLaunchKernelGGL(k_residual_parts, grid, block, 0, st, n, d_b, t, part);
        hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, st, part, G, res);
        double h[2];
        HIP_TRY(hipMemcpyAsync(h, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return (h[1] > 0.0) ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
    };
    Outcome o;
    o.rel = true_residual();
    o.converged = o.rel <= tol;
    auto iteration = [&]() {
        spmv(y, t);
        hipLaunchKernelGGL(k_minres_vec1, grid, block, 0, st, n, scal, y, t, r1, v, part);
        hipLaunchKernelGGL(k_minres_alfa, dim3(1), dim3(256), 0, st, scal, part, G);
        precond(t, r2, y);
        hipLaunchKernelGGL(k_minres_scal2, dim3(1), dim3(256), 0, st, scal, part, 2 * G);
        std::swap(r1, r2);   // r1 <- r2, r2 <- t, t <- old r1
        std::swap(r2, t);
        hipLaunchKernelGGL(k_minres_vec3, grid, block, 0, st, n, scal, v, w2, w, w1, d_x);
        double* wt = w1;
        w1 = w2;
        w2 = w;
        w = wt;
    };
    // The buffer rotations have period 3: three iterations from a given rotation phase are one
    // fixed launch sequence.  One graph per phase (iterations done mod 3) is captured when first
    // needed and replayed; a batch of check_interval iterations replays its phase's graph
    // batch / 3 times and launches the remaining batch % 3 iterations eagerly, so the true
    // residual is checked at the same iterations as without graphs.
    struct GraphGuard {
        hipGraphExec_t e[3] = {nullptr, nullptr, nullptr};
        hipGraph_t g[3] = {nullptr, nullptr, nullptr};
        ~GraphGuard()
        {
            for (int q = 0; q < 3; ++q) {
                if (e[q]) (void)hipGraphExecDestroy(e[q]);
                if (g[q]) (void)hipGraphDestroy(g[q]);
            }
        }
    } gg;
    bool use_graph = launch_opt;
    if (const char* e = std::getenv("APF_CSR_GRAPH")) use_graph = use_graph && std::atoi(e) != 0;
    while (!o.converged && o.iterations < max_it) {
        const int batch = std::min(check_interval, max_it - o.iterations);
        int eager = batch;
        if (use_graph && batch >= 3) {
            const int ph = o.iterations % 3;
            if (!gg.e[ph]) {
                HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                for (int q = 0; q < 3; ++q) iteration();
                HIP_TRY(hipStreamEndCapture(st, &gg.g[ph]));
                HIP_TRY(hipGraphInstantiate(&gg.e[ph], gg.g[ph], nullptr, nullptr, 0));
            }
            for (int b = 0; b + 3 <= batch; b += 3) HIP_TRY(hipGraphLaunch(gg.e[ph], st));
            eager = batch % 3;
        }
        for (int b = 0; b < eager; ++b) iteration();
        o.iterations += batch;
        HIP_TRY(hipGetLastError());
        o.rel = true_residual();
        o.converged = o.rel <= tol;
        if (!std::isfinite(o.rel)) break;
    }
    MinresScalars hs{};
    HIP_TRY(hipMemcpyAsync(&hs, scal, sizeof(hs), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    o.estimate = hs.beta1 > 0 ? hs.phibar / hs.beta1 : 0.0;
    return o;
}

int check_solve_args(apf_ctx* ctx, const apf_csr_matrix* K, int64_t nY, const apf_csr_matrix* H, double tol,
                     int max_it, int check_interval, int degree, int prec, const char* who)
{
    const std::string w(who);
    if (!ctx || !K) return fail(ctx, APF_BAD_ARG, w + ": null argument");
    if (K->ctx != ctx || (H && H->ctx != ctx)) return fail(ctx, APF_BAD_ARG, w + ": matrix of another context");
    if (nY < 1 || nY >= K->n) return fail(ctx, APF_BAD_ARG, w + ": n_lambda must lie in [1, n)");
    if (H && H->n != K->n - nY) return fail(ctx, APF_BAD_ARG, w + ": Hhat must have n - n_lambda rows");
    if (!(tol > 0.0) || max_it < 1 || check_interval < 1)
        return fail(ctx, APF_BAD_ARG, w + ": bad tolerance / iteration limits");
    if (degree < 1 || degree > 64) return fail(ctx, APF_BAD_ARG, w + ": Chebyshev degree must be in [1, 64]");
    if (prec != APF_FP64 && prec != APF_FP32_REFINE) return fail(ctx, APF_BAD_ARG, w + ": unknown precision");
    return APF_OK;
}

void fill_stats(apf_minres_stats* stats, const Outcome& o, double secs)
{
    if (!stats) return;
    stats->iterations = o.iterations;
    stats->converged = o.converged ? 1 : 0;
    stats->true_relative_residual = o.rel;
    stats->estimate = o.estimate;
    stats->seconds = secs;
}

}  // namespace

// ------------------------------------------------------------------ C ABI
extern "C" {

int apf_csr_matrix_create(apf_ctx* ctx, int64_t n, const int* outer, const int* inner, const double* values,
                          int storage, apf_csr_matrix** out)
{
    if (!ctx || !outer || !out || n < 1) return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: bad argument");
    if (storage != APF_CSR_FULL && storage != APF_CSR_TRIANGLE)
        return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: unknown storage");
    try {
        const int64_t nnz_in = outer[n] - outer[0];
        if (outer[0] != 0 || nnz_in < 0) return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: bad outer index array");
        if (nnz_in > 0 && (!inner || !values)) return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: null arrays");
        for (int64_t i = 0; i < n; ++i)
            if (outer[i + 1] < outer[i]) return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: outer index decreases");
        for (int64_t e = 0; e < nnz_in; ++e)
            if (inner[e] < 0 || inner[e] >= n)
                return fail(ctx, APF_BAD_ARG, "apf_csr_matrix_create: inner index out of range");
        // rows of the matrix used: the stored entries, or the symmetric completion of those
        // with inner >= outer (Eigen ColMajor storage: the lower triangle, Eigen::Lower)
        std::vector<int64_t> rowptr(n + 1, 0);
        std::vector<int> col;
        std::vector<double> val;
        if (storage == APF_CSR_FULL) {
            for (int64_t i = 0; i <= n; ++i) rowptr[i] = outer[i];
            col.assign(inner, inner + nnz_in);
            val.assign(values, values + nnz_in);
        } else {
            for (int64_t i = 0; i < n; ++i)
                for (int64_t e = outer[i]; e < outer[i + 1]; ++e) {
                    const int j = inner[e];
                    if (j < i) continue;
                    ++rowptr[i + 1];
                    if (j != i) ++rowptr[j + 1];
                }
            for (int64_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
            col.resize(rowptr[n]);
            val.resize(rowptr[n]);
            std::vector<int64_t> pos(rowptr.begin(), rowptr.end() - 1);
            for (int64_t i = 0; i < n; ++i)
                for (int64_t e = outer[i]; e < outer[i + 1]; ++e) {
                    const int j = inner[e];
                    if (j < i) continue;
                    col[pos[i]] = j;
                    val[pos[i]++] = values[e];
                    if (j != i) {
                        col[pos[j]] = static_cast<int>(i);
                        val[pos[j]++] = values[e];
                    }
                }
        }
        auto M = std::make_unique<apf_csr_matrix>();
        M->ctx = ctx;
        M->n = n;
        M->nnz = rowptr[n];
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipMalloc(&M->rowptr, (n + 1) * sizeof(int64_t)));
        HIP_TRY(hipMalloc(&M->col, std::max<int64_t>(1, M->nnz) * sizeof(int)));
        HIP_TRY(hipMalloc(&M->val, std::max<int64_t>(1, M->nnz) * sizeof(double)));
        HIP_TRY(hipMemcpy(M->rowptr, rowptr.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        if (M->nnz) {
            HIP_TRY(hipMemcpy(M->col, col.data(), M->nnz * sizeof(int), hipMemcpyHostToDevice));
            HIP_TRY(hipMemcpy(M->val, val.data(), M->nnz * sizeof(double), hipMemcpyHostToDevice));
        }
        *out = M.release();
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_matrix_create: ") + e.what());
    } catch (const std::exception& e) {
        return fail(ctx, APF_BAD_ARG, std::string("apf_csr_matrix_create: ") + e.what());
    }
}

void apf_csr_matrix_destroy(apf_csr_matrix* M)
{
    if (!M) return;
    (void)hipSetDevice(M->ctx->device);
    if (M->owned) {
        (void)hipFree(M->rowptr);
        (void)hipFree(M->col);
        (void)hipFree(M->val);
    }
    if (M->work) (void)hipFree(M->work);
    delete M;
}

int apf_csr_matrix_info(const apf_csr_matrix* M, int64_t* n, int64_t* nnz)
{
    if (!M) return fail(nullptr, APF_BAD_ARG, "apf_csr_matrix_info: null matrix");
    if (n) *n = M->n;
    if (nnz) *nnz = M->nnz;
    return APF_OK;
}

int apf_csr_matrix_download(const apf_csr_matrix* M, int64_t* rowptr, int* col, double* values)
{
    if (!M) return fail(nullptr, APF_BAD_ARG, "apf_csr_matrix_download: null matrix");
    try {
        HIP_TRY(hipSetDevice(M->ctx->device));
        HIP_TRY(hipStreamSynchronize(M->ctx->stream));
        if (rowptr) HIP_TRY(hipMemcpy(rowptr, M->rowptr, (M->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (col && M->nnz) HIP_TRY(hipMemcpy(col, M->col, M->nnz * sizeof(int), hipMemcpyDeviceToHost));
        if (values && M->nnz) HIP_TRY(hipMemcpy(values, M->val, M->nnz * sizeof(double), hipMemcpyDeviceToHost));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(M->ctx, APF_HIP_ERROR, std::string("apf_csr_matrix_download: ") + e.what());
    }
}

int apf_csr_minres_solve(apf_ctx* ctx, apf_csr_matrix* K, int64_t n_lambda, const apf_csr_matrix* Hhat,
                         const double* b, double* x, const double* x0, double tol, int max_it, int check_interval,
                         int cheb_degree, int precond_prec, apf_minres_stats* stats)
{
    if (int c = check_solve_args(ctx, K, n_lambda, Hhat, tol, max_it, check_interval, cheb_degree, precond_prec,
                                 "apf_csr_minres_solve"))
        return c;
    if (!b || !x) return fail(ctx, APF_BAD_ARG, "apf_csr_minres_solve: null vector");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t st = ctx->stream;
        const WorkLayout L = work_layout(K->n);
        // work: the layout plus b and x
        const size_t vec = align256(static_cast<size_t>(K->n) * sizeof(double));
        const size_t need = L.total + 2 * vec;
        if (K->work_bytes < need) {
            if (K->work) HIP_TRY(hipFree(K->work));
            K->work = nullptr;
            K->work_bytes = 0;
            HIP_TRY(hipMalloc(&K->work, need));
            K->work_bytes = need;
        }
        unsigned char* work = static_cast<unsigned char*>(K->work);
        double* d_b = at<double>(work, L.total);
        double* d_x = at<double>(work, L.total + vec);
        HIP_TRY(hipMemcpyAsync(d_b, b, K->n * sizeof(double), hipMemcpyHostToDevice, st));
        if (x0) HIP_TRY(hipMemcpyAsync(d_x, x0, K->n * sizeof(double), hipMemcpyHostToDevice, st));
        const Dev Kd{K->n, K->rowptr, K->col, K->val};
        Dev Hd{};
        if (Hhat) Hd = Dev{Hhat->n, Hhat->rowptr, Hhat->col, Hhat->val};
        // K's lambda block as a compact CSR for the lambda block's Chebyshev steps: K's lambda rows
        // also hold B^T, which those steps would read and skip (APF_CSR_COMPACT=0: read K's rows)
        struct Blk {
            int64_t* ptr = nullptr;
            int* col = nullptr;
            double* val = nullptr;
            int* cnt = nullptr;
            ~Blk()
            {
                if (ptr) (void)hipFree(ptr);
                if (col) (void)hipFree(col);
                if (val) (void)hipFree(val);
                if (cnt) (void)hipFree(cnt);
            }
        } blk;
        bool compact = true;
        if (const char* e = std::getenv("APF_CSR_COMPACT")) compact = std::atoi(e) != 0;
        Dev Ad{};
        if (compact) {
            const int64_t nY = n_lambda;
            HIP_TRY(hipMalloc(&blk.cnt, nY * sizeof(int)));
            hipLaunchKernelGGL(k_blk_count, dim3(kRowGrid), dim3(256), 0, st, nY, K->rowptr, K->col, blk.cnt);
            std::vector<int> cnt(nY);
            HIP_TRY(hipMemcpyAsync(cnt.data(), blk.cnt, nY * sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            std::vector<int64_t> bptr(nY + 1, 0);
            for (int64_t i = 0; i < nY; ++i) bptr[i + 1] = bptr[i] + cnt[i];
            const int64_t nnzA = bptr[nY];
            HIP_TRY(hipMalloc(&blk.ptr, (nY + 1) * sizeof(int64_t)));
            HIP_TRY(hipMalloc(&blk.col, std::max<int64_t>(1, nnzA) * sizeof(int)));
            HIP_TRY(hipMalloc(&blk.val, std::max<int64_t>(1, nnzA) * sizeof(double)));
            HIP_TRY(hipMemcpyAsync(blk.ptr, bptr.data(), (nY + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_blk_fill, dim3(kRowGrid), dim3(256), 0, st, nY, K->rowptr, K->col, K->val, blk.ptr,
                               blk.col, blk.val);
            HIP_TRY(hipGetLastError());
            Ad = Dev{nY, blk.ptr, blk.col, blk.val};
        }
        const Outcome o = run_minres(st, ctx->stream2, Kd, n_lambda, Hhat ? &Hd : nullptr, work, d_b, d_x,
                                     x0 != nullptr, tol, max_it, check_interval, cheb_degree,
                                     precond_prec == APF_FP32_REFINE, true, compact ? &Ad : nullptr);
        HIP_TRY(hipMemcpyAsync(x, d_x, K->n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        fill_stats(stats, o, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        if (!o.converged)
            return fail(ctx, APF_NOT_CONVERGED,
                        "apf_csr_minres_solve: not converged after " + std::to_string(o.iterations) +
                            " iterations, true relative residual " + std::to_string(o.rel));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_minres_solve: ") + e.what());
    } catch (const std::exception& e) {
        return fail(ctx, APF_BAD_ARG, std::string("apf_csr_minres_solve: ") + e.what());
    }
}

}  // extern "C"


// ------------------------------------------------------------------ block SPD solve: C ABI
namespace {
struct BlockWork {
    double *b = nullptr, *x = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
    double *c0 = nullptr, *c1 = nullptr, *cd = nullptr, *part = nullptr, *sc = nullptr;
    std::vector<void*> all;
    ~BlockWork()
    {
        for (void* v : all) (void)hipFree(v);
    }
    double* get(size_t count)
    {
        void* v = nullptr;
        HIP_TRY(hipMalloc(&v, std::max<size_t>(1, count) * sizeof(double)));
        all.push_back(v);
        return static_cast<double*>(v);
    }
};

int check_block(apf_ctx* ctx, const apf_csr_matrix* A, int64_t n_block, const char* who)
{
    if (!ctx || !A) return fail(ctx, APF_BAD_ARG, std::string(who) + ": null argument");
    if (A->ctx != ctx) return fail(ctx, APF_BAD_ARG, std::string(who) + ": matrix of another context");
    if (n_block < 1 || n_block > A->n) return fail(ctx, APF_BAD_ARG, std::string(who) + ": n_block must lie in [1, n]");
    return APF_OK;
}
}  // namespace

extern "C" int apf_csr_spmv(apf_ctx* ctx, const apf_csr_matrix* A, const double* x, double* y)
{
    if (int c = check_block(ctx, A, A ? A->n : 1, "apf_csr_spmv")) return c;
    if (!x || !y) return fail(ctx, APF_BAD_ARG, "apf_csr_spmv: null vector");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        BlockWork W;
        double* dx = W.get(A->n);
        double* dy = W.get(A->n);
        HIP_TRY(hipMemcpyAsync(dx, x, A->n * sizeof(double), hipMemcpyHostToDevice, st));
        const int64_t rows_per_block = (256 / 16) * 2;
        const int64_t g = (A->n + rows_per_block - 1) / rows_per_block;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_spmv<16, 4, 2>), dim3(static_cast<unsigned>(g)), dim3(256), 0, st, A->n,
                           A->rowptr, A->col, A->val, dx, dy);
        HIP_TRY(hipMemcpyAsync(y, dy, A->n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_spmv: ") + e.what());
    }
}

extern "C" int apf_csr_block_energy(apf_ctx* ctx, const apf_csr_matrix* A, int64_t n_block, const double* d,
                                    double* out)
{
    if (int c = check_block(ctx, A, n_block, "apf_csr_block_energy")) return c;
    if (!d || !out) return fail(ctx, APF_BAD_ARG, "apf_csr_block_energy: null argument");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        BlockWork W;
        double* dd = W.get(n_block);
        double* dy = W.get(n_block);
        double* part = W.get(kRowGrid);
        double* sc = W.get(4);
        HIP_TRY(hipMemcpyAsync(dd, d, n_block * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_block_spmv, dim3(kRowGrid), dim3(256), 0, st, n_block, A->rowptr, A->col, A->val, dd,
                           dy, part);
        hipLaunchKernelGGL(k_scalar_from_partials, dim3(1), dim3(256), 0, st, part, kRowGrid, sc, 0);
        HIP_TRY(hipMemcpyAsync(out, sc, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_block_energy: ") + e.what());
    }
}

extern "C" int apf_csr_block_pcg_solve(apf_ctx* ctx, const apf_csr_matrix* A, int64_t n_block, const double* b,
                                       double* x, const double* x0, double tol, int max_it, int check_interval,
                                       int cheb_degree, apf_minres_stats* stats)
{
    if (int c = check_block(ctx, A, n_block, "apf_csr_block_pcg_solve")) return c;
    if (!b || !x) return fail(ctx, APF_BAD_ARG, "apf_csr_block_pcg_solve: null vector");
    if (!(tol > 0.0) || max_it < 1 || check_interval < 1)
        return fail(ctx, APF_BAD_ARG, "apf_csr_block_pcg_solve: bad tolerance / iteration limits");
    if (cheb_degree < 1 || cheb_degree > 64)
        return fail(ctx, APF_BAD_ARG, "apf_csr_block_pcg_solve: Chebyshev degree must be in [1, 64]");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        hipStream_t st = ctx->stream;
        const int64_t n = n_block;
        const int G = kRowGrid;
        BlockWork W;
        double* db = W.get(n);
        double* dx = W.get(n);
        double* r = W.get(n);
        double* z = W.get(n);
        double* p = W.get(n);
        double* q = W.get(n);
        double* dinv = W.get(n);
        double* c0 = W.get(n);
        double* c1 = W.get(n);
        double* cd = W.get(n);
        double* part = W.get(2 * G);
        double* sc = W.get(8);   // [0] rz, [1] p.Ap, [2] new rz, [3] r.r, [4] lambda_max, [5..6] residual sums
        HIP_TRY(hipMemcpyAsync(db, b, n * sizeof(double), hipMemcpyHostToDevice, st));
        if (x0) HIP_TRY(hipMemcpyAsync(dx, x0, n * sizeof(double), hipMemcpyHostToDevice, st));
        else hipLaunchKernelGGL(k_fill, dim3(G), dim3(256), 0, st, n, 0.0, dx);
        // Jacobi data and the Gershgorin bound of D^-1 A
        hipLaunchKernelGGL(k_block_jacobi, dim3(G), dim3(256), 0, st, n, A->rowptr, A->col, A->val, dinv, part);
        hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, st, part, G, sc, 4);
        double lm = 0.0;
        HIP_TRY(hipMemcpyAsync(&lm, sc + 4, sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (!(lm > 0.0) || !std::isfinite(lm))
            return fail(ctx, APF_BAD_ARG, "apf_csr_block_pcg_solve: the block is not positive definite "
                                          "(non-positive diagonal)");
        const Cheb cc = chebyshev_coefficients(cheb_degree);
        // z = q(D^-1 A) D^-1 r, with the partials of z . r (new rz) into sc[2]
        auto precond = [&]() {
            for (int k = 0; k < cheb_degree; ++k) {
                double* yo = (k % 2 == 0) ? c1 : c0;
                double* yn = (k % 2 == 0) ? c0 : c1;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cheb_step<double>), dim3(G), dim3(256), 0, st, n, A->rowptr,
                                   A->col, A->val, dinv, r, yo, yn, cd, sc, 4, k, cc.cdir[k], cc.cres[k],
                                   k == cheb_degree - 1 ? 1 : 0, z, part);
            }
            hipLaunchKernelGGL(k_scalar_from_partials, dim3(1), dim3(256), 0, st, part, G, sc, 2);
        };
        // r = b - A x; returns ||r|| / ||b||
        auto true_residual = [&]() {
            hipLaunchKernelGGL(k_block_spmv, dim3(G), dim3(256), 0, st, n, A->rowptr, A->col, A->val, dx, q, part);
            hipLaunchKernelGGL(k_block_residual, dim3(G), dim3(256), 0, st, n, db, q, r, part);
            hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(256), 0, st, part, G, sc + 5);
            double h[2];
            HIP_TRY(hipMemcpyAsync(h, sc + 5, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            return (h[1] > 0.0) ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
        };
        Outcome o;
        o.rel = true_residual();
        o.converged = o.rel <= tol;
        int restarts = 0;
        auto iteration = [&]() {
            hipLaunchKernelGGL(k_block_spmv, dim3(G), dim3(256), 0, st, n, A->rowptr, A->col, A->val, p, q, part);
            hipLaunchKernelGGL(k_scalar_from_partials, dim3(1), dim3(256), 0, st, part, G, sc, 1);
            hipLaunchKernelGGL(k_pcg_update, dim3(G), dim3(256), 0, st, n, sc, p, q, dx, r, part + G);
            precond();
            hipLaunchKernelGGL(k_pcg_dir, dim3(G), dim3(256), 0, st, n, sc, 0, z, p);
            hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(64), 0, st, sc, 2, 0);
        };
        struct GraphGuard {
            hipGraphExec_t e = nullptr;
            hipGraph_t g = nullptr;
            ~GraphGuard()
            {
                if (e) (void)hipGraphExecDestroy(e);
                if (g) (void)hipGraphDestroy(g);
            }
        } gg;
        bool use_graph = true;
        if (const char* e = std::getenv("APF_CSR_GRAPH")) use_graph = std::atoi(e) != 0;
        while (!o.converged && o.iterations < max_it) {
            // one CG cycle from the current true residual
            precond();
            hipLaunchKernelGGL(k_pcg_dir, dim3(G), dim3(256), 0, st, n, sc, 1, z, p);
            hipLaunchKernelGGL(k_scalar_copy, dim3(1), dim3(64), 0, st, sc, 2, 0);
            double rr0 = -1.0;
            bool cycle_done = false;
            while (!cycle_done && o.iterations < max_it) {
                const int batch = std::min(check_interval, max_it - o.iterations);
                if (use_graph && batch == check_interval) {
                    // a full batch: the same launches on the same buffers every time, captured
                    // once and replayed
                    if (!gg.e) {
                        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                        for (int it = 0; it < batch; ++it) iteration();
                        HIP_TRY(hipStreamEndCapture(st, &gg.g));
                        HIP_TRY(hipGraphInstantiate(&gg.e, gg.g, nullptr, nullptr, 0));
                    }
                    HIP_TRY(hipGraphLaunch(gg.e, st));
                } else {
                    for (int it = 0; it < batch; ++it) iteration();
                }
                o.iterations += batch;
                hipLaunchKernelGGL(k_scalar_from_partials, dim3(1), dim3(256), 0, st, part + G, G, sc, 3);
                double rr = 0.0;
                HIP_TRY(hipMemcpyAsync(&rr, sc + 3, sizeof(double), hipMemcpyDeviceToHost, st));
                HIP_TRY(hipStreamSynchronize(st));
                if (!std::isfinite(rr)) throw std::runtime_error("the block solve diverged (non-finite residual)");
                if (rr0 < 0.0) rr0 = rr;
                // recursive residual at a tenth of the tolerance, or stalled: confirm on the true one
                const double bn2 = [&] {
                    double h[2];
                    HIP_TRY(hipMemcpyAsync(h, sc + 5, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
                    HIP_TRY(hipStreamSynchronize(st));
                    return h[1] > 0.0 ? h[1] : 1.0;
                }();
                if (std::sqrt(rr / bn2) <= 0.1 * tol || o.iterations >= max_it) cycle_done = true;
            }
            o.rel = true_residual();
            o.converged = o.rel <= tol;
            if (!o.converged && ++restarts > 8) break;
        }
        HIP_TRY(hipMemcpyAsync(x, dx, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        fill_stats(stats, o, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        if (!o.converged)
            return fail(ctx, APF_NOT_CONVERGED,
                        "apf_csr_block_pcg_solve: not converged after " + std::to_string(o.iterations) +
                            " iterations, true relative residual " + std::to_string(o.rel));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_block_pcg_solve: ") + e.what());
    } catch (const std::exception& e) {
        return fail(ctx, APF_BAD_ARG, std::string("apf_csr_block_pcg_solve: ") + e.what());
    }
}

// Resolve the device address of `scratch` on the context's device (once per device).
// Called from apf_ctx_create with that device current.
int apf_internal_noalloc_init(int device)
{
    if (device < 0 || device >= 64) return APF_BAD_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return APF_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        g_last_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return APF_HIP_ERROR;
    }
    g_scratch_ptr[device] = ptr;
    return APF_OK;
}

extern "C" int apf_csr_minres_solve_noalloc(apf_ctx* ctx, const apf_csr_matrix* K, int64_t n_lambda,
                                            const double* d_b, double* d_x, int use_x0, double tol, int max_it,
                                            int check_interval, int cheb_degree, int precond_prec,
                                            apf_minres_stats* stats)
{
    if (int c = check_solve_args(ctx, K, n_lambda, nullptr, tol, max_it, check_interval, cheb_degree, precond_prec,
                                 "apf_csr_minres_solve_noalloc"))
        return c;
    if (!d_b || !d_x) return fail(ctx, APF_BAD_ARG, "apf_csr_minres_solve_noalloc: null vector");
    if (K->n > APF_NOALLOC_MAX_N)
        return fail(ctx, APF_BAD_ARG,
                    "apf_csr_minres_solve_noalloc: " + std::to_string(K->n) + " unknowns exceed the scratch limit of " +
                        std::to_string(APF_NOALLOC_MAX_N));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mutex);
        base = (ctx->device >= 0 && ctx->device < 64) ? g_scratch_ptr[ctx->device] : nullptr;
    }
    if (!base) return fail(ctx, APF_HIP_ERROR, "apf_csr_minres_solve_noalloc: scratch not resolved for this device");
    if (work_layout(K->n).total > kScratchBytes)
        return fail(ctx, APF_BAD_ARG, "apf_csr_minres_solve_noalloc: work layout exceeds the scratch array");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        const Dev Kd{K->n, K->rowptr, K->col, K->val};
        const Outcome o = run_minres(ctx->stream, nullptr, Kd, n_lambda, nullptr, static_cast<unsigned char*>(base),
                                     d_b, d_x,
                                     use_x0 != 0, tol, max_it, check_interval, cheb_degree,
                                     precond_prec == APF_FP32_REFINE, false);
        fill_stats(stats, o, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        if (!o.converged)
            return fail(ctx, APF_NOT_CONVERGED,
                        "apf_csr_minres_solve_noalloc: not converged after " + std::to_string(o.iterations) +
                            " iterations, true relative residual " + std::to_string(o.rel));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_csr_minres_solve_noalloc: ") + e.what());
    } catch (const std::exception& e) {
        return fail(ctx, APF_BAD_ARG, std::string("apf_csr_minres_solve_noalloc: ") + e.what());
    }
}

// ------------------------------------------------------------------ device buffers
// Plain device memory for callers that hand device vectors to
// apf_csr_minres_solve_noalloc (the allocation happens here, outside that call).
extern "C" int apf_device_alloc(apf_ctx* ctx, int64_t bytes, void** out)
{
    if (!ctx || !out || bytes < 1) return fail(ctx, APF_BAD_ARG, "apf_device_alloc: bad argument");
    *out = nullptr;
    if (hipSetDevice(ctx->device) != hipSuccess || hipMalloc(out, static_cast<size_t>(bytes)) != hipSuccess)
        return fail(ctx, APF_HIP_ERROR, "apf_device_alloc: hipMalloc failed");
    return APF_OK;
}

extern "C" void apf_device_free(apf_ctx* ctx, void* p)
{
    if (!ctx || !p) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(p);
}

extern "C" int apf_device_upload(apf_ctx* ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return fail(ctx, APF_BAD_ARG, "apf_device_upload: bad argument");
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice) != hipSuccess)
        return fail(ctx, APF_HIP_ERROR, "apf_device_upload: copy failed");
    return APF_OK;
}

extern "C" int apf_device_download(apf_ctx* ctx, void* dst, const void* src, int64_t bytes)
{
    if (!ctx || !dst || !src || bytes < 0) return fail(ctx, APF_BAD_ARG, "apf_device_download: bad argument");
    if (hipSetDevice(ctx->device) != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess ||
        hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(ctx, APF_HIP_ERROR, "apf_device_download: copy failed");
    return APF_OK;
}
