// Device side of the space-time multigrid preconditioner (see apf_mg.hpp).
//
// All spatial work is multi-column: a block Z[nf][ld] holds one column per
// time dof (Y block, shift 0) or per time eigenmode (X block, shift s_i), so
// every sparse product streams the spatial matrix once for all columns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "apf_internal.hpp"
#include "apf_mg.hpp"
#include "apf_mg_dev.hpp"

namespace apf {

namespace {

// Value of lane t (wave-uniform) of v, through scalar registers.
__device__ inline int lane_value(int v, int t) { return __builtin_amdgcn_readlane(v, t); }
__device__ inline float lane_value(float v, int t)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), t));
}
__device__ inline double lane_value(double v, int t)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffull), t));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), t));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// One wavefront per sparse row, lanes along the columns of a block X[n][ld]:
// NP passes of 64 columns are handled together (cc holds each lane's column
// per pass).  The row's entries (column and NV values) are staged 64 at a time,
// one per lane, and read back through scalar registers; the column loads of
// MG_U entries are issued before their FMAs, so a row costs a few memory
// latencies instead of one per entry.  coef(v, x, acc) applies one entry with
// values v to the loaded x, in the row's entry order.  Entries past the row's
// end repeat its last column with zero values.
constexpr int MG_U = 8;

template <int NP, int NV, typename R, typename Coef>
__device__ inline void row_entries(int b, int e, const int* __restrict__ col, const R* const (&vals)[NV],
                                   const R* __restrict__ X, int ld, const int (&cc)[NP], R (&acc)[NP], Coef coef)
{
    const int lane = threadIdx.x % 64;
    for (int e0 = b; e0 < e; e0 += 64) {
        const int my = e0 + lane;
        const int cnt = min(64, e - e0);
        const int cj = col[min(my, e - 1)];
        R v[NV];
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = (my < e) ? vals[k][my] : R(0);
        auto step = [&](int t0, auto width) {
            constexpr int U = decltype(width)::value;
            R xv[U][NP];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const R* xr = X + static_cast<int64_t>(lane_value(cj, t0 + u)) * ld;
#pragma unroll
                for (int q = 0; q < NP; ++q) xv[u][q] = xr[cc[q]];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                R w[NV];
#pragma unroll
                for (int k = 0; k < NV; ++k) w[k] = lane_value(v[k], t0 + u);
                coef(w, xv[u], acc);
            }
        };
        int t0 = 0;
        for (; t0 + MG_U <= cnt; t0 += MG_U) step(t0, std::integral_constant<int, MG_U>{});
        for (; t0 < cnt; t0 += 4) step(t0, std::integral_constant<int, 4>{});
    }
}

// Rows of (K + s_c M) X over columns [0, nc); row_apply hands the product of
// each column to the epilogue f(c, value).  Lanes past the last column read
// column nc - 1 and are not handed on.
template <int NP, typename R, typename F>
__device__ inline void row_apply(int row, const int* __restrict__ ptr, const int* __restrict__ col,
                                 const R* __restrict__ K, const R* __restrict__ M, const R* __restrict__ sig, int nc,
                                 int ld, const R* __restrict__ X, F f)
{
    const int lane = threadIdx.x % 64;
    const int b = ptr[row], e = ptr[row + 1];
    const R* const vals[2] = {K, M};
    for (int cg = 0; cg < nc; cg += NP * 64) {
        R acc[NP];
        R sg[NP];
        int cc[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            acc[q] = 0;
            cc[q] = min(cg + q * 64 + lane, nc - 1);
            sg[q] = sig[cc[q]];
        }
        row_entries<NP, 2>(b, e, col, vals, X, ld, cc, acc, [&](const R(&w)[2], const R(&x)[NP], R(&a)[NP]) {
#pragma unroll
            for (int q = 0; q < NP; ++q) a[q] += (w[0] + sg[q] * w[1]) * x[q];
        });
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int c = cg + q * 64 + lane;
            if (c < nc) f(c, acc[q]);
        }
    }
}

// Y = (K + s_c M) X   (mode 0)   or   Y = B - (K + s_c M) X   (mode 1), columns [0, nc).
template <typename R, int NP>
__global__ void __launch_bounds__(256) k_mg_apply(int nf, const int* __restrict__ ptr, const int* __restrict__ col,
                                                  const R* __restrict__ K, const R* __restrict__ M,
                                                  const R* __restrict__ sig, int nc, int ld, const R* __restrict__ X,
                                                  const R* __restrict__ B, R* __restrict__ Y, int mode)
{
    const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    if (row >= nf) return;
    row_apply<NP>(row, ptr, col, K, M, sig, nc, ld, X, [&](int c, R v) {
        const int64_t o = static_cast<int64_t>(row) * ld + c;
        Y[o] = mode ? B[o] - v : v;
    });
}

// Chebyshev start from x: r = B - A x, d = D^-1 r / theta (one pass)
template <typename R, int NP>
__global__ void __launch_bounds__(256)
    k_mg_resid_start(int nf, const int* __restrict__ ptr, const int* __restrict__ col, const R* __restrict__ K,
                     const R* __restrict__ M, const R* __restrict__ sig, int nc, int ld, const R* __restrict__ X,
                     const R* __restrict__ B, R* __restrict__ r, R* __restrict__ d, const R* __restrict__ dK,
                     const R* __restrict__ dM, R inv_theta)
{
    const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    if (row >= nf) return;
    const R dk = dK[row], dm = dM[row];
    row_apply<NP>(row, ptr, col, K, M, sig, nc, ld, X, [&](int c, R v) {
        const int64_t o = static_cast<int64_t>(row) * ld + c;
        const R rv = B[o] - v;
        r[o] = rv;
        d[o] = rv * inv_theta / (dk + sig[c] * dm);
    });
}

// One Chebyshev step with its product: x += d ; r -= A d ; dn = c1 d + c2 D^-1 r.
// A d reads d of the neighbouring rows, so the new direction goes to a second
// buffer dn (the caller swaps them).  The first step from a zero start (r0 =
// B, x0 = 0) reads B for r and writes x = d.
template <typename R, int NP>
__global__ void __launch_bounds__(256)
    k_mg_cheb_fused(int nf, const int* __restrict__ ptr, const int* __restrict__ col, const R* __restrict__ K,
                    const R* __restrict__ M, const R* __restrict__ sig, int nc, int ld, const R* __restrict__ d,
                    R* __restrict__ x, R* __restrict__ r, R* __restrict__ dn, const R* __restrict__ dK,
                    const R* __restrict__ dM, R c1, R c2, int last, const R* __restrict__ B0)
{
    const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    if (row >= nf) return;
    const R dk = dK[row], dm = dM[row];
    row_apply<NP>(row, ptr, col, K, M, sig, nc, ld, d, [&](int c, R v) {
        const int64_t o = static_cast<int64_t>(row) * ld + c;
        const R dv = d[o];
        R rv;
        if (B0) {
            x[o] = dv;
            rv = B0[o] - v;
        } else {
            x[o] += dv;
            rv = r[o] - v;
        }
        r[o] = rv;
        if (!last) dn[o] = c1 * dv + c2 * rv / (dk + sig[c] * dm);
    });
}

// zero start: d = D^-1 B / theta (x0 = 0 and r0 = B are folded into the first step)
template <typename R>
__global__ void k_cheb_start_zero(int nf, int nc, int ld, const R* __restrict__ dK, const R* __restrict__ dM,
                                  const R* __restrict__ sig, R inv_theta, const R* __restrict__ B, R* __restrict__ d)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(nf) * nc) return;
    const int i = static_cast<int>(t / nc), c = static_cast<int>(t % nc);
    const int64_t o = static_cast<int64_t>(i) * ld + c;
    d[o] = B[o] * inv_theta / (dK[i] + sig[c] * dM[i]);
}

// Y[:, 0:nc] = X[:, 0:nc]
template <typename R>
__global__ void k_block_copy(int nf, int nc, int ld, const R* __restrict__ X, R* __restrict__ Y)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(nf) * nc) return;
    const int i = static_cast<int>(t / nc), c = static_cast<int>(t % nc);
    const int64_t o = static_cast<int64_t>(i) * ld + c;
    Y[o] = X[o];
}

// Y = A X for a plain CSR (prolongation / restriction), optionally accumulated
template <typename R, int NP>
__global__ void __launch_bounds__(256) k_csr_mm(int nrows, const int* __restrict__ ptr, const int* __restrict__ col,
                                                const R* __restrict__ val, int nc, int ldx, const R* __restrict__ X,
                                                int ldy, R* __restrict__ Y, int accumulate)
{
    const int row = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
    const int lane = threadIdx.x % 64;
    if (row >= nrows) return;
    const int b = ptr[row], e = ptr[row + 1];
    const R* const vals[1] = {val};
    for (int cg = 0; cg < nc; cg += NP * 64) {
        R acc[NP];
        int cc[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            acc[q] = 0;
            cc[q] = min(cg + q * 64 + lane, nc - 1);
        }
        row_entries<NP, 1>(b, e, col, vals, X, ldx, cc, acc, [&](const R(&w)[1], const R(&x)[NP], R(&a)[NP]) {
#pragma unroll
            for (int q = 0; q < NP; ++q) a[q] += w[0] * x[q];
        });
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int c = cg + q * 64 + lane;
            if (c < nc) {
                const int64_t o = static_cast<int64_t>(row) * ldy + c;
                Y[o] = accumulate ? Y[o] + acc[q] : acc[q];
            }
        }
    }
}

// coarsest level: X[:,c] = inv[shift_idx[c]] B[:,c]
template <typename R>
__global__ void k_coarse_solve(int n, int nc, int ld, const R* __restrict__ inv, int64_t blk_total,
                               const int* __restrict__ comp_start, const int* __restrict__ node_comp,
                               const int64_t* __restrict__ blk_off, const int* __restrict__ sidx,
                               const R* __restrict__ B, R* __restrict__ X)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<int64_t>(n) * nc) return;
    const int i = static_cast<int>(t / nc), c = static_cast<int>(t % nc);
    const int comp = node_comp[i];
    const int s = comp_start[comp], b = comp_start[comp + 1] - s;
    const R* A = inv + static_cast<int64_t>(sidx[c]) * blk_total + blk_off[comp] + static_cast<int64_t>(i - s) * b;
    R acc = 0;
    for (int j = 0; j < b; ++j) acc += A[j] * B[static_cast<int64_t>(s + j) * ld + c];
    X[static_cast<int64_t>(i) * ld + c] = acc;
}

// ---- fused coarse tail of the V-cycle
// Levels tail_from..L-1 of one column per workgroup: the right-hand side and
// solution of every level and three work vectors of the first level stay in
// LDS; the level matrices are read from global memory (L2-resident: a few
// hundred KB shared by all columns).  The arithmetic is that of the per-level
// kernels (Chebyshev start, fused steps, restriction, coarsest dense solve,
// prolongation, post-smoothing) in the same order per row.
constexpr int MG_TAIL_THREADS = 1024;
constexpr int MG_TAIL_MAX_NU = 8;
constexpr size_t MG_TAIL_LDS_CAP = 60 * 1024;  // within the default dynamic LDS limit

template <typename R>
struct MGTailLevel {
    int nf;
    int off;  // LDS offset of b (x follows at off + nf)
    const int* ptr;
    const int* col;
    const R* K;
    const R* M;
    const R* dK;
    const R* dM;
    const int* Rptr;  // restriction to the next level (rows: the next level's nodes)
    const int* Rcol;
    const R* Rval;
    const int* Pptr;  // prolongation from the next level (rows: this level's nodes)
    const int* Pcol;
    const R* Pval;
    R inv_theta;
    R c1[MG_TAIL_MAX_NU];
    R c2[MG_TAIL_MAX_NU];
};

template <typename R>
__device__ inline R tail_row(int i, const int* __restrict__ ptr, const int* __restrict__ col,
                             const R* __restrict__ K, const R* __restrict__ M, R s, const R* x)
{
    R acc = 0;
    const int b = ptr[i], e = ptr[i + 1];
#pragma unroll 4
    for (int k = b; k < e; ++k) acc += (K[k] + s * M[k]) * x[col[k]];
    return acc;
}

template <typename R>
__device__ inline R tail_csr_row(int i, const int* __restrict__ ptr, const int* __restrict__ col,
                                 const R* __restrict__ val, const R* x)
{
    R acc = 0;
    const int b = ptr[i], e = ptr[i + 1];
#pragma unroll 4
    for (int k = b; k < e; ++k) acc += val[k] * x[col[k]];
    return acc;
}

// Chebyshev-Jacobi smoothing of one level in LDS (as MGDevice::smooth); leaves
// r = b - A x in r.  (Keeping each thread's row in registers for a phase's three
// passes, and 256-thread workgroups, were both measured slower at 2+1D p=2.)
template <typename R>
__device__ void tail_smooth(const MGTailLevel<R>& L, R s, R* sm, R* r, R* d, R* dn, int nu, bool zero_start)
{
    const int nf = L.nf;
    const R* b = sm + L.off;
    R* x = sm + L.off + nf;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
        const R dg = L.dK[i] + s * L.dM[i];
        if (zero_start) {
            d[i] = b[i] * L.inv_theta / dg;
        } else {
            const R rv = b[i] - tail_row(i, L.ptr, L.col, L.K, L.M, s, x);
            r[i] = rv;
            d[i] = rv * L.inv_theta / dg;
        }
    }
    __syncthreads();
    R* dc = d;
    R* dx = dn;
    for (int k = 0; k < nu; ++k) {
        const bool last = k == nu - 1;
        const R c1 = L.c1[k], c2 = L.c2[k];
        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
            const R v = tail_row(i, L.ptr, L.col, L.K, L.M, s, dc);
            const R dv = dc[i];
            R rv;
            if (zero_start && k == 0) {
                x[i] = dv;
                rv = b[i] - v;
            } else {
                x[i] += dv;
                rv = r[i] - v;
            }
            r[i] = rv;
            if (!last) dx[i] = c1 * dv + c2 * rv / (L.dK[i] + s * L.dM[i]);
        }
        __syncthreads();
        R* tmp = dc;
        dc = dx;
        dx = tmp;
    }
}

template <typename R>
__global__ void __launch_bounds__(MG_TAIL_THREADS)
    k_mg_tail(const MGTailLevel<R>* __restrict__ lv, int nl, int nu, int ld, const R* __restrict__ sig,
              const int* __restrict__ sidx, const R* __restrict__ inv, int64_t blk_total,
              const int* __restrict__ comp_start, const int* __restrict__ node_comp,
              const int64_t* __restrict__ blk_off, const R* __restrict__ Bin, R* __restrict__ Xout, int tmp_off)
{
    extern __shared__ unsigned char tail_lds[];
    R* sm = reinterpret_cast<R*>(tail_lds);
    const int c = blockIdx.x;
    const R s = sig[c];
    const int nf0 = lv[0].nf;
    R* r = sm + tmp_off;
    R* d = r + nf0;
    R* dn = d + nf0;
    for (int i = threadIdx.x; i < nf0; i += blockDim.x) sm[lv[0].off + i] = Bin[static_cast<int64_t>(i) * ld + c];
    __syncthreads();
    for (int j = 0; j + 1 < nl; ++j) {
        const MGTailLevel<R>& F = lv[j];
        const MGTailLevel<R>& C = lv[j + 1];
        tail_smooth(F, s, sm, r, d, dn, nu, true);
        for (int i = threadIdx.x; i < C.nf; i += blockDim.x)
            sm[C.off + i] = tail_csr_row(i, F.Rptr, F.Rcol, F.Rval, r);
        __syncthreads();
    }
    {
        const MGTailLevel<R>& C = lv[nl - 1];
        const R* A0 = inv + static_cast<int64_t>(sidx[c]) * blk_total;
        const R* b = sm + C.off;
        for (int i = threadIdx.x; i < C.nf; i += blockDim.x) {
            const int comp = node_comp[i];
            const int st = comp_start[comp], bs = comp_start[comp + 1] - st;
            const R* A = A0 + blk_off[comp] + static_cast<int64_t>(i - st) * bs;
            R acc = 0;
            for (int j = 0; j < bs; ++j) acc += A[j] * b[st + j];
            sm[C.off + C.nf + i] = acc;
        }
        __syncthreads();
    }
    for (int j = nl - 2; j >= 0; --j) {
        const MGTailLevel<R>& F = lv[j];
        const MGTailLevel<R>& C = lv[j + 1];
        const R* xc = sm + C.off + C.nf;
        R* x = sm + F.off + F.nf;
        for (int i = threadIdx.x; i < F.nf; i += blockDim.x) x[i] += tail_csr_row(i, F.Pptr, F.Pcol, F.Pval, xc);
        __syncthreads();
        tail_smooth(F, s, sm, r, d, dn, nu, false);
    }
    for (int i = threadIdx.x; i < nf0; i += blockDim.x)
        Xout[static_cast<int64_t>(i) * ld + c] = sm[lv[0].off + nf0 + i];
}

constexpr int MG_TT = 1024;  // threads of the transpose blocks
constexpr int MG_LB = 4;     // staging loads in flight per thread

// gather: Z[s][i] = sum_tn r_X[tn nf + s] W[tn][i]  (i < T)  and  Z[s][T + c] = r_Y[c nf + s].
// A block of MG_TT threads owns th consecutive spatial rows: it reads the
// vector coalesced along s into an LDS tile (row stride th + 1), keeps W in LDS
// too when WL, and writes Z rows coalesced along the columns.
// With yo (Y only) the X columns are neither read nor written.
template <typename R, bool WL>
__global__ void __launch_bounds__(MG_TT) k_mg_gather(int nf, int T, int nyc, int ld, int64_t nY, int th, int yo,
                                                   const double* __restrict__ r, const double* __restrict__ W,
                                                   R* __restrict__ Z)
{
    extern __shared__ double lds_g[];
    const int ncol = T + nyc, ts = th + 1;
    double* tile = lds_g;            // [ncol][th + 1]
    double* Wl = lds_g + ncol * ts;  // [T][T] (WL)
    const int s0 = blockIdx.x * th;
    const int ns = min(th, nf - s0);
    // MG_LB elements per thread at a time: the loads are unconditional (indices
    // clamped into range) so that they are all issued before the first store;
    // rows past the last one store zero
    const int cb = yo ? T : 0;
    const int nst = (ncol - cb) * th;
    for (int e0 = threadIdx.x; e0 < nst; e0 += MG_LB * MG_TT) {
        double v[MG_LB];
#pragma unroll
        for (int u = 0; u < MG_LB; ++u) {
            const int e = min(e0 + u * MG_TT, nst - 1);
            const int c = cb + e / th, sl = min(e % th, ns - 1);
            v[u] = r[(c < T ? nY + static_cast<int64_t>(c) * nf : static_cast<int64_t>(c - T) * nf) + s0 + sl];
        }
#pragma unroll
        for (int u = 0; u < MG_LB; ++u) {
            const int e = e0 + u * MG_TT;
            if (e < nst) tile[(cb + e / th) * ts + e % th] = (e % th < ns) ? v[u] : 0.0;
        }
    }
    if (WL && !yo)
        for (int e = threadIdx.x; e < T * T; e += blockDim.x) Wl[e] = W[e];
    __syncthreads();
    const double* Wp = WL ? Wl : W;
    for (int e = threadIdx.x; e < (yo ? 0 : ns * T); e += blockDim.x) {
        const int sl = e / T, c = e % T;
        double v = 0.0;
#pragma unroll 8
        for (int tn = 0; tn < T; ++tn) v += tile[tn * ts + sl] * Wp[tn * T + c];
        Z[static_cast<int64_t>(s0 + sl) * ld + c] = static_cast<R>(v);
    }
    for (int e = threadIdx.x; e < ns * nyc; e += blockDim.x) {
        const int sl = e / nyc, c = e % nyc;
        Z[static_cast<int64_t>(s0 + sl) * ld + T + c] = static_cast<R>(tile[(T + c) * ts + sl]);
    }
}

// scatter back: z_X[tn nf + s] = sum_i Z[s][i] W[tn][i] ; z_Y[(k Nt + l) nf + s] = sum_l' Minv_k[l][l'] Z[s][T + k Nt + l'],
// and dot(z, r) partials (one per block).  Tiles as in k_mg_gather; the entries
// of r are loaded before the products that they multiply.
template <typename R, bool WL>
__global__ void __launch_bounds__(MG_TT) k_mg_scatter(int nf, int T, int Nt, int m, int ld, int64_t nY, int th, int yo,
                                                    const R* __restrict__ Z, const double* __restrict__ W,
                                                    const double* __restrict__ MtInvH, const double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ part)
{
    extern __shared__ double lds_s[];
    const int ncol = T + m * Nt, ts = th + 1;
    double* tile = lds_s;            // [ncol][th + 1]
    double* Wl = lds_s + ncol * ts;  // [T][T] (WL)
    const int s0 = blockIdx.x * th;
    const int ns = min(th, nf - s0);
    const int cb = yo ? T : 0, nce = ncol - cb;
    const int nst = ns * nce;
    for (int e0 = threadIdx.x; e0 < nst; e0 += MG_LB * MG_TT) {
        double v[MG_LB];
#pragma unroll
        for (int u = 0; u < MG_LB; ++u) {
            const int e = min(e0 + u * MG_TT, nst - 1);
            v[u] = static_cast<double>(Z[static_cast<int64_t>(s0 + e / nce) * ld + cb + e % nce]);
        }
#pragma unroll
        for (int u = 0; u < MG_LB; ++u) {
            const int e = e0 + u * MG_TT;
            if (e < nst) tile[(cb + e % nce) * ts + e / nce] = v[u];
        }
    }
    if (WL && !yo)
        for (int e = threadIdx.x; e < T * T; e += blockDim.x) Wl[e] = W[e];
    __syncthreads();
    const double* Wp = WL ? Wl : W;
    double dot = 0.0;
    for (int e = threadIdx.x; e < (yo ? 0 : T * th); e += blockDim.x) {
        const int c = e / th, sl = e % th;  // c plays tn
        if (sl >= ns) continue;
        const int64_t idx = nY + static_cast<int64_t>(c) * nf + s0 + sl;
        const double rv = r[idx];
        double v = 0.0;
#pragma unroll 8
        for (int i = 0; i < T; ++i) v += tile[i * ts + sl] * Wp[c * T + i];
        z[idx] = v;
        dot += v * rv;
    }
    for (int e = threadIdx.x; e < m * Nt * th; e += blockDim.x) {
        const int kl = e / th, sl = e % th;
        if (sl >= ns) continue;
        const int k = kl / Nt, l = kl % Nt;
        const int64_t idx = static_cast<int64_t>(kl) * nf + s0 + sl;
        const double rv = r[idx];
        const double* Mi = MtInvH + (static_cast<int64_t>(k) * Nt + l) * Nt;
        double v = 0.0;
        for (int q = 0; q < Nt; ++q) v += Mi[q] * tile[(T + k * Nt + q) * ts + sl];
        z[idx] = v;
        dot += v * rv;
    }
    __shared__ double sh[MG_TT / 64];
    for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
    if (threadIdx.x % 64 == 0) sh[threadIdx.x / 64] = dot;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) s += sh[w];
        part[blockIdx.x] = s;
    }
}

// Launch f with NP, the passes of 64 columns that cover nc columns (at most 4 per round).
template <typename F>
void with_passes(int nc, F f)
{
    if (nc <= 64) f(std::integral_constant<int, 1>{});
    else if (nc <= 128) f(std::integral_constant<int, 2>{});
    else if (nc <= 192) f(std::integral_constant<int, 3>{});
    else f(std::integral_constant<int, 4>{});
}

template <typename R>
R* mg_alloc(System* sys, size_t n)
{
    return dalloc<R>(sys, n);
}

template <typename R>
std::vector<R> conv(const std::vector<double>& v)
{
    return std::vector<R>(v.begin(), v.end());
}

template <typename T>
void up(T* dst, const std::vector<T>& src)
{
    if (!src.empty()) HIP_TRY(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
}

inline unsigned grid_for(int64_t n, int block) { return static_cast<unsigned>((n + block - 1) / block); }

}  // namespace

template <typename R>
void MGDevice<R>::build(System* sys, const MGSetupHost& H, int nyc_)
{
    if (ready) {
        refresh(H);
        return;
    }
    const int L = static_cast<int>(H.levels.size());
    // tuning knobs: Chebyshev degree and smoothing interval ratio
    if (const char* e = std::getenv("APF_MG_DEGREE")) nu = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("APF_MG_RATIO")) cheb_ratio = std::max(1.5, std::atof(e));
    T = H.T;
    nyc = nyc_;
    ncol = T + nyc;
    ld = ncol;
    levels.resize(L);
    for (int l = 0; l < L; ++l) {
        const MGLevelHost& h = H.levels[l];
        Level& d = levels[l];
        d.nf = h.nf;
        d.lmax = h.lmax;
        d.nnz = static_cast<int64_t>(h.K.size());
        d.Pnnz = static_cast<int64_t>(h.P.val.size());
        d.Rnnz = static_cast<int64_t>(h.R.val.size());
        d.ptr = mg_alloc<int>(sys, h.sptr.size());
        up(d.ptr, h.sptr);
        d.col = mg_alloc<int>(sys, h.scol.size());
        up(d.col, h.scol);
        d.K = mg_alloc<R>(sys, h.K.size());
        up(d.K, conv<R>(h.K));
        d.M = mg_alloc<R>(sys, h.M.size());
        up(d.M, conv<R>(h.M));
        d.dK = mg_alloc<R>(sys, h.nf);
        up(d.dK, conv<R>(h.dK));
        d.dM = mg_alloc<R>(sys, h.nf);
        up(d.dM, conv<R>(h.dM));
        const size_t blk = static_cast<size_t>(h.nf) * ld;
        d.X = mg_alloc<R>(sys, blk);
        d.B = mg_alloc<R>(sys, blk);
        d.Rr = mg_alloc<R>(sys, blk);
        d.D = mg_alloc<R>(sys, blk);
        d.Tm = mg_alloc<R>(sys, blk);
        if (l + 1 < L) {
            d.Pn = h.P.nrows;
            d.Pptr = mg_alloc<int>(sys, h.P.ptr.size());
            up(d.Pptr, h.P.ptr);
            d.Pcol = mg_alloc<int>(sys, h.P.col.size());
            up(d.Pcol, h.P.col);
            d.Pval = mg_alloc<R>(sys, h.P.val.size());
            up(d.Pval, conv<R>(h.P.val));
            d.Rn = h.R.nrows;
            d.Rptr = mg_alloc<int>(sys, h.R.ptr.size());
            up(d.Rptr, h.R.ptr);
            d.Rcol = mg_alloc<int>(sys, h.R.col.size());
            up(d.Rcol, h.R.col);
            d.Rval = mg_alloc<R>(sys, h.R.val.size());
            up(d.Rval, conv<R>(h.R.val));
        }
    }
    // per-column shifts: X columns carry sqrt(theta_i), Y columns 0
    std::vector<R> sig(ncol, 0), zero(ncol, 0);
    std::vector<int> sidx(ncol, 0), zidx(ncol, 0);
    for (int i = 0; i < T; ++i) {
        sig[i] = static_cast<R>(H.shift[i]);
        sidx[i] = 1 + i;
    }
    d_sig = mg_alloc<R>(sys, ncol);
    up(d_sig, sig);
    d_zero = mg_alloc<R>(sys, ncol);
    up(d_zero, zero);
    d_sidx = mg_alloc<int>(sys, ncol);
    up(d_sidx, sidx);
    d_zidx = mg_alloc<int>(sys, ncol);
    up(d_zidx, zidx);
    nc_coarse = H.levels.back().nf;
    blk_total = H.blk_total;
    cinv_size = static_cast<int64_t>(H.coarse_inv.size());
    W_size = static_cast<int64_t>(H.W.size());
    MtInvH_size = static_cast<int64_t>(H.MtInvH.size());
    d_cinv = mg_alloc<R>(sys, H.coarse_inv.size());
    up(d_cinv, conv<R>(H.coarse_inv));
    d_comp_start = mg_alloc<int>(sys, H.comp_start.size());
    up(d_comp_start, H.comp_start);
    d_node_comp = mg_alloc<int>(sys, std::max<size_t>(H.node_comp.size(), 1));
    if (!H.node_comp.empty()) up(d_node_comp, H.node_comp);
    d_blk_off = mg_alloc<int64_t>(sys, std::max<size_t>(H.blk_off.size(), 1));
    if (!H.blk_off.empty()) up(d_blk_off, H.blk_off);
    d_W = mg_alloc<double>(sys, H.W.size());
    up(d_W, H.W);
    d_MtInvH = mg_alloc<double>(sys, H.MtInvH.size());
    up(d_MtInvH, H.MtInvH);
    setup_tail(sys);
    ready = true;
}

// Choose the first level from which the rest of the V-cycle fits one workgroup's
// LDS (b and x of every level plus three work vectors of the first, at most
// MG_TAIL_LDS_CAP) and upload the level descriptors.  APF_MG_TAIL=0 keeps the
// per-level launches.
template <typename R>
void MGDevice<R>::setup_tail(System* sys)
{
    const int L = static_cast<int>(levels.size());
    tail_from = -1;
    if (const char* e = std::getenv("APF_MG_TAIL"))
        if (std::atoi(e) == 0) return;
    if (L < 2) return;
    int from = -1;
    size_t bytes = 0;
    for (int l = 0; l + 1 < L; ++l) {
        size_t words = 3 * static_cast<size_t>(levels[l].nf);
        for (int j = l; j < L; ++j) words += 2 * static_cast<size_t>(levels[j].nf);
        if (words * sizeof(R) <= MG_TAIL_LDS_CAP) {
            from = l;
            bytes = words * sizeof(R);
            break;
        }
    }
    if (from < 0) return;
    std::vector<MGTailLevel<R>> h(L - from);
    int off = 0;
    for (int l = from; l < L; ++l) {
        const Level& d = levels[l];
        MGTailLevel<R>& t = h[l - from];
        t = MGTailLevel<R>{};
        t.nf = d.nf;
        t.off = off;
        off += 2 * d.nf;
        t.ptr = d.ptr;
        t.col = d.col;
        t.K = d.K;
        t.M = d.M;
        t.dK = d.dK;
        t.dM = d.dM;
        t.Rptr = d.Rptr;
        t.Rcol = d.Rcol;
        t.Rval = d.Rval;
        t.Pptr = d.Pptr;
        t.Pcol = d.Pcol;
        t.Pval = d.Pval;
        // the coefficients of smooth(), in the same type and order
        const R lmax = static_cast<R>(d.lmax);
        const R lmin = lmax / static_cast<R>(cheb_ratio);
        const R theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2;
        const R sigma1 = theta / delta;
        t.inv_theta = static_cast<R>(1) / theta;
        R rho = 1 / sigma1;
        // the step coefficients do not depend on the degree (only the last step differs)
        for (int k = 0; k < MG_TAIL_MAX_NU; ++k) {
            const R rho_n = 1 / (2 * sigma1 - rho);
            t.c1[k] = rho_n * rho;
            t.c2[k] = 2 * rho_n / delta;
            rho = rho_n;
        }
    }
    if (!d_tail) d_tail = mg_alloc<MGTailLevel<R>>(sys, h.size());
    up(static_cast<MGTailLevel<R>*>(d_tail), h);
    tail_from = from;
    tail_lds = bytes;
}

template <typename R>
void MGDevice<R>::refresh(const MGSetupHost& H)
{
    for (size_t l = 0; l < levels.size(); ++l) {
        const MGLevelHost& h = H.levels[l];
        Level& d = levels[l];
        d.lmax = h.lmax;
        up(d.K, conv<R>(h.K));
        up(d.M, conv<R>(h.M));
        up(d.dK, conv<R>(h.dK));
        up(d.dM, conv<R>(h.dM));
    }
    std::vector<R> sig(ncol, 0);
    for (int i = 0; i < T; ++i) sig[i] = static_cast<R>(H.shift[i]);
    up(d_sig, sig);
    up(d_cinv, conv<R>(H.coarse_inv));
    up(d_W, H.W);
    up(d_MtInvH, H.MtInvH);
    // the smoothing bounds may have changed: rewrite the tail descriptors (d_tail is
    // already allocated, so no System is needed)
    if (d_tail) setup_tail(nullptr);
}

template <typename S, typename R>
__global__ void k_convert(int64_t n, const S* __restrict__ a, R* __restrict__ b)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        b[i] = static_cast<R>(a[i]);
}

template <typename R>
template <typename S>
void MGDevice<R>::refresh_from(const MGDevice<S>& src, hipStream_t st)
{
    if (!ready || !src.ready || src.levels.size() != levels.size() || src.ncol != ncol || src.blk_total != blk_total ||
        src.cinv_size != cinv_size)
        throw std::runtime_error("multigrid hierarchies of different structure");
    auto cv = [&](int64_t n, const S* a, R* b) {
        if (n > 0)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<S, R>), dim3(static_cast<unsigned>(std::min<int64_t>((n + 255) / 256, 1024))),
                               dim3(256), 0, st, n, a, b);
    };
    for (size_t l = 0; l < levels.size(); ++l) {
        Level& d = levels[l];
        const auto& e = src.levels[l];
        if (e.nf != d.nf || e.nnz != d.nnz || e.Pnnz != d.Pnnz || e.Rnnz != d.Rnnz)
            throw std::runtime_error("multigrid hierarchies of different structure");
        d.lmax = e.lmax;
        cv(d.nnz, e.K, d.K);
        cv(d.nnz, e.M, d.M);
        cv(d.nf, e.dK, d.dK);
        cv(d.nf, e.dM, d.dM);
        cv(d.Pnnz, e.Pval, d.Pval);
        cv(d.Rnnz, e.Rval, d.Rval);
    }
    cv(ncol, src.d_sig, d_sig);
    cv(cinv_size, src.d_cinv, d_cinv);
    if (W_size) HIP_TRY(hipMemcpyAsync(d_W, src.d_W, W_size * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (MtInvH_size)
        HIP_TRY(hipMemcpyAsync(d_MtInvH, src.d_MtInvH, MtInvH_size * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipGetLastError());
    // the smoothing coefficients follow the new bounds (d_tail is already allocated)
    if (d_tail) {
        HIP_TRY(hipStreamSynchronize(st));
        setup_tail(nullptr);
    }
}

template <typename R>
void MGDevice<R>::smooth(hipStream_t st, int l, int c0, int nc, const R* sig, bool zero_start, int deg)
{
    Level& d = levels[l];
    R* X = d.X + c0;
    R* B = d.B + c0;
    R* Rr = d.Rr + c0;
    const R lmax = static_cast<R>(d.lmax);
    const R lmin = lmax / static_cast<R>(cheb_ratio);
    const R theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2;
    const R sigma1 = theta / delta;
    const int64_t nel = static_cast<int64_t>(d.nf) * nc;
    const int blk = 256;
    const unsigned rows_grid = grid_for(d.nf, 4);
    if (zero_start) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cheb_start_zero<R>), dim3(grid_for(nel, blk)), dim3(blk), 0, st, d.nf, nc,
                           ld, d.dK, d.dM, sig, static_cast<R>(1) / theta, B, d.D + c0);
    } else {
        with_passes(nc, [&](auto np) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_resid_start<R, decltype(np)::value>), dim3(rows_grid), dim3(256), 0,
                               st, d.nf, d.ptr, d.col, d.K, d.M, sig, nc, ld, X, B, Rr, d.D + c0, d.dK, d.dM,
                               static_cast<R>(1) / theta);
        });
    }
    // the direction alternates between D and Tm
    R* dcur = d.D + c0;
    R* dnext = d.Tm + c0;
    R rho = 1 / sigma1;
    for (int k = 0; k < deg; ++k) {
        const R rho_n = 1 / (2 * sigma1 - rho);
        const R c1 = rho_n * rho, c2 = 2 * rho_n / delta;
        with_passes(nc, [&](auto np) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_cheb_fused<R, decltype(np)::value>), dim3(rows_grid), dim3(256), 0,
                               st, d.nf, d.ptr, d.col, d.K, d.M, sig, nc, ld, dcur, X, Rr, dnext, d.dK, d.dM, c1, c2,
                               (k == deg - 1) ? 1 : 0, (zero_start && k == 0) ? B : nullptr);
        });
        std::swap(dcur, dnext);
        rho = rho_n;
    }
}

// V-cycle on level l: B_l -> X_l (columns [c0, c0 + nc); sig and sidx already offset)
template <typename R>
void MGDevice<R>::vcycle(hipStream_t st, int l, int c0, int nc, const R* sig, const int* sidx, int deg)
{
    Level& d = levels[l];
    const int L = static_cast<int>(levels.size());
    if (l == tail_from && deg <= MG_TAIL_MAX_NU) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_tail<R>), dim3(nc), dim3(MG_TAIL_THREADS), tail_lds, st,
                           static_cast<const MGTailLevel<R>*>(d_tail), L - tail_from, deg, ld, sig, sidx, d_cinv,
                           blk_total, d_comp_start, d_node_comp, d_blk_off, d.B + c0, d.X + c0,
                           static_cast<int>(tail_lds / sizeof(R)) - 3 * d.nf);
        return;
    }
    if (l == L - 1) {
        const int64_t nel = static_cast<int64_t>(d.nf) * nc;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_coarse_solve<R>), dim3(grid_for(nel, 256)), dim3(256), 0, st, d.nf, nc, ld,
                           d_cinv, blk_total, d_comp_start, d_node_comp, d_blk_off, sidx, d.B + c0, d.X + c0);
        return;
    }
    Level& c = levels[l + 1];
    smooth(st, l, c0, nc, sig, true, deg);  // leaves r = B - A x in Rr
    with_passes(nc, [&](auto np) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_csr_mm<R, decltype(np)::value>), dim3(grid_for(d.Rn, 4)), dim3(256), 0, st,
                           d.Rn, d.Rptr, d.Rcol, d.Rval, nc, ld, d.Rr + c0, ld, c.B + c0, 0);
    });
    vcycle(st, l + 1, c0, nc, sig, sidx, deg);
    with_passes(nc, [&](auto np) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_csr_mm<R, decltype(np)::value>), dim3(grid_for(d.Pn, 4)), dim3(256), 0, st,
                           d.Pn, d.Pptr, d.Pcol, d.Pval, nc, ld, c.X + c0, ld, d.X + c0, 1);
    });
    smooth(st, l, c0, nc, sig, false, deg);
}

// z = P^-1 r for the saddle-point vector r = [r_Y; r_X]; writes dot(z, r) partials.
template <typename R>
void MGDevice<R>::apply(hipStream_t st, const Space& S, const double* r, double* z, double* part, int nparts,
                        bool y_only, int degree)
{
    const int deg = degree > 0 ? degree : nu;
    Level& f = levels[0];
    const int nf = f.nf;
    const int th = tile_rows();
    const unsigned tiles = static_cast<unsigned>((nf + th - 1) / th);
    // the transpose tile, and W when it fits in 40 KiB
    const bool wl = static_cast<size_t>(T) * T * sizeof(double) <= 40 * 1024;
    const size_t lds = (static_cast<size_t>(ncol) * (th + 1) + (wl ? static_cast<size_t>(T) * T : 0)) * sizeof(double);
    const int yo = y_only ? 1 : 0;
    if (wl)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_gather<R, true>), dim3(tiles), dim3(MG_TT), lds, st, nf, T, nyc, ld, S.nY,
                           th, yo, r, d_W, f.B);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_gather<R, false>), dim3(tiles), dim3(MG_TT), lds, st, nf, T, nyc, ld,
                           S.nY, th, yo, r, d_W, f.B);
    // stage 1: all columns, Y columns with shift 0, X columns with shift s_i (only the Y
    // columns when the caller reads only z's Y part)
    if (y_only) vcycle(st, 0, T, nyc, d_sig + T, d_sidx + T, deg);
    else vcycle(st, 0, 0, ncol, d_sig, d_sidx, deg);
    // stage 2 (X columns): apply K, then the shifted V-cycle again.  The second V-cycle
    // writes only the X columns [0, T) of the fine level's X, so afterwards X holds the
    // X columns of stage 2 and the Y columns of stage 1: the scatter reads it in place.
    if (!y_only) {
        with_passes(T, [&](auto np) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_apply<R, decltype(np)::value>), dim3(grid_for(nf, 4)), dim3(256), 0,
                               st, nf, f.ptr, f.col, f.K, f.M, d_zero, T, ld, f.X, f.X, f.B, 0);
        });
        vcycle(st, 0, 0, T, d_sig, d_sidx, deg);
    }
    const R* Zs = f.X;
    if (wl)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_scatter<R, true>), dim3(tiles), dim3(MG_TT), lds, st, nf, T, S.Nt,
                           S.mesh.m(), ld, S.nY, th, yo, Zs, d_W, d_MtInvH, r, z, part);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_scatter<R, false>), dim3(tiles), dim3(MG_TT), lds, st, nf, T, S.Nt,
                           S.mesh.m(), ld, S.nY, th, yo, Zs, d_W, d_MtInvH, r, z, part);
}

template <typename R>
void MGDevice<R>::bench_step(hipStream_t st, int reps, double* ms, double* bytes)
{
    Level& d = levels[0];
    const int nc = ncol;
    const unsigned rows_grid = grid_for(d.nf, 4);
    auto launch = [&]() {
        with_passes(nc, [&](auto np) {
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_mg_cheb_fused<R, decltype(np)::value>), dim3(rows_grid), dim3(256), 0,
                               st, d.nf, d.ptr, d.col, d.K, d.M, d_sig, nc, ld, d.D, d.X, d.Rr, d.Tm, d.dK, d.dM,
                               static_cast<R>(0.5), static_cast<R>(0.5), 0, static_cast<const R*>(nullptr));
        });
    };
    launch();  // warm
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r) launch();
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = static_cast<double>(t) / reps;
    // d, x, r read; x, r, dn written (nf x nc each); the level's pattern, K, M, diagonals
    *bytes = 6.0 * sizeof(R) * static_cast<double>(d.nf) * nc +
             static_cast<double>(d.nnz) * (sizeof(int) + 2 * sizeof(R)) +
             static_cast<double>(d.nf) * (sizeof(int) + 2 * sizeof(R));
}

template <typename R>
int MGDevice<R>::tile_rows() const
{
    // rows per transpose tile: the LDS tile ncol x (th + 1) doubles stays within
    // 40 KiB, so with W beside it two blocks share a CU
    int th = 40 * 1024 / (ncol * static_cast<int>(sizeof(double))) - 1;
    th = std::max(1, std::min(64, th));
    return th;
}

template <typename R>
int MGDevice<R>::n_partials(int nf) const
{
    const int th = tile_rows();
    return (nf + th - 1) / th;
}

template struct MGDevice<double>;
template struct MGDevice<float>;
template void MGDevice<float>::refresh_from<double>(const MGDevice<double>&, hipStream_t);

}  // namespace apf
