// Device kernels shared by the MINRES drivers (the tensor-mesh system in apf_gpu.hip and
// the caller's-CSR solver in apf_csr.hip): CSR SpMV, deterministic block reductions and
// the Paige-Saunders MINRES vector/scalar steps.  Internal linkage: each translation
// unit that includes this header gets its own copies.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "apf_internal.hpp"

#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function"

namespace {

using apf::MinresScalars;

// ============================================================== SpMV
// CSR SpMV y = K x.  A group of LANES lanes owns RP rows (rows gi, gi + G, ...
// of the block's G*RP rows, G = groups per block); the first LANES*U entries of
// all RP rows are loaded together, U (value, column) pairs per lane and row,
// before any x is gathered, so a wave has RP row fetches in flight instead of
// one.  Longer rows continue in LANES*U chunks.  Values and column indices are
// streamed once, x is gathered through the cache.
template <int LANES, int U, int RP>
__global__ void __launch_bounds__(256) k_spmv(int64_t n, const int64_t* __restrict__ rowptr,
                                              const int* __restrict__ col, const double* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y)
{
    constexpr int G = 256 / LANES;
    const int lane = threadIdx.x % LANES, gi = threadIdx.x / LANES;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * (G * RP) + gi;
    int64_t b[RP];
    int len[RP];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        const int64_t row = base + k * G;
        const int64_t rc = row < n ? row : n - 1;  // rows past the end: length 0
        b[k] = rowptr[rc];
        len[k] = row < n ? static_cast<int>(rowptr[rc + 1] - b[k]) : 0;
    }
    double sum[RP];
    double a[RP][U];
    int c[RP][U];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        sum[k] = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = lane + u * LANES;
            a[k][u] = (j < len[k]) ? val[b[k] + j] : 0.0;
            c[k][u] = (j < len[k]) ? col[b[k] + j] : 0;
        }
    }
#pragma unroll
    for (int k = 0; k < RP; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u) sum[k] += a[k][u] * x[c[k][u]];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        for (int j0 = LANES * U; j0 < len[k]; j0 += LANES * U) {
            double a2[U];
            int c2[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = j0 + lane + u * LANES;
                a2[u] = (j < len[k]) ? val[b[k] + j] : 0.0;
                c2[u] = (j < len[k]) ? col[b[k] + j] : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) sum[k] += a2[u] * x[c2[u]];
        }
#pragma unroll
        for (int off = LANES / 2; off > 0; off >>= 1) sum[k] += __shfl_down(sum[k], off, LANES);
        const int64_t row = base + k * G;
        if (row < n && lane == 0) y[row] = sum[k];
    }
}

// ============================================================== reductions
__device__ inline double block_sum(double v)
{
    __shared__ double sh[64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < static_cast<int>(blockDim.x / 64); ++i) r += sh[i];
    __syncthreads();
    return r;
}

// deterministic final reduction of per-block partials (single block)
__device__ inline double reduce_partials(const double* part, int nparts)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
    return block_sum(s);
}

__global__ void k_residual_parts(int64_t n, const double* __restrict__ b, const double* __restrict__ Kx,
                                 double* __restrict__ part)
{
    double s_r = 0.0, s_b = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double r = b[i] - Kx[i];
        s_r += r * r;
        s_b += b[i] * b[i];
    }
    const double a = block_sum(s_r);
    const double c = block_sum(s_b);
    if (threadIdx.x == 0) {
        part[blockIdx.x] = a;
        part[gridDim.x + blockIdx.x] = c;
    }
}

__global__ void k_residual_final(const double* part, int nparts, double* out)
{
    const double a = reduce_partials(part, nparts);
    const double c = reduce_partials(part + nparts, nparts);
    if (threadIdx.x == 0) { out[0] = a; out[1] = c; }
}

// Paige-Saunders MINRES (the algorithm behind Eigen::MINRES used at
// src/linear_algebra/eigen_backend/preconditioners/parabolic_graph_norm_minres.hpp:240-243),
// with all scalar recurrences kept on the device so the iteration needs no
// host synchronisation.
__global__ void k_minres_vec1(int64_t n, const MinresScalars* __restrict__ st, const double* __restrict__ y,
                              double* __restrict__ t, const double* __restrict__ r1, double* __restrict__ v,
                              double* __restrict__ part)
{
    const double s = 1.0 / st->beta;
    const double c = (st->itn >= 1) ? st->beta / st->oldb : 0.0;
    double dot = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double vi = s * y[i];
        const double ti = s * t[i] - c * r1[i];
        v[i] = vi;
        t[i] = ti;
        dot += vi * ti;
    }
    const double bs = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

__global__ void k_minres_upd(int64_t n, const MinresScalars* __restrict__ st, double* __restrict__ t,
                             const double* __restrict__ r2)
{
    const double c = st->alfa / st->beta;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        t[i] -= c * r2[i];
}

__global__ void k_minres_alfa(MinresScalars* st, const double* part, int nparts)
{
    const double a = reduce_partials(part, nparts);
    if (threadIdx.x == 0) st->alfa = a;
}

__global__ void k_minres_scal2(MinresScalars* st, const double* part, int nparts)
{
    const double b2 = reduce_partials(part, nparts);
    if (threadIdx.x != 0) return;
    MinresScalars s = *st;
    s.oldb = s.beta;
    s.beta = sqrt(fmax(b2, 0.0));
    s.oldeps = s.epsln;
    s.delta = s.cs * s.dbar + s.sn * s.alfa;
    const double gbar = s.sn * s.dbar - s.cs * s.alfa;
    s.epsln = s.sn * s.beta;
    s.dbar = -s.cs * s.beta;
    double gamma = hypot(gbar, s.beta);
    if (gamma < 1e-300) gamma = 1e-300;
    s.gamma = gamma;
    s.cs = gbar / gamma;
    s.sn = s.beta / gamma;
    s.phi = s.cs * s.phibar;
    s.phibar = s.sn * s.phibar;
    s.itn += 1;
    *st = s;
}

__global__ void k_minres_vec3(int64_t n, const MinresScalars* __restrict__ st, const double* __restrict__ v,
                              const double* __restrict__ w2, const double* __restrict__ w,
                              double* __restrict__ wnew /* aliases old w1 */, double* __restrict__ x)
{
    const double oe = st->oldeps, de = st->delta, ig = 1.0 / st->gamma, ph = st->phi;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const double wn = (v[i] - oe * w2[i] - de * w[i]) * ig;
        wnew[i] = wn;
        x[i] += ph * wn;
    }
}

__global__ void k_minres_init(MinresScalars* st, const double* part, int nparts)
{
    const double b2 = reduce_partials(part, nparts);
    if (threadIdx.x != 0) return;
    MinresScalars s{};
    s.beta = sqrt(fmax(b2, 0.0));
    s.beta1 = s.beta;
    s.oldb = 0.0;
    s.dbar = 0.0;
    s.epsln = 0.0;
    s.phibar = s.beta;
    s.cs = -1.0;
    s.sn = 0.0;
    s.itn = 0;
    *st = s;
}

__global__ void k_axpby(int64_t n, double a, const double* x, double b, double* y)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

__global__ void k_dot_parts(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                            double* __restrict__ part)
{
    double s = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        s += a[i] * b[i];
    const double t = block_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_dot_final(const double* part, int nparts, double* out)
{
    const double s = reduce_partials(part, nparts);
    if (threadIdx.x == 0) out[0] = s;
}

}  // namespace

#pragma clang diagnostic pop
