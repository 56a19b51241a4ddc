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
#include <cstdlib>
#include <numeric>
#include <string>
#include <type_traits>
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
    int refine_steps;
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

// max of non-negative statistics in which a NaN counts as +infinity (fmax would drop it),
// so that a non-finite patch solution or residual fails the acceptance test
__device__ inline double max_nonfinite_wins(double a, double b)
{
    return (a != a || b != b) ? INFINITY : fmax(a, b);
}

__device__ inline void atomic_max_pos(double* addr, double v)
{
    // v >= 0: the IEEE order of non-negative doubles matches their bit order
    atomicMax(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

template <int DIM, int PP>
struct LocalSizes {
    static constexpr int NS = DIM == 1 ? PP + 1 : (PP + 1) * (PP + 2) / 2;  // spatial Lagrange
    static constexpr int NT = PP + 1;                                        // time modes
    static constexpr int NF = DIM == 1 ? PP + 2 : (PP + 1) * (PP + 3);      // local flux
    static constexpr int NQ = NS;                                            // local scalar
    static constexpr int NV = DIM + 1;
};

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename FT>
struct Vec4;
template <>
struct Vec4<float> {
    using type = v4f;
};
template <>
struct Vec4<double> {
    using type = v4d;
};

// D += A B for one 16x16x4 tile: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; element r of the result is D[drow(l / 16, r)][l % 16],
// with drow = 4 (l / 16) + r for FP32 and l / 16 + 4 r for FP64 (measured lane
// layouts of v_mfma_f32_16x16x4f32 and v_mfma_f64_16x16x4f64 on gfx950).
__device__ inline v4f mfma_16x16x4(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
__device__ inline v4d mfma_16x16x4(double a, double b, v4d c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
template <typename FT>
__device__ constexpr int drow(int lk, int r);
template <>
__device__ constexpr int drow<float>(int lk, int r) { return 4 * lk + r; }
template <>
__device__ constexpr int drow<double>(int lk, int r) { return lk + 4 * r; }

// Patch slots in processing order q (largest patch first): order[q] is the
// patch, n[q] its order, np[q] the order padded to a multiple of 16, off2[q]
// the offset of its np x np matrix slots and off1[q] of its np-row vectors.
struct PatchSlots {
    const int* n;
    const int* np;
    const int64_t* off2;
    const int64_t* off1;
};

// In-place Gauss-Jordan inversion of the n x n matrix A (row stride ld) with
// partial row pivoting (ties to the lowest row).  Lanes own columns
// (j = lane + 64 t), waves own rows (i = wave + nw s).  Wave 0 picks the pivot,
// swaps it into place and scales it; then every wave eliminates its rows: two
// workgroup barriers per step, no integer division in the update.
template <typename FT>
__device__ void gj_invert(FT* A, int n, int ld, int* perm)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int nch = (n + 255) / 256;
    for (int k = 0; k < n; ++k) {
        if (wave == 0) {
            double best = -1.0, bval = 0.0;
            int bi = n;
            for (int r = k + lane; r < n; r += 64) {
                const double v = static_cast<double>(A[static_cast<int64_t>(r) * ld + k]);
                if (!(fabs(v) <= best)) {  // a NaN entry wins, so the pivot row is always a real row
                    best = fabs(v);
                    bval = v;
                    bi = r;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off, 64);
                const double ov = __shfl_xor(bval, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (!(ob <= best) || (ob == best && oi < bi)) {
                    best = ob;
                    bval = ov;
                    bi = oi;
                }
            }
            const int pr = min(max(bi, k), n - 1);
            const FT inv = static_cast<FT>(1) / static_cast<FT>(bval);
            for (int j = lane; j < n; j += 64) {
                const FT v = A[static_cast<int64_t>(pr) * ld + j];
                if (pr != k) A[static_cast<int64_t>(pr) * ld + j] = A[static_cast<int64_t>(k) * ld + j];
                A[static_cast<int64_t>(k) * ld + j] = ((j == k) ? static_cast<FT>(1) : v) * inv;
            }
            if (lane == 0) perm[k] = pr;
        }
        __syncthreads();
        // the column block holding k goes last: its update overwrites A[i][k]
        const int kc = k >> 8;
        for (int cc = 1; cc <= nch; ++cc) {
            const int jb = ((kc + cc) % nch) * 256;
            FT prow[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int j = jb + lane + 64 * t;
                prow[t] = (j < n) ? A[static_cast<int64_t>(k) * ld + j] : static_cast<FT>(0);
            }
            for (int i = wave; i < n; i += nw) {
                if (i == k) continue;
                FT* row = A + static_cast<int64_t>(i) * ld;
                const FT f = row[k];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int j = jb + lane + 64 * t;
                    if (j < n) row[j] = ((j == k) ? static_cast<FT>(0) : row[j]) - f * prow[t];
                }
            }
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges; each thread owns whole rows
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        FT* row = A + static_cast<int64_t>(i) * ld;
        for (int k = n - 1; k >= 0; --k) {
            const int pr = perm[k];
            if (pr != k) {
                const FT t = row[k];
                row[k] = row[pr];
                row[pr] = t;
            }
        }
    }
    __syncthreads();
}

struct Metric2 {
    double det, G[4], H[4];
};

__device__ inline Metric2 metric_of(int dim, int c, const double* verts, const int* cells, const double* Mcell)
{
    const CellMetric g = cell_metric(dim, c, verts, cells, Mcell);
    Metric2 r;
    r.det = g.det;
    for (int i = 0; i < 4; ++i) {
        r.G[i] = g.G[i];
        r.H[i] = g.H[i];
    }
    return r;
}

// theta = lambda_tilde + u and u local coefficients of space-time cell (k, c):
// [Ns][Nt], Lagrange in time.
__device__ inline void gather_cell(const EstParams& P, int k, int c, const int* cell_free, const double* lt,
                                   const double* x, int m_, int l, double& th, double& uu)
{
    const int s = cell_free[c * P.Ns + m_];
    if (s < 0) {
        th = 0.0;
        uu = 0.0;
        return;
    }
    const double lam = lt[(static_cast<int64_t>(k) * P.Nt + l) * P.nf + s];
    const double u = x[P.nY + (static_cast<int64_t>(k) * P.p + l) * P.nf + s];
    th = lam + u;
    uu = u;
}

constexpr int PATCH_THREADS = 256;

// FP64 patch matrix of one time mode (flux mass, divergence coupling,
// mean-zero row; dense_local_error_blocks_2d.hpp:719-768, 812-879) of patch a
// into K (np x np, zero padded, exactly symmetric).  Only the entries of the
// patch's pattern are written: the slot was zeroed when it was allocated and
// no other entry is ever written.  With Alds, the same contributions also go to
// a copy in LDS (FP32 or FP64, row stride S).  Ends with a barrier.
//
// A flux-flux entry has two contributions only when both of its unknowns lie on
// one edge of the patch (the two cells of that edge); two cells share at most
// one edge, so every other entry comes from one cell and is stored directly.
// Only the same-edge entries are zeroed first and summed with atomics.
template <typename AT = float>
__device__ void build_patch_matrix(const EstParams& P, const EstTables& T, const PatchTopo& topo, int a, int n, int np,
                                   const double* verts, const int* cells, const double* Mcell, double* K,
                                   AT* Alds = nullptr, int S = 0)
{
    const int Nf = P.Nf, Nq = P.Nq, dim = P.dim;
    const int nF = topo.nflux[a];
    const bool interior = topo.boundary[a] == 0;
    const int cb = topo.cell_ptr[a], ce = topo.cell_ptr[a + 1];
    // cell metrics once per cell (patches of up to 64 cells; larger ones recompute)
    constexpr int GMAX = 64;
    __shared__ double gmH[GMAX][5];
    const int nc = ce - cb;
    for (int ci = threadIdx.x; ci < min(nc, GMAX); ci += blockDim.x) {
        const Metric2 g = metric_of(dim, topo.cells[cb + ci], verts, cells, Mcell);
        for (int u = 0; u < 4; ++u) gmH[ci][u] = g.H[u];
        gmH[ci][4] = g.det;
    }
    // local flux unknowns on the cell's edges: face i / Nfm for i < Nfc Nfm
    const int nedge = P.Nfc * P.Nfm;
    auto same_edge = [&](int i, int j) { return i < nedge && j < nedge && i / P.Nfm == j / P.Nfm; };
    // zero the same-edge entries the atomic sums below accumulate into
    for (int e = threadIdx.x; e < nc * nedge * P.Nfm; e += blockDim.x) {
        const int ci = cb + e / (nedge * P.Nfm), w = e % (nedge * P.Nfm);
        const int i = w / P.Nfm, j = (i / P.Nfm) * P.Nfm + w % P.Nfm;
        if (j < i) continue;
        const int fi = topo.fmap[static_cast<int64_t>(ci) * Nf + i], fj = topo.fmap[static_cast<int64_t>(ci) * Nf + j];
        if (fi < 0 || fj < 0) continue;
        K[static_cast<int64_t>(fi) * np + fj] = 0.0;
        K[static_cast<int64_t>(fj) * np + fi] = 0.0;
    }
    if (Alds)
        for (int e = threadIdx.x; e < np * S; e += blockDim.x) Alds[e] = static_cast<AT>(0);
    __syncthreads();
    // every cell at once: a same-edge entry gets at most two contributions (the
    // cells of the edge), so its FP64 atomic sum is exact-order independent
    const int per_cell = Nf * Nf + Nq * Nf + Nq;
    for (int e = threadIdx.x; e < nc * per_cell; e += blockDim.x) {
        const int ci = cb + e / per_cell, w = e % per_cell;
        const int* fm = topo.fmap + static_cast<int64_t>(ci) * Nf;
        const int8_t* fs = topo.fsign + static_cast<int64_t>(ci) * Nf;
        if (w < Nf * Nf) {
            // flux mass: int M^-1 s_i . s_j = sum_ab H_ab RTab_ab[i][j], i <= j mirrored
            const int i = w / Nf, j = w % Nf;
            if (j < i) continue;
            const int fi = fm[i], fj = fm[j];
            if (fi < 0 || fj < 0) continue;
            double H[4];
            if (ci - cb < GMAX) {
                for (int u = 0; u < 4; ++u) H[u] = gmH[ci - cb][u];
            } else {
                const Metric2 g = metric_of(dim, topo.cells[ci], verts, cells, Mcell);
                for (int u = 0; u < 4; ++u) H[u] = g.H[u];
            }
            // (ab / dim) * 2 + ab % dim = ab for dim = 2, and dim = 1 has ab = 0 only
            double v = 0.0;
#pragma unroll
            for (int ab = 0; ab < 4; ++ab)
                if (ab < dim * dim) v += H[ab] * T.RTab[(static_cast<int64_t>(ab) * Nf + i) * Nf + j];
            v *= static_cast<double>(fs[i] * fs[j]);
            if (same_edge(i, j)) {
                atomicAdd(&K[static_cast<int64_t>(fi) * np + fj], v);
                if (i != j) atomicAdd(&K[static_cast<int64_t>(fj) * np + fi], v);
                if (Alds) {
                    atomicAdd(&Alds[fi * S + fj], static_cast<AT>(v));
                    if (i != j) atomicAdd(&Alds[fj * S + fi], static_cast<AT>(v));
                }
            } else {
                K[static_cast<int64_t>(fi) * np + fj] = v;
                K[static_cast<int64_t>(fj) * np + fi] = v;
                if (Alds) {
                    Alds[fi * S + fj] = static_cast<AT>(v);
                    Alds[fj * S + fi] = static_cast<AT>(v);
                }
            }
        } else if (w < Nf * Nf + Nq * Nf) {
            // divergence coupling int q_i div s_j (orthonormal scalars)
            const int w2 = w - Nf * Nf;
            const int qq = w2 / Nf, j = w2 % Nf;
            const int fj = fm[j];
            if (fj < 0) continue;
            const int r0 = nF + (ci - cb) * Nq;
            const double v = fs[j] * T.Bhat[qq * Nf + j];
            K[static_cast<int64_t>(r0 + qq) * np + fj] = v;
            K[static_cast<int64_t>(fj) * np + r0 + qq] = v;
            if (Alds) {
                Alds[(r0 + qq) * S + fj] = static_cast<AT>(v);
                Alds[fj * S + r0 + qq] = static_cast<AT>(v);
            }
        } else if (interior) {
            // mean-zero row
            const int qq = w - Nf * Nf - Nq * Nf;
            const int r0 = nF + (ci - cb) * Nq;
            const double det = (ci - cb < GMAX) ? gmH[ci - cb][4] : metric_of(dim, topo.cells[ci], verts, cells, Mcell).det;
            const double v = det * T.qbar[qq];
            K[static_cast<int64_t>(r0 + qq) * np + (n - 1)] = v;
            K[static_cast<int64_t>(n - 1) * np + r0 + qq] = v;
            if (Alds) {
                Alds[(r0 + qq) * S + (n - 1)] = static_cast<AT>(v);
                Alds[(n - 1) * S + r0 + qq] = static_cast<AT>(v);
            }
        }
    }
    __syncthreads();
}

// Phase 1 (large patches), one workgroup per spatial vertex patch of one size
// bucket: the patch matrix into its slot of Kd and its inverse in FT into the
// slot of Ki (Gauss-Jordan in LDS when in_lds, else in place in Ki).
template <typename FT>
__global__ void __launch_bounds__(PATCH_THREADS)
    k_patch_factor(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, int q0, int in_lds, const double* verts,
                   const int* cells, const double* Mcell, double* Kd, FT* Ki)
{
    extern __shared__ double lds[];
    const int q = q0 + blockIdx.x;
    const int a = topo.order[q];
    const int n = ps.n[q], np = ps.np[q];
    double* K = Kd + ps.off2[q];
    FT* Kg = Ki + ps.off2[q];
    const int64_t nn = static_cast<int64_t>(np) * np;
    build_patch_matrix(P, T, topo, a, n, np, verts, cells, Mcell, K);
    FT* A = in_lds ? reinterpret_cast<FT*>(lds) : Kg;
    int* perm = reinterpret_cast<int*>(in_lds ? reinterpret_cast<char*>(lds) + nn * sizeof(FT)
                                              : reinterpret_cast<char*>(lds));
    for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) A[e] = static_cast<FT>(K[e]);
    __syncthreads();
    gj_invert<FT>(A, n, np, perm);
    if (in_lds)
        for (int64_t e = threadIdx.x; e < nn; e += blockDim.x) Kg[e] = A[e];
}

__device__ inline float lane_value(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
__device__ inline double lane_value(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b), src));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), src));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

__device__ inline double first_lane(double v)
{
    const long long b = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b)));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(b >> 32)));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Shared state of the register-resident Gauss-Jordan below.
template <typename FT, int NW, int CPL>
struct GJShared {
    FT prow[64 * CPL];
    double cabs[NW], cval[NW];
    int cidx[NW];
    int pivrow[64 * CPL], rinv[64 * CPL];
};

// In-place Gauss-Jordan inversion of an n x n matrix held in registers, row
// i = wave + NW il in A[il][.], column j = lane + 64 t in A[.][t]
// (n <= min(NW RPW, 64 CPL)).  Partial pivoting among the rows not yet used as
// pivots (ties to the lowest row), no row interchange: step k scales pivot row
// r_k, broadcasts it through LDS and eliminates column k from every other row,
// the multiplier A[i][k] coming from lane k % 64 by readlane.  Afterwards A
// holds T with inv[k][r_k'] = T[r_k][k'] (sh.pivrow[k] = r_k, sh.rinv[r_k] = k).
// Two barriers per step; the pivot candidates of column k+1 are found while
// column k is eliminated.  Wave id, pivot row and multipliers are
// wave-uniform and kept in scalar registers.
template <typename FT, int NW, int RPW, int CPL>
__device__ void gj_registers(FT (&A)[RPW][CPL], int n, GJShared<FT, NW, CPL>& sh)
{
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    unsigned used = 0;  // bit il: row wave + NW il is a pivot row already, or beyond n
#pragma unroll
    for (int il = 0; il < RPW; ++il)
        if (wave + NW * il >= n) used |= 1u << il;
    auto candidates = [&](int kc) {
        const int tk = kc >> 6, lk = kc & 63;
        double best = -1.0, bval = 0.0;
        int bi = 0x7fffffff;
#pragma unroll
        for (int il = 0; il < RPW; ++il) {
            if (used & (1u << il)) continue;
            FT v = 0;
#pragma unroll
            for (int t = 0; t < CPL; ++t)
                if (t == tk) v = A[il][t];
            const double av = fabs(static_cast<double>(v));
            if (!(av <= best)) {  // a NaN entry wins, so the pivot row is always a real row
                best = av;
                bval = static_cast<double>(v);
                bi = wave + NW * il;
            }
        }
        best = lane_value(best, lk);
        bval = lane_value(bval, lk);
        bi = __builtin_amdgcn_readlane(bi, lk);
        if (lane == 0) {
            sh.cabs[wave] = best;
            sh.cval[wave] = bval;
            sh.cidx[wave] = bi;
        }
    };
    for (int i = threadIdx.x; i < 64 * CPL; i += blockDim.x) sh.rinv[i] = i;
    candidates(0);
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        double best = -1.0, pv = 0.0;
        int r = 0x7fffffff;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const double b = sh.cabs[w];
            const int ri = sh.cidx[w];
            if (!(b <= best) || (b == best && ri < r)) {
                best = b;
                pv = sh.cval[w];
                r = ri;
            }
        }
        // every index below stays inside the matrix whatever the data
        r = __builtin_amdgcn_readfirstlane(min(max(r, 0), n - 1));
        pv = first_lane(pv);
        if (wave == r % NW) {
            const int ilr = r / NW;
            const FT inv = static_cast<FT>(1) / static_cast<FT>(pv);
#pragma unroll
            for (int il = 0; il < RPW; ++il)
                if (il == ilr) {
#pragma unroll
                    for (int t = 0; t < CPL; ++t) {
                        const int j = lane + 64 * t;
                        const FT v = ((j == k) ? static_cast<FT>(1) : A[il][t]) * inv;
                        A[il][t] = v;
                        sh.prow[j] = v;
                    }
                    used |= 1u << il;
                }
        }
        if (threadIdx.x == 0) {
            sh.pivrow[k] = r;
            sh.rinv[r] = k;
        }
        __syncthreads();
        FT pr[CPL];
#pragma unroll
        for (int t = 0; t < CPL; ++t) pr[t] = sh.prow[lane + 64 * t];
        const int tk = k >> 6, lk = k & 63;
#pragma unroll
        for (int il = 0; il < RPW; ++il) {
            const int i = wave + NW * il;
            if (i == r || i >= n) continue;
            FT ck = 0;
#pragma unroll
            for (int t = 0; t < CPL; ++t)
                if (t == tk) ck = A[il][t];
            const FT f = lane_value(ck, lk);
#pragma unroll
            for (int t = 0; t < CPL; ++t) {
                FT v = A[il][t];
                if (t == tk && lane == lk) v = static_cast<FT>(0);
                A[il][t] = v - f * pr[t];
            }
        }
        if (k + 1 < n) candidates(k + 1);
        __syncthreads();
    }
}

// Operand staging of wg_gemm: k-chunks of both factors (M, N <= 128).
struct GemmShared {
    static constexpr int KC = 16;
    double a[128][KC + 1];
    double b[KC][128 + 4];
};

// C(i, j) = sum_k A(i, k) B(k, j) for i < M, j < N (M, N <= 128), handed to
// store(i, j, value), with blockDim.x * TPT >= 32 * 32.  k-chunks of both
// factors are staged in LDS with all their loads in flight, then every thread
// updates its TPT 4 x 4 register tiles from LDS.
template <int TPT, typename FA, typename FB, typename FS>
__device__ void wg_gemm(int M, int N, int K, FA A, FB B, FS store, GemmShared& g)
{
    constexpr int KC = GemmShared::KC;
    const int mt = (M + 3) / 4, nt = (N + 3) / 4;
    double acc[TPT][4][4] = {};
    for (int k0 = 0; k0 < K; k0 += KC) {
        const int kc = min(KC, K - k0);
        __syncthreads();
        for (int e = threadIdx.x; e < M * KC; e += blockDim.x) {
            const int i = e / KC, kk = e % KC;
            g.a[i][kk] = (kk < kc) ? A(i, k0 + kk) : 0.0;
        }
        for (int e = threadIdx.x; e < N * KC; e += blockDim.x) {
            const int j = e / KC, kk = e % KC;
            g.b[kk][j] = (kk < kc) ? B(k0 + kk, j) : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < TPT; ++u) {
            const int tile = threadIdx.x + u * blockDim.x;
            if (tile < mt * nt) {
            const int i0 = (tile % mt) * 4, j0 = (tile / mt) * 4;
            for (int kk = 0; kk < kc; ++kk) {
                double av[4], bv[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) av[r] = (i0 + r < M) ? g.a[i0 + r][kk] : 0.0;
#pragma unroll
                for (int c = 0; c < 4; ++c) bv[c] = (j0 + c < N) ? g.b[kk][j0 + c] : 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r)
#pragma unroll
                    for (int c = 0; c < 4; ++c) acc[u][r][c] += av[r] * bv[c];
            }
            }
        }
    }
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int tile = threadIdx.x + u * blockDim.x;
        const int i0 = (tile % mt) * 4, j0 = (tile / mt) * 4;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (tile < mt * nt && i0 + r < M && j0 + c < N) store(i0 + r, j0 + c, acc[u][r][c]);
    }
}

// Phase 1 (patches of padded order <= min(NW RPW, 64 CPL)): the whole patch
// matrix in registers, inverted by one register-resident Gauss-Jordan in FT,
// written to the FT inverse slot with the pivot permutation folded in.
template <typename FT, int NW, int RPW, int CPL>
__global__ void __launch_bounds__(NW * 64)
    k_patch_factor_full(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, int q0, const double* verts,
                        const int* cells, const double* Mcell, double* Kd, FT* Ki)
{
    __shared__ GJShared<FT, NW, CPL> sh;
    const int q = q0 + blockIdx.x;
    const int a = topo.order[q];
    const int n = ps.n[q], np = ps.np[q];
    double* K = Kd + ps.off2[q];
    FT* Kg = Ki + ps.off2[q];
    build_patch_matrix(P, T, topo, a, n, np, verts, cells, Mcell, K);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    FT R[RPW][CPL];
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            R[il][t] = (i < n && j < n) ? static_cast<FT>(K[static_cast<int64_t>(i) * np + j]) : static_cast<FT>(0);
        }
    gj_registers<FT, NW, RPW, CPL>(R, n, sh);
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            if (i < n && j < n)
                Kg[static_cast<int64_t>(sh.rinv[i]) * np + sh.pivrow[j]] = R[il][t];
            else if (i < np && j < np)
                Kg[static_cast<int64_t>(i) * np + j] = static_cast<FT>(0);
        }
}

// Largest value over the wave (all 64 lanes) by DPP row shifts and row
// broadcasts, no LDS: row_shr 1, 2, 3, 4, 8 within rows of 16, then
// row_bcast 15 and 31 across rows; lane 63 holds the result.
__device__ inline unsigned wave_umax(unsigned v)
{
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x113, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xe, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xc, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
    v = max(v, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
    return static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Pivot-search key bits of |v| (a non-negative value orders like its bit
// pattern): the FP32 word, or the high word of the FP64 value.
__device__ inline unsigned pivot_bits(float v) { return __float_as_uint(fabsf(v)); }
__device__ inline unsigned pivot_bits(double v) { return static_cast<unsigned>(__double2hiint(fabs(v))); }
// 1 / v: the FP32 hardware reciprocal (the FP64 refinement of the solve absorbs
// its last-bit error), the correctly rounded FP64 quotient
__device__ inline float pivot_recip(float v) { return __builtin_amdgcn_rcpf(v); }
__device__ inline double pivot_recip(double v) { return 1.0 / v; }
// 2^-floor(e / 2) for the binary exponent e of v > 0 (1 for v = 0)
__device__ inline float equil_scale(float v)
{
    int e = 0;
    (void)frexpf(v, &e);
    return (v > 0.0f) ? ldexpf(1.0f, -(e >> 1)) : 1.0f;
}
__device__ inline double equil_scale(double v)
{
    int e = 0;
    (void)frexp(v, &e);
    return (v > 0.0) ? ldexp(1.0, -(e >> 1)) : 1.0;
}

// Phase 1, blocked (padded order np <= 64 RPL, FT = FP32 or FP64): Gauss-Jordan
// in 16-column panels on the matrix in LDS.  Wave 0 eliminates a panel in
// registers (lane l owns rows l + 64 s, all 16 panel columns; partial pivoting
// among unused rows, no row interchange), so a panel step needs no workgroup
// barrier and every row's multiplier is in its own lane.  The panel's
// accumulated transform E (identity outside the pivot columns R, which the
// eliminated panel P holds) is then applied to the other columns by all waves
// as one rank-16 MFMA update  A[:, T] += (P - I_R) A[R, T]  with A[R, T]
// copied first.  Same arithmetic as the unblocked elimination, reordered.
// FP64 at np = 128 takes 157 KB of LDS: one workgroup per CU.
template <typename FT, int RPL>
__global__ void __launch_bounds__(256)
    k_patch_factor_blocked(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, int q0, const double* verts,
                           const int* cells, const double* Mcell, double* Kd, FT* Ki)
{
    using V4 = typename Vec4<FT>::type;
    constexpr int NPM = 64 * RPL;  // largest padded order
    // LDS row strides: S = np + 4 (= 4 mod 16) spreads the MFMA tile accesses
    // (rows 4 (l / 16) + r, columns l % 16) and the row-tile operand reads over
    // distinct banks; RS = 16 mod 64 does the same for the pivot-row operand.
    // The scaling vector lives in the pivot-row buffer while the matrix is
    // scaled, then in the stride padding of each row (column np, never read by
    // the elimination).
    __shared__ __align__(16) FT A[NPM * (NPM + 4)];
    __shared__ __align__(16) FT Rb[16 * (NPM + 16)];
    __shared__ int pivrow[NPM], rinv[NPM];
    FT* dsc = Rb;
    const int q = q0 + blockIdx.x;
    const int a = topo.order[q];
    const int n = ps.n[q], np = ps.np[q];
    const int S = np + 4, RS = np + ((16 - np) & 63);
    double* K = Kd + ps.off2[q];
    // the FP64 matrix into its slot and its FT copy into LDS in one pass
    build_patch_matrix<FT>(P, T, topo, a, n, np, verts, cells, Mcell, K, A, S);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int nrt = np >> 4;
    // symmetric equilibration D K D with d_i = 2^-floor(e_i / 2), e_i the binary
    // exponent of max_j |K_ij| (so d_i^2 max_j |K_ij| lies in [1/2, 2)): the
    // flux, divergence and mean rows differ in scale by powers of h, and an FP32
    // Gauss-Jordan inverse of the scaled matrix is several times more accurate.
    // Power-of-two factors make the scaling and the final unscaling exact.
    // Each wave reduces four rows (wave + 4u of a group of 16) at a time.
    for (int i0 = wave; i0 < np; i0 += 16) {  // np is a multiple of 16
        FT mx[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            mx[u] = static_cast<FT>(0);
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int j = lane + 64 * t;
                if (j < np) mx[u] = fmax(mx[u], fabs(A[(i0 + 4 * u) * S + j]));
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
#pragma unroll
            for (int u = 0; u < 4; ++u) mx[u] = fmax(mx[u], __shfl_xor(mx[u], off, 64));
        if (lane == 0) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                dsc[i0 + 4 * u] = equil_scale(mx[u]);
                rinv[i0 + 4 * u] = i0 + 4 * u;
            }
        }
    }
    __syncthreads();
    auto dkeep = [&](int i) -> FT& { return A[i * S + np]; };
    for (int i = wave; i < np; i += 4) {
        const FT di = dsc[i];
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int j = lane + 64 * t;
            if (j < np) A[i * S + j] *= di * dsc[j];
        }
        if (lane == 0) dkeep(i) = di;
    }
    __syncthreads();
    unsigned used = 0;  // wave 0: bit s = row lane + 64 s used as pivot (or >= n)
#pragma unroll
    for (int s = 0; s < RPL; ++s)
        if (lane + 64 * s >= n) used |= 1u << s;
    for (int k0 = 0; k0 < n; k0 += 16) {
        const int kc = min(16, n - k0);
        if (wave == 0) {
            // wave 0 eliminates the panel in registers: lane l owns rows l + 64 s
            FT Pn[RPL][16];
#pragma unroll
            for (int s = 0; s < RPL; ++s)
#pragma unroll
                for (int c = 0; c < 16; ++c) Pn[s][c] = A[(lane + 64 * s) * S + k0 + c];
            // one pivot step, branch-free so that the unrolled panel is one basic
            // block (step kk + 1's pivot search can overlap step kk's updates); kk
            // is a compile-time constant in the unrolled loops below
            int prs[16];  // pivot row of each step (wave-uniform)
            auto pivot_step = [&](const int kk) __attribute__((always_inline)) {
                // pivot: largest |A[i][k]| over unused rows, as one 32-bit key
                // per candidate (the key bits of |v| with their 7 low bits
                // replaced by 127 - i: a pivot within 1 - 2^-16 (FP32) or
                // 1 - 2^-13 (FP64) of the largest is as good, ties go to the
                // lowest row, a NaN wins) and one DPP max
                unsigned key = 0u;
#pragma unroll
                for (int s = 0; s < RPL; ++s) {
                    const unsigned kv = (pivot_bits(Pn[s][kk]) & ~0x7Fu) | static_cast<unsigned>(127 - (lane + 64 * s));
                    key = (used & (1u << s)) ? key : max(key, kv);
                }
                key = wave_umax(key);
                const int r = min(max(127 - static_cast<int>(key & 0x7Fu), 0), n - 1);
                const int owner = r & 63, sr = r >> 6;
                // the pivot row, scaled by the reciprocal pivot
                FT prow[16];
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    const FT src = (RPL == 1 || sr == 0) ? Pn[0][c] : Pn[RPL - 1][c];
                    prow[c] = lane_value(src, owner);
                }
                const FT inv = pivot_recip(prow[kk]);
#pragma unroll
                for (int c = 0; c < 16; ++c) prow[c] = (c == kk) ? inv : prow[c] * inv;
                // every row eliminates (padding rows are zero and stay zero),
                // then the owner lane puts the scaled pivot row in place
                const bool mine = lane == owner;
#pragma unroll
                for (int s = 0; s < RPL; ++s) {
                    const FT f = Pn[s][kk];
                    const bool w = mine && (RPL == 1 || s == sr);
#pragma unroll
                    for (int c = 0; c < 16; ++c) {
                        const FT e = ((c == kk) ? static_cast<FT>(0) : Pn[s][c]) - f * prow[c];
                        Pn[s][c] = w ? prow[c] : e;
                    }
                }
                used |= (mine ? 1u : 0u) << sr;
                prs[kk] = r;
            };
            // full panels without per-step guards (straight-line code: the
            // guarded loop costs ~40 register moves per step); the last panel
            // of an order that is not a multiple of 16 is guarded
            if (kc == 16) {
#pragma clang loop unroll(full)
                for (int kk = 0; kk < 16; ++kk) pivot_step(kk);
            } else {
#pragma clang loop unroll(full)
                for (int kk = 0; kk < 16; ++kk)
                    if (kk < kc) pivot_step(kk);
            }
            // lane kk records step kk's pivot row
            int myr = 0;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
                if (kk < kc) myr = (lane == kk) ? prs[kk] : myr;
            if (lane < kc) {
                pivrow[k0 + lane] = myr;
                rinv[myr] = k0 + lane;
            }
#pragma unroll
            for (int s = 0; s < RPL; ++s)
#pragma unroll
                for (int c = 0; c < 16; ++c) A[(lane + 64 * s) * S + k0 + c] = Pn[s][c];
        }
        __syncthreads();
        // pivot rows outside the panel, before the update (4 panel rows per wave)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = wave + 4 * u;
            const int pr = (c < kc) ? pivrow[k0 + c] : 0;
#pragma unroll
            for (int t = 0; t < RPL; ++t) {
                const int j = lane + 64 * t;
                if (j < np) Rb[c * RS + j] = (c < kc && (j < k0 || j >= k0 + 16)) ? A[pr * S + j] : static_cast<FT>(0);
            }
        }
        __syncthreads();
        // A[:, T] += (P - I_R) Rb on MFMA; the panel's own column tile is left
        // alone.  Each wave keeps the (P - I_R) operand of its row tile.
        const int pt = k0 >> 4;
        for (int rt = wave; rt < nrt; rt += 4) {
            FT av[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int c = 4 * t + lk, i = rt * 16 + li;
                av[t] = (c < kc) ? A[i * S + k0 + c] - ((i == pivrow[k0 + c]) ? static_cast<FT>(1) : static_cast<FT>(0))
                                 : static_cast<FT>(0);
            }
#pragma unroll
            for (int ct = 0; ct < NPM / 16; ++ct) {
                if (ct >= nrt || ct == pt) continue;
                V4 acc;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[r] = A[(rt * 16 + drow<FT>(lk, r)) * S + ct * 16 + li];
#pragma unroll
                for (int t = 0; t < 4; ++t) acc = mfma_16x16x4(av[t], Rb[(4 * t + lk) * RS + ct * 16 + li], acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) A[(rt * 16 + drow<FT>(lk, r)) * S + ct * 16 + li] = acc[r];
            }
        }
        __syncthreads();
    }
    // K^-1 = D (D K D)^-1 D, with the pivot permutation folded in; rows by
    // wave, columns by lane
    FT* Kg = Ki + ps.off2[q];
    for (int i = wave; i < np; i += 4) {
        const int r = (i < n) ? rinv[i] : 0;
        const FT dr = (i < n) ? dkeep(r) : static_cast<FT>(0);
#pragma unroll
        for (int t = 0; t < RPL; ++t) {
            const int j = lane + 64 * t;
            if (j >= np) continue;
            if (i < n && j < n) {
                const int c = pivrow[j];
                Kg[static_cast<int64_t>(r) * np + c] = A[i * S + j] * (dr * dkeep(c));
            } else {
                Kg[static_cast<int64_t>(i) * np + j] = static_cast<FT>(0);
            }
        }
    }
}

// Phase 1 (patches whose flux block and Schur block both have order <=
// min(NW RPW, 64 CPL)).  With the patch unknowns ordered [flux; scalar; mean
// multiplier] the matrix is K = [A C^T; C D], A the SPD flux mass.  Block
// elimination gives
//   K^-1 = [A^-1 + G H^T, -H; -H^T, S^-1],  G = A^-1 C^T,  S = D - C G,  H^T = S^-1 G^T,
// so the dense inversions shrink from order n to nF (A) and n - nF (S), both
// register-resident Gauss-Jordan; the products between them are workgroup
// GEMMs on the FP64 inverse slot: A^-1 in the top-left block, G in the
// top-right, S and then S^-1 in the bottom-right, H^T in the bottom-left.
// Everything is FP64: the top-left block is A^-1 minus a matrix of similar
// size (the inverse on the divergence-free fluxes), and an FP32 A^-1 loses
// too much there.  With packed FP32 off, FP64 and FP32 VALU rates are equal.
// For FT = float the result is also rounded into Ki32 for the FP32 products.
template <typename FT, int NW, int RPW, int CPL>
__global__ void __launch_bounds__(NW * 64)
    k_patch_factor_schur(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, int q0, const double* verts,
                         const int* cells, const double* Mcell, double* Kd, double* Ki, float* Ki32)
{
    __shared__ GJShared<double, NW, CPL> sh;
    __shared__ GemmShared gs;
    const int q = q0 + blockIdx.x;
    const int a = topo.order[q];
    const int n = ps.n[q], np = ps.np[q];
    const int nF = topo.nflux[a], m = n - nF;
    double* K = Kd + ps.off2[q];
    double* Kg = Ki + ps.off2[q];
    build_patch_matrix(P, T, topo, a, n, np, verts, cells, Mcell, K);
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double R[RPW][CPL];
    // 1. A^-1 into the top-left block
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            R[il][t] = (i < nF && j < nF) ? K[static_cast<int64_t>(i) * np + j] : 0.0;
        }
    gj_registers<double, NW, RPW, CPL>(R, nF, sh);
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            if (i < nF && j < nF) Kg[static_cast<int64_t>(sh.rinv[i]) * np + sh.pivrow[j]] = R[il][t];
        }
    __syncthreads();
    // 2. G = A^-1 C^T  (nF x m, top-right)
    wg_gemm<32 * 32 / (NW * 64)>(
        nF, m, nF, [&](int i, int k) { return Kg[static_cast<int64_t>(i) * np + k]; },
        [&](int k, int s) { return K[static_cast<int64_t>(nF + s) * np + k]; },
        [&](int i, int s, double v) { Kg[static_cast<int64_t>(i) * np + nF + s] = v; }, gs);
    __syncthreads();
    // 3. S = D - C G  (m x m, bottom-right)
    wg_gemm<32 * 32 / (NW * 64)>(
        m, m, nF, [&](int s, int k) { return K[static_cast<int64_t>(nF + s) * np + k]; },
        [&](int k, int s2) { return Kg[static_cast<int64_t>(k) * np + nF + s2]; },
        [&](int s, int s2, double v) {
            const int64_t e = static_cast<int64_t>(nF + s) * np + nF + s2;
            Kg[e] = K[e] - v;
        }, gs);
    __syncthreads();
    // 4. S^-1 in place
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            R[il][t] = (i < m && j < m) ? Kg[static_cast<int64_t>(nF + i) * np + nF + j] : 0.0;
        }
    gj_registers<double, NW, RPW, CPL>(R, m, sh);
#pragma unroll
    for (int il = 0; il < RPW; ++il)
#pragma unroll
        for (int t = 0; t < CPL; ++t) {
            const int i = wave + NW * il, j = lane + 64 * t;
            if (i < m && j < m) Kg[static_cast<int64_t>(nF + sh.rinv[i]) * np + nF + sh.pivrow[j]] = R[il][t];
        }
    __syncthreads();
    // 5. H^T = S^-1 G^T  (m x nF, bottom-left)
    wg_gemm<32 * 32 / (NW * 64)>(
        m, nF, m, [&](int s, int k) { return Kg[static_cast<int64_t>(nF + s) * np + nF + k]; },
        [&](int k, int j) { return Kg[static_cast<int64_t>(j) * np + nF + k]; },
        [&](int s, int j, double v) { Kg[static_cast<int64_t>(nF + s) * np + j] = v; }, gs);
    __syncthreads();
    // 6. top-left += G H^T
    wg_gemm<32 * 32 / (NW * 64)>(
        nF, nF, m, [&](int i, int k) { return Kg[static_cast<int64_t>(i) * np + nF + k]; },
        [&](int k, int j) { return Kg[static_cast<int64_t>(nF + k) * np + j]; },
        [&](int i, int j, double v) { Kg[static_cast<int64_t>(i) * np + j] += v; }, gs);
    __syncthreads();
    // 7. off-diagonal blocks -H, -H^T; zero padding rows and columns
    for (int e = threadIdx.x; e < m * nF; e += blockDim.x) {
        const int s = e / nF, j = e % nF;
        const double v = -Kg[static_cast<int64_t>(nF + s) * np + j];
        Kg[static_cast<int64_t>(nF + s) * np + j] = v;
        Kg[static_cast<int64_t>(j) * np + nF + s] = v;
    }
    for (int e = threadIdx.x; e < np * np; e += blockDim.x) {
        const int i = e / np, j = e % np;
        if (i >= n || j >= n) Kg[e] = 0.0;
    }
    if constexpr (std::is_same<FT, float>::value) {
        __syncthreads();
        float* K32 = Ki32 + ps.off2[q];
        for (int e = threadIdx.x; e < np * np; e += blockDim.x) K32[e] = static_cast<float>(Kg[e]);
    }
}

// Phase 1 for large patches (order > 128 with a diagonal block beyond the
// Schur kernel, padded order <= 2048): Gauss-Jordan inversion blocked in
// 16-column panels across launches, batched over every such patch.  Each
// member g owns a workspace region at ws + off[g] (doubles):
//   T    [np][np]  FT  the equilibrated matrix D K D, inverted in place;
//   d    [np]          d_i = 1 / sqrt(max_j |K_ij|) (1 on padding rows);
//   piv  [np]      int pivot row of column k (implicit pivoting: no swaps);
//   mark [np]      int panel that took row i as pivot, -1 while unused.
// Padding rows/columns n..np-1 carry an identity block, so every column has a
// pivot.  Panel c0 (k_big_panel) eliminates columns c0..c0+15 with partial
// pivoting over the unused rows; T's panel columns then hold P = M [e_r0..e_r15]
// for the panel's Gauss-Jordan transform M = I + (P - E_R) E_R^T, and the
// other columns take  T[:,j] <- (rows R zeroed) T[:,j] + P T[R,j]
// (k_big_update, MFMA).  k_big_finish undoes the implicit pivoting and the
// scaling: K^-1[k][r_k'] = d_k T[r_k][k'] d_r_k'.
struct BigPatches {
    const int* q;        // slot of member g
    const int64_t* off;  // workspace offset (doubles) of member g
    double* ws;
};

struct BigView {
    int np;
    void* T;
    double* d;
    int* piv;
    int* mark;
};

__device__ inline BigView big_view(const PatchSlots& ps, const BigPatches& bp, int g)
{
    BigView v;
    v.np = ps.np[bp.q[g]];
    double* base = bp.ws + bp.off[g];
    v.T = base;
    v.d = base + static_cast<int64_t>(v.np) * v.np;
    v.piv = reinterpret_cast<int*>(v.d + v.np);
    v.mark = v.piv + v.np;
    return v;
}

template <typename FT>
__global__ void __launch_bounds__(PATCH_THREADS)
    k_big_prepare(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, BigPatches bp, const double* verts,
                  const int* cells, const double* Mcell, double* Kd)
{
    const int q = bp.q[blockIdx.x];
    const int a = topo.order[q], n = ps.n[q], np = ps.np[q];
    double* K = Kd + ps.off2[q];
    build_patch_matrix(P, T, topo, a, n, np, verts, cells, Mcell, K);
    const BigView v = big_view(ps, bp, blockIdx.x);
    FT* A = static_cast<FT*>(v.T);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int i = wave; i < np; i += nw) {
        double mx = 0.0;
        if (i < n)
            for (int j = lane; j < n; j += 64) mx = fmax(mx, fabs(K[static_cast<int64_t>(i) * np + j]));
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        if (lane == 0) {
            v.d[i] = (i < n && mx > 0.0) ? 1.0 / sqrt(mx) : 1.0;
            v.mark[i] = -1;
            v.piv[i] = i;
        }
    }
    __syncthreads();
    for (int i = wave; i < np; i += nw) {
        const double di = v.d[i];
        const double* Kr = K + static_cast<int64_t>(i) * np;
        FT* Ar = A + static_cast<int64_t>(i) * np;
        for (int j = lane; j < np; j += 64)
            Ar[j] = (i < n && j < n) ? static_cast<FT>(di * Kr[j] * v.d[j]) : static_cast<FT>(i == j ? 1 : 0);
    }
}

// one workgroup of 512 threads per member; thread t owns rows t + 512 s (s < RPT)
// of the 16-column panel in registers
template <typename FT, int RPT>
__global__ void __launch_bounds__(512) k_big_panel(PatchSlots ps, BigPatches bp, int c0)
{
    const BigView v = big_view(ps, bp, blockIdx.x);
    const int np = v.np;
    if (c0 >= np) return;
    FT* A = static_cast<FT*>(v.T);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    __shared__ double s_best[8];
    __shared__ int s_row[8];
    __shared__ FT s_prow[16];
    FT x[RPT][16];
    bool used[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        const int i = tid + 512 * s;
        used[s] = true;
#pragma unroll
        for (int j = 0; j < 16; ++j) x[s][j] = static_cast<FT>(0);
        if (i < np) {
            used[s] = v.mark[i] >= 0;
            const FT* Ar = A + static_cast<int64_t>(i) * np + c0;
#pragma unroll
            for (int j = 0; j < 16; ++j) x[s][j] = Ar[j];
        }
    }
    const int panel = c0 >> 4;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        // largest |x| over the unused rows, ties to the lower row; a NaN wins so
        // that some row is always taken
        double best = -1.0;
        int br = 0x7fffffff;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            const double av = fabs(static_cast<double>(x[s][k]));
            if (!used[s] && !(av <= best)) {
                best = av;
                br = tid + 512 * s;
            }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off, 64);
            const int orow = __shfl_xor(br, off, 64);
            if (!(ob <= best) || (ob == best && orow < br)) {
                best = ob;
                br = orow;
            }
        }
        if (lane == 0) {
            s_best[wave] = best;
            s_row[wave] = br;
        }
        __syncthreads();
        best = s_best[0];
        br = s_row[0];
        for (int w = 1; w < 8; ++w)
            if (!(s_best[w] <= best) || (s_best[w] == best && s_row[w] < br)) {
                best = s_best[w];
                br = s_row[w];
            }
        const int r = min(max(br, 0), np - 1);
#pragma unroll
        for (int s = 0; s < RPT; ++s)
            if (tid + 512 * s == r) {
#pragma unroll
                for (int j = 0; j < 16; ++j) s_prow[j] = x[s][j];
                used[s] = true;
                v.mark[r] = panel;
                v.piv[c0 + k] = r;
            }
        __syncthreads();
        const FT inv = static_cast<FT>(1) / s_prow[k];
        FT pr[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) pr[j] = s_prow[j] * inv;
#pragma unroll
        for (int s = 0; s < RPT; ++s) {
            if (tid + 512 * s == r) {
#pragma unroll
                for (int j = 0; j < 16; ++j) x[s][j] = (j == k) ? inv : pr[j];
            } else {
                const FT f = x[s][k];
#pragma unroll
                for (int j = 0; j < 16; ++j) x[s][j] = (j == k) ? -f * inv : x[s][j] - f * pr[j];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int s = 0; s < RPT; ++s) {
        const int i = tid + 512 * s;
        if (i < np) {
            FT* Ar = A + static_cast<int64_t>(i) * np + c0;
#pragma unroll
            for (int j = 0; j < 16; ++j) Ar[j] = x[s][j];
        }
    }
}

// workgroup (member, 64-column strip): T[:, strip] <- (rows R zeroed) + P T[R, strip],
// the panel's own columns excepted; 4 waves over the 16-row tiles
template <typename FT>
__global__ void __launch_bounds__(256) k_big_update(PatchSlots ps, BigPatches bp, int c0, int nstrips)
{
    using V4 = typename Vec4<FT>::type;
    const int g = blockIdx.x / nstrips, strip = blockIdx.x - g * nstrips;
    const BigView v = big_view(ps, bp, g);
    const int np = v.np, j0 = strip * 64;
    if (c0 >= np || j0 >= np) return;
    FT* A = static_cast<FT*>(v.T);
    __shared__ FT s_ar[16][64];
    __shared__ int s_r[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    if (tid < 16) s_r[tid] = v.piv[c0 + tid];
    __syncthreads();
    for (int e = tid; e < 16 * 64; e += blockDim.x) {
        const int k = e >> 6, c = e & 63, j = j0 + c;
        s_ar[k][c] = (j < np) ? A[static_cast<int64_t>(s_r[k]) * np + j] : static_cast<FT>(0);
    }
    __syncthreads();
    const int panel = c0 >> 4;
    for (int rt = wave; rt < (np >> 4); rt += 4) {
        V4 acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = V4{0, 0, 0, 0};
        const FT* Pr = A + static_cast<int64_t>(rt * 16 + li) * np + c0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const FT a = Pr[4 * t + lk];
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) acc[ct] = mfma_16x16x4(a, s_ar[4 * t + lk][ct * 16 + li], acc[ct]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = rt * 16 + drow<FT>(lk, r);
            const bool in_r = v.mark[row] == panel;
            FT* Ar = A + static_cast<int64_t>(row) * np;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const int j = j0 + ct * 16 + li;
                if (j < np && (j < c0 || j >= c0 + 16)) Ar[j] = (in_r ? static_cast<FT>(0) : Ar[j]) + acc[ct][r];
            }
        }
    }
}

// workgroup (member, 16 output rows): K^-1[k][r_k'] = d_k T[r_k][k'] d_r_k' into Ki,
// zero on the padding
template <typename FT>
__global__ void __launch_bounds__(256) k_big_finish(PatchSlots ps, BigPatches bp, int nrb, FT* Ki)
{
    const int g = blockIdx.x / nrb, rb = blockIdx.x - g * nrb;
    const int q = bp.q[g];
    const BigView v = big_view(ps, bp, g);
    const int np = v.np, n = ps.n[q];
    if (rb * 16 >= np) return;
    const FT* A = static_cast<const FT*>(v.T);
    FT* out = Ki + ps.off2[q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = rb * 16 + wave; k < min(np, rb * 16 + 16); k += 4) {
        const FT* src = A + static_cast<int64_t>(v.piv[k]) * np;
        FT* dst = out + static_cast<int64_t>(k) * np;
        const double dk = v.d[k];
        for (int c = lane; c < np; c += 64) {
            const int pc = v.piv[c];
            dst[pc] = (k < n && pc < n) ? static_cast<FT>(dk * static_cast<double>(src[c]) * v.d[pc])
                                        : static_cast<FT>(0);
        }
    }
}

// Phase 2, one workgroup per (patch, group of sg slabs): right-hand sides of
// the patch problems of every slab and time mode in the group
// (vec_f_2d.hpp:27-78, vec_g_2d.hpp:30-95), in the Legendre time basis,
// mode kk scaled by 1 / (h |L_kk|^2), into RHS[off1 * ncols + row * ncols + col],
// col = slab * NT + kk.  The time projections Theta W and U W' of every
// (cell, slab) are formed once while staging; flux rows shared by two cells
// accumulate with LDS atomics; rows n..np-1 stay zero.
constexpr int RHS_THREADS = 256;

template <int DIM, int PP>
__global__ void __launch_bounds__(RHS_THREADS)
    k_patch_rhs(EstParams P, EstTables T, PatchTopo topo, PatchSlots ps, int q0, int sg, int nsg, int nc_max,
                const double* verts, const int* cells, const double* Mcell, const int* cell_free, const double* times,
                const double* lt, const double* x, double* RHS)
{
    using L = LocalSizes<DIM, PP>;
    constexpr int NS = L::NS, NT = L::NT, NF = L::NF, NQ = L::NQ, NV = L::NV;
    __shared__ double Rt[NV * NF * NS], Qt[NV * NQ * NS], Et[DIM * NQ * NS], Pt[NV * NQ];
    __shared__ double Wt[NT * NT], Wdt[NT * NT], Ln[NT];
    constexpr int RMAX = 32;  // slabs per workgroup with staged step sizes
    __shared__ double s_h[RMAX], s_scale[RMAX * NT];
    extern __shared__ double lds[];
    const int q = q0 + blockIdx.x / nsg;
    const int k0 = (blockIdx.x % nsg) * sg;
    const int a = topo.order[q];
    const int np = ps.np[q];
    const int nsl = min(sg, P.m - k0);
    const int ncols = P.m * NT;
    const int wcol = sg * NT;
    const int nF = topo.nflux[a];
    const int cb = topo.cell_ptr[a], nc = topo.cell_ptr[a + 1] - cb;
    double* acc = lds;                                                  // [np][sg * NT]
    double* stage = acc + static_cast<int64_t>(np) * wcol;             // [nc][sg][NS][NT][2]: Theta W, U W'
    double* geo = stage + static_cast<int64_t>(nc_max) * sg * NS * NT * 2;  // [nc][4]: det, gv
    // per cell: local vertex, flux map and flux signs of the patch
    int* c_lv = reinterpret_cast<int*>(geo + nc_max * 4);                 // [nc]
    int* c_fmap = c_lv + nc_max;                                          // [nc][NF]
    int* c_fsign = c_fmap + nc_max * NF;                                  // [nc][NF]
    // the reference tables and the step sizes: every load is issued before the
    // first store (indices clamped, sizes known at compile time), so the
    // prologue costs one memory latency instead of one per table
    {
        constexpr int NR = NV * NF * NS, NQS = NV * NQ * NS, NE = DIM * NQ * NS, NPV = NV * NQ, NW = NT * NT;
        constexpr int UR = (NR + RHS_THREADS - 1) / RHS_THREADS, UQ = (NQS + RHS_THREADS - 1) / RHS_THREADS;
        constexpr int UE = (NE + RHS_THREADS - 1) / RHS_THREADS;
        static_assert(NPV <= RHS_THREADS && NW <= RHS_THREADS && RMAX * NT <= RHS_THREADS, "one entry per thread");
        const int t = threadIdx.x;
        double vR[UR], vQ[UQ], vE[UE];
#pragma unroll
        for (int u = 0; u < UR; ++u) vR[u] = T.R[min(t + u * RHS_THREADS, NR - 1)];
#pragma unroll
        for (int u = 0; u < UQ; ++u) vQ[u] = T.Q[min(t + u * RHS_THREADS, NQS - 1)];
#pragma unroll
        for (int u = 0; u < UE; ++u) vE[u] = T.E[min(t + u * RHS_THREADS, NE - 1)];
        const double vP = T.Pv[min(t, NPV - 1)];
        const double vW = T.W[min(t, NW - 1)], vWd = T.Wd[min(t, NW - 1)];
        // step size and mode scale 1 / (h |L_kk|^2) of slab sl = t / NT of the group
        const int sl = min(t / NT, nsl - 1), kk = t % NT;
        const double t0 = times[k0 + sl], t1 = times[k0 + sl + 1], ln = T.Lnorm[kk];
#pragma unroll
        for (int u = 0; u < UR; ++u)
            if (t + u * RHS_THREADS < NR) Rt[t + u * RHS_THREADS] = vR[u];
#pragma unroll
        for (int u = 0; u < UQ; ++u)
            if (t + u * RHS_THREADS < NQS) Qt[t + u * RHS_THREADS] = vQ[u];
#pragma unroll
        for (int u = 0; u < UE; ++u)
            if (t + u * RHS_THREADS < NE) Et[t + u * RHS_THREADS] = vE[u];
        if (t < NPV) Pt[t] = vP;
        if (t < NW) {
            Wt[t] = vW;
            Wdt[t] = vWd;
        }
        if (t < NT) Ln[t] = ln;
        if (t < nsl * NT) {
            const double h = t1 - t0;
            if (kk == 0) s_h[sl] = h;
            s_scale[t] = 1.0 / (h * ln);
        }
    }
    // the patch topology: the first RHS_THREADS entries of each map in one
    // batch of loads (indices clamped; a patch has at least one cell), the rest
    // (patches of more than RHS_THREADS / NF cells) by the loops after it
    {
        const int t = threadIdx.x, nmap = nc * NF;
        const int64_t mb = static_cast<int64_t>(cb) * NF;
        const int fm = topo.fmap[mb + min(t, nmap - 1)];
        const int fs = topo.fsign[mb + min(t, nmap - 1)];
        const int lv = topo.lv[cb + min(t, nc - 1)];
        for (int e = t; e < np * wcol; e += blockDim.x) acc[e] = 0.0;
        if (t < nmap) {
            c_fmap[t] = fm;
            c_fsign[t] = fs;
        }
        if (t < nc) c_lv[t] = lv;
        for (int e = t + RHS_THREADS; e < nmap; e += RHS_THREADS) {
            c_fmap[e] = topo.fmap[mb + e];
            c_fsign[e] = topo.fsign[mb + e];
        }
        for (int ci = t + RHS_THREADS; ci < nc; ci += RHS_THREADS) c_lv[ci] = topo.lv[cb + ci];
    }
    for (int ci = threadIdx.x; ci < nc; ci += blockDim.x) {
        const int c = topo.cells[cb + ci];
        const int v = topo.lv[cb + ci];
        const Metric2 g = metric_of(DIM, c, verts, cells, Mcell);
        // grad psi_a pulled through the stiffness metric: gv = G grad^lam_v
        double gv[2] = {0.0, 0.0};
        for (int a2 = 0; a2 < DIM; ++a2)
            for (int b = 0; b < DIM; ++b) gv[a2] += g.G[a2 * 2 + b] * T.dbary[v * DIM + b];
        geo[ci * 4 + 0] = g.det;
        geo[ci * 4 + 1] = gv[0];
        geo[ci * 4 + 2] = gv[1];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nc * nsl * NS; e += blockDim.x) {
        const int ci = e / (nsl * NS), r = e % (nsl * NS);
        const int sl = r / NS, mm = r % NS;
        const int c = topo.cells[cb + ci];
        const int k = k0 + sl;
        double th[NT], uu[NT];
#pragma unroll
        for (int l = 0; l < NT; ++l) gather_cell(P, k, c, cell_free, lt, x, mm, l, th[l], uu[l]);
        double* s = stage + ((static_cast<int64_t>(ci) * sg + sl) * NS + mm) * NT * 2;
#pragma unroll
        for (int kk = 0; kk < NT; ++kk) {
            double tw = 0.0, uw = 0.0;
#pragma unroll
            for (int l = 0; l < NT; ++l) {
                tw += th[l] * Wt[l * NT + kk];
                uw += uu[l] * Wdt[l * NT + kk];
            }
            s[kk * 2] = tw;
            s[kk * 2 + 1] = uw;
        }
    }
    __syncthreads();
    constexpr int ROWS = (NF + NQ) * NT;
    // a wave holds PW (cell, slab) pairs at a time, lane = slot ROWS + rk: lane
    // owns local (row, mode) rk (rows = flux dofs, then scalar dofs) of the pairs
    // wave PW + slot + (waves PW) j.  With more than 32 rows x modes PW = 1 and
    // the lane loops over rk = lane, lane + 64, ..  Its indices and table rows are
    // fixed, and the pair advances without divisions.
    static_assert(ROWS <= 64 * 4, "local rows x modes: at most four rows per lane");
    constexpr int PW = (ROWS <= 32) ? 64 / ROWS : 1;
    const int lane64 = threadIdx.x % 64, slot = (PW > 1) ? lane64 / ROWS : 0;
    const int pstep = (blockDim.x / 64) * PW;
    for (int rk = (PW > 1) ? lane64 % ROWS : lane64; rk < ROWS && slot < PW; rk += (PW > 1) ? ROWS : 64) {
        const int row = rk / NT, kk = rk - (rk / NT) * NT;
        int ci = 0, sl = (threadIdx.x / 64) * PW + slot;
        while (sl >= nsl && ci < nc) {
            sl -= nsl;
            ++ci;
        }
        for (; ci < nc;) {
            const int v = c_lv[ci];
            const double h = s_h[sl];
            const double scale = s_scale[sl * NT + kk];
            const double* st = stage + (static_cast<int64_t>(ci) * sg + sl) * NS * NT * 2;
            if (row < NF) {
                const int fi = c_fmap[ci * NF + row];
                if (fi >= 0) {
                    // F = -h R_v (Theta W)
                    double sum = 0.0;
#pragma unroll
                    for (int mm = 0; mm < NS; ++mm) sum += Rt[(v * NF + row) * NS + mm] * st[(mm * NT + kk) * 2];
                    const int sgn = c_fsign[ci * NF + row];
                    atomicAdd(&acc[fi * wcol + sl * NT + kk], -h * sgn * sum * scale);
                }
            } else {
                const int qq = row - NF;
                const double det = geo[ci * 4 + 0];
                double sum = (kk == 0) ? P.ell * det * h * Pt[v * NQ + qq] : 0.0;
                double du = 0.0, dth = 0.0;
#pragma unroll
                for (int mm = 0; mm < NS; ++mm) {
                    du += Qt[(v * NQ + qq) * NS + mm] * st[(mm * NT + kk) * 2 + 1];
                    double eb = 0.0;
#pragma unroll
                    for (int b = 0; b < DIM; ++b) eb += geo[ci * 4 + 1 + b] * Et[(b * NQ + qq) * NS + mm];
                    dth += eb * st[(mm * NT + kk) * 2];
                }
                sum += -det * du - h * dth;
                acc[(nF + ci * NQ + qq) * wcol + sl * NT + kk] = sum * scale;
            }
            // next (cell, slab) pair of this thread: PW per wave further on
            sl += pstep;
            while (sl >= nsl && ci < nc) {
                sl -= nsl;
                ++ci;
            }
        }
    }
    __syncthreads();
    const int ncw = nsl * NT;
    double* out = RHS + ps.off1[q] * ncols + static_cast<int64_t>(k0) * NT;
    for (int e = threadIdx.x; e < np * ncw; e += blockDim.x) {
        const int row = e / ncw, c = e % ncw;
        __builtin_nontemporal_store(acc[row * wcol + c], &out[static_cast<int64_t>(row) * ncols + c]);
    }
}

// Solve workgroups: 8 waves, or 16 when the column tiles split in two halves
// (one wave per row tile and half), so that each wave fits 128 VGPRs and four
// waves share a SIMD.
__host__ __device__ constexpr int solve_threads(int CT) { return (CT % 2 == 0) ? 1024 : 512; }

// acc[ct] += M(rows rt*16.., k) B(k, cols ct*16..) over k < np.  mrow points
// at row rt*16 + l%16 of M.  The k order is permuted: at MFMA step t, lane
// group g = l/16 supplies k = kb + 16 g + t for both operands, so each lane
// reads 16 consecutive entries of its row as vector loads (one address, one
// memory latency for 16 * CT MFMAs).  A group past np supplies zeros.  B
// comes from LDS (or the workspace) with row stride ldb.
__device__ inline void load16(const float* p, float (&v)[16])
{
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const float4 w = reinterpret_cast<const float4*>(p)[u];
        v[4 * u] = w.x;
        v[4 * u + 1] = w.y;
        v[4 * u + 2] = w.z;
        v[4 * u + 3] = w.w;
    }
}
__device__ inline void load16(const double* p, double (&v)[16])
{
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const double2 w = reinterpret_cast<const double2*>(p)[u];
        v[2 * u] = w.x;
        v[2 * u + 1] = w.y;
    }
}

template <int CT, typename MT, typename BT, typename V4>
__device__ inline void tile_product(const MT* mrow, int np, const BT* Bs, int ldb, int li, int lk, V4 (&acc)[CT],
                                    bool negate)
{
    for (int kb = 0; kb < np; kb += 64) {
        const int k0 = kb + 16 * lk;
        MT av[16];
        if (k0 < np) {
            load16(mrow + k0, av);
        } else {
#pragma unroll
            for (int t = 0; t < 16; ++t) av[t] = static_cast<MT>(0);
        }
        const BT* brow = Bs + (k0 < np ? k0 : 0) * ldb + li;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const MT a = negate ? -av[t] : av[t];
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[ct] = mfma_16x16x4(a, static_cast<MT>(brow[t * ldb + ct * 16]), acc[ct]);
        }
    }
}

// ---- full-pivot retry of patches that miss the acceptance
// (solve_dense_local_error_scalar_constraint_system_2d, dense_local_error_blocks_2d.hpp:1014-1048:
// after a failed PartialPivLU solve the reference refactors with FullPivLU at threshold
// 1e-12 and solves again, and throws if the residual still fails.)

// slots q (processing order) whose patch missed the acceptance (or all, with force)
__global__ void k_fullpiv_list(int n_patches, const int* __restrict__ order, const double* __restrict__ patch_res,
                               int force, int* __restrict__ list, int* __restrict__ count)
{
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n_patches; q += gridDim.x * blockDim.x) {
        const double r = patch_res[order[q]];
        if (force || !(r <= 1e-10)) list[atomicAdd(count, 1)] = q;
    }
}

constexpr int FP_THREADS = 1024;

// block-wide argmax of (|v|, index): NaN wins, ties go to the lower index
__device__ inline void argmax_merge(double& v, int& i, double v2, int i2)
{
    const bool n1 = v != v, n2 = v2 != v2;
    bool take;
    if (n1 || n2) take = n2 && (!n1 || i2 < i);
    else take = v2 > v || (v2 == v && i2 < i);
    if (take) {
        v = v2;
        i = i2;
    }
}

__device__ inline void block_argmax(double& v, int& i, double* sv, int* si)
{
    for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off, 64);
        const int i2 = __shfl_xor(i, off, 64);
        argmax_merge(v, i, v2, i2);
    }
    const int w = threadIdx.x / 64, nw = blockDim.x / 64;
    __syncthreads();
    if (threadIdx.x % 64 == 0) {
        sv[w] = v;
        si[w] = i;
    }
    __syncthreads();
    v = sv[0];
    i = si[0];
    for (int k = 1; k < nw; ++k) argmax_merge(v, i, sv[k], si[k]);
}

// One workgroup per listed patch: Gauss-Jordan elimination with complete pivoting of
// [K | B] (the patch's FP64 matrix of order n and all its slabs x modes right-hand
// sides) in a global workspace; a pivot at most 1e-12 times the first one ends the
// elimination (rank r), and the unknowns of the remaining columns are zero, as the
// rank-revealing solve returns them.  Then the residual of the original K gives the
// acceptance statistic per (slab, vertex) problem (written to patch_res), and the flux
// coefficients go to sig as in k_patch_solve.
__global__ void __launch_bounds__(FP_THREADS)
    k_patch_fullpiv(EstParams P, PatchTopo topo, PatchSlots ps, const int* __restrict__ list, const double* Kd,
                    const double* RHS, double* ws, int64_t ws_stride, int* perm_ws, int perm_stride,
                    double* __restrict__ sig, double* __restrict__ patch_res)
{
    __shared__ double sv[FP_THREADS / 64];
    __shared__ int si[FP_THREADS / 64];
    __shared__ int s_rank;
    const int q = list[blockIdx.x];
    const int a = topo.order[q];
    const int n = ps.n[q], np = ps.np[q];
    const int Nt = P.Nt, Nf = P.Nf;
    const int ncols = P.m * Nt;
    const int ld = n + ncols;
    const double* Kq = Kd + ps.off2[q];
    const double* Bq = RHS + ps.off1[q] * ncols;
    double* W = ws + static_cast<int64_t>(blockIdx.x) * ws_stride;  // [n][ld]
    double* X = W + static_cast<int64_t>(n) * ld;                     // [n][ncols]
    unsigned long long* colmax = reinterpret_cast<unsigned long long*>(X + static_cast<int64_t>(n) * ncols);
    int* perm = perm_ws + static_cast<int64_t>(blockIdx.x) * perm_stride;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int64_t e = tid; e < static_cast<int64_t>(n) * ld; e += nt) {
        const int i = static_cast<int>(e / ld), j = static_cast<int>(e % ld);
        W[e] = (j < n) ? Kq[static_cast<int64_t>(i) * np + j] : Bq[static_cast<int64_t>(i) * ncols + (j - n)];
    }
    for (int j = tid; j < n; j += nt) perm[j] = j;
    for (int c = tid; c < 2 * ncols; c += nt) colmax[c] = 0ull;
    if (tid == 0) s_rank = n;
    __syncthreads();
    double first = 0.0;
    for (int k = 0; k < n; ++k) {
        const int m = n - k;
        double bv = -1.0;
        int bi = 0x7fffffff;
        for (int64_t e = tid; e < static_cast<int64_t>(m) * m; e += nt) {
            const int i = k + static_cast<int>(e / m), j = k + static_cast<int>(e % m);
            argmax_merge(bv, bi, fabs(W[static_cast<int64_t>(i) * ld + j]), i * n + j);
        }
        block_argmax(bv, bi, sv, si);
        if (k == 0) first = bv;
        if (!(bv > 1e-12 * first)) {
            if (tid == 0) s_rank = k;
            break;
        }
        const int pr = bi / n, pc = bi % n;
        // rows k <-> pr over all columns, columns k <-> pc over all rows
        if (pr != k)
            for (int j = tid; j < ld; j += nt) {
                const double t = W[static_cast<int64_t>(k) * ld + j];
                W[static_cast<int64_t>(k) * ld + j] = W[static_cast<int64_t>(pr) * ld + j];
                W[static_cast<int64_t>(pr) * ld + j] = t;
            }
        __syncthreads();
        if (pc != k) {
            for (int i = tid; i < n; i += nt) {
                const double t = W[static_cast<int64_t>(i) * ld + k];
                W[static_cast<int64_t>(i) * ld + k] = W[static_cast<int64_t>(i) * ld + pc];
                W[static_cast<int64_t>(i) * ld + pc] = t;
            }
            if (tid == 0) {
                const int t = perm[k];
                perm[k] = perm[pc];
                perm[pc] = t;
            }
        }
        __syncthreads();
        const double piv = W[static_cast<int64_t>(k) * ld + k];
        __syncthreads();
        for (int j = k + 1 + tid; j < ld; j += nt) W[static_cast<int64_t>(k) * ld + j] /= piv;
        __syncthreads();
        // eliminate column k from every other row (columns > k; column k is not read again)
        const int w = ld - k - 1;
        for (int64_t e = tid; e < static_cast<int64_t>(n) * w; e += nt) {
            const int i = static_cast<int>(e / w), j = k + 1 + static_cast<int>(e % w);
            if (i == k) continue;
            W[static_cast<int64_t>(i) * ld + j] -= W[static_cast<int64_t>(i) * ld + k] * W[static_cast<int64_t>(k) * ld + j];
        }
        __syncthreads();
    }
    __syncthreads();
    const int rank = s_rank;
    // x[perm[i]] = (i < rank) ? W[i][n + c] : 0
    for (int64_t e = tid; e < static_cast<int64_t>(n) * ncols; e += nt) {
        const int i = static_cast<int>(e / ncols), c = static_cast<int>(e % ncols);
        X[static_cast<int64_t>(perm[i]) * ncols + c] = (i < rank) ? W[static_cast<int64_t>(i) * ld + n + c] : 0.0;
    }
    __syncthreads();
    // residual b - K x of the original matrix, |r| and |b| maxima per column
    for (int64_t e = tid; e < static_cast<int64_t>(n) * ncols; e += nt) {
        const int i = static_cast<int>(e / ncols), c = static_cast<int>(e % ncols);
        const double b = Bq[static_cast<int64_t>(i) * ncols + c];
        double r = b;
        for (int j = 0; j < n; ++j) r -= Kq[static_cast<int64_t>(i) * np + j] * X[static_cast<int64_t>(j) * ncols + c];
        const double ar = (r != r) ? INFINITY : fabs(r);
        const double ab = (b != b) ? INFINITY : fabs(b);
        atomicMax(&colmax[c], static_cast<unsigned long long>(__double_as_longlong(ar)));
        atomicMax(&colmax[ncols + c], static_cast<unsigned long long>(__double_as_longlong(ab)));
    }
    __syncthreads();
    if (tid == 0) {
        double worst = 0.0;
        for (int s0 = 0; s0 < ncols; s0 += Nt) {
            double rm = 0.0, bm = 0.0;
            for (int c = s0; c < s0 + Nt; ++c) {
                rm = fmax(rm, __longlong_as_double(static_cast<long long>(colmax[c])));
                bm = fmax(bm, __longlong_as_double(static_cast<long long>(colmax[ncols + c])));
            }
            const double qv = (bm < INFINITY) ? rm / fmax(1.0, bm) : INFINITY;
            worst = max_nonfinite_wins(worst, qv);
        }
        patch_res[a] = worst;
    }
    // flux coefficients of the patch's cells (as k_patch_solve writes them)
    const int cb = topo.cell_ptr[a], nc = topo.cell_ptr[a + 1] - cb;
    for (int64_t e = tid; e < static_cast<int64_t>(nc) * Nf * ncols; e += nt) {
        const int c = static_cast<int>(e % ncols);
        const int row = static_cast<int>(e / ncols);
        const int ci = row / Nf, i = row - ci * Nf;
        const int fi = topo.fmap[static_cast<int64_t>(cb + ci) * Nf + i];
        const double sg = topo.fsign[static_cast<int64_t>(cb + ci) * Nf + i];
        const int64_t sc = topo.cells[cb + ci];
        const int64_t off = ((sc * P.NV + topo.lv[cb + ci]) * Nf + i) * ncols;
        sig[off + c] = (fi >= 0) ? sg * X[static_cast<int64_t>(fi) * ncols + c] : 0.0;
    }
}

// Phase 3, one workgroup per (patch, block of 16 CT columns holding whole
// slabs): X = Ki B on MFMA in FT, then refine_steps rounds of R = B - Kd X
// (FP64 MFMA) and X += Ki R (FT); the residual of the final X gives the
// acceptance statistic |K x - b|_inf / max(1, |b|_inf) per (slab, vertex)
// problem (dense_local_error_blocks_2d.hpp:923-949).  A wave item is one row
// tile and one group of CT / CS column tiles (CS = 2 with 16 waves for even CT).
// The flux coefficients go to sig.  X and R live in LDS, or, with GWS
// (patches whose tiles exceed the LDS), in this workgroup's slice of the
// global workspace ws (np x CW doubles, then np x LR FT).
template <typename FT, int CT, bool GWS>
__global__ void __launch_bounds__(solve_threads(CT))
    k_patch_solve(EstParams P, PatchTopo topo, PatchSlots ps, int q0, int count, int ncg, int cpw, const double* Kd,
                  const FT* Ki, const double* RHS, double* sig, double* patch_res, int* extra_rounds, double* ws,
                  int64_t ws_stride)
{
    using V4 = typename Vec4<FT>::type;
    constexpr int CW = 16 * CT;
    constexpr int CS = (CT % 2 == 0) ? 2 : 1;  // column-tile groups per row tile, one wave each
    constexpr int CTW = CT / CS;
    // FP32 residual tile row stride = 2 mod 4 words: the B-operand reads of
    // lane groups 0 and 1 (rows 16 apart) then fall in different bank halves
    constexpr int LR = std::is_same<FT, float>::value ? CW + 2 : CW;
    extern __shared__ double lds[];
    __shared__ unsigned long long colmax[2 * CW];  // [0, CW): |R| per column, [CW, 2 CW): |b| per column
    const int Nf = P.Nf, Nt = P.Nt;
    // Workgroups are dealt to the 8 XCDs round-robin: keep every column block
    // of one patch on the same XCD so its matrices are read into one L2.
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int qb = (slot / ncg) * 8 + xcd;
    if (qb >= count) return;
    const int q = q0 + qb;
    const int c0 = (slot % ncg) * cpw;
    const int a = topo.order[q];
    const int np = ps.np[q];
    const int ncols = P.m * Nt;
    const int ncw = min(cpw, ncols - c0);
    const double* Kq = Kd + ps.off2[q];
    const FT* Iq = Ki + ps.off2[q];
    const double* Bq = RHS + ps.off1[q] * ncols + c0;
    double* Xs = lds;  // [np][CW]
    if constexpr (GWS) Xs = ws + (static_cast<int64_t>(qb) * ncg + slot % ncg) * ws_stride;
    FT* Rs = reinterpret_cast<FT*>(Xs + static_cast<int64_t>(np) * CW);  // [np][LR]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int nrt = np >> 4;
    // B into LDS, 8 loads in flight per thread
    for (int e0 = threadIdx.x; e0 < np * CW; e0 += 8 * blockDim.x) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * blockDim.x, r = e / CW, c = e % CW;
            v[u] = (e < np * CW && c < ncw) ? Bq[static_cast<int64_t>(r) * ncols + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < np * CW) Rs[(e / CW) * LR + e % CW] = static_cast<FT>(v[u]);
        }
    }
    __syncthreads();
    // X = Ki B; wave item w = (row tile, column-tile group)
    for (int w = wave; w < nrt * CS; w += nw) {
        const int rt = w / CS, cg = (w % CS) * CTW;
        V4 acc[CTW];
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct) acc[ct] = V4{0, 0, 0, 0};
        tile_product<CTW>(Iq + static_cast<int64_t>(rt * 16 + li) * np, np, Rs + cg * 16, LR, li, lk, acc, false);
#pragma unroll
        for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                Xs[(rt * 16 + drow<FT>(lk, r)) * CW + (cg + ct) * 16 + li] = static_cast<double>(acc[ct][r]);
    }
    __syncthreads();
    // refine_steps rounds, then more (at most MAX_EXTRA) while the acceptance
    // statistic of this block's (slab, vertex) problems exceeds 1e-10
    constexpr int MAX_EXTRA = 4;
    __shared__ int s_done;
    __shared__ double s_worst;
    const int steps = P.refine_steps;
    int extra = 0;
    for (int it = 0;; ++it) {
        const bool check = it >= steps;
        if (check) {
            for (int e = threadIdx.x; e < 2 * CW; e += blockDim.x) colmax[e] = 0ull;
            __syncthreads();
        }
        // R = B - Kd X in FP64
        for (int w = wave; w < nrt * CS; w += nw) {
            const int rt = w / CS, cg = (w % CS) * CTW;
            v4d acc[CTW];
            double bm[CTW];
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct) {
                const int col = (cg + ct) * 16 + li;
                bm[ct] = 0.0;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = rt * 16 + drow<double>(lk, r);
                    acc[ct][r] = (col < ncw) ? Bq[static_cast<int64_t>(row) * ncols + col] : 0.0;
                    bm[ct] = max_nonfinite_wins(bm[ct], fabs(acc[ct][r]));
                }
            }
            tile_product<CTW>(Kq + static_cast<int64_t>(rt * 16 + li) * np, np, Xs + cg * 16, CW, li, lk, acc, true);
            if (check) {
#pragma unroll
                for (int ct = 0; ct < CTW; ++ct) {
                    double rm = 0.0, bb = bm[ct];
#pragma unroll
                    for (int r = 0; r < 4; ++r) rm = max_nonfinite_wins(rm, fabs(acc[ct][r]));
                    for (int off = 16; off < 64; off <<= 1) {
                        rm = max_nonfinite_wins(rm, __shfl_xor(rm, off, 64));
                        bb = max_nonfinite_wins(bb, __shfl_xor(bb, off, 64));
                    }
                    if (lk == 0) {
                        // non-negative doubles order like their bit patterns
                        atomicMax(&colmax[(cg + ct) * 16 + li],
                                  static_cast<unsigned long long>(__double_as_longlong(rm)));
                        atomicMax(&colmax[CW + (cg + ct) * 16 + li],
                                  static_cast<unsigned long long>(__double_as_longlong(bb)));
                    }
                }
            }
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Rs[(rt * 16 + drow<double>(lk, r)) * LR + (cg + ct) * 16 + li] = static_cast<FT>(acc[ct][r]);
        }
        __syncthreads();
        if (check) {
            // acceptance per (slab, vertex) problem: the Nt mode columns of one slab
            if (threadIdx.x == 0) {
                double worst = 0.0;
                for (int s0 = 0; s0 < ncw; s0 += Nt) {
                    double rm = 0.0, bm = 0.0;
                    for (int c = s0; c < s0 + Nt; ++c) {
                        rm = fmax(rm, __longlong_as_double(static_cast<long long>(colmax[c])));
                        bm = fmax(bm, __longlong_as_double(static_cast<long long>(colmax[CW + c])));
                    }
                    // non-finite residuals or right-hand sides fail, as the reference's
                    // finiteness checks do (dense_local_error_blocks_2d.hpp:931-951)
                    const double q = (bm < INFINITY) ? rm / fmax(1.0, bm) : INFINITY;
                    worst = max_nonfinite_wins(worst, q);
                }
                s_worst = worst;
                s_done = (worst <= 1e-10 || extra >= MAX_EXTRA) ? 1 : 0;
            }
            __syncthreads();
            if (s_done) break;
            ++extra;
        }
        // X += Ki R in FT
        for (int w = wave; w < nrt * CS; w += nw) {
            const int rt = w / CS, cg = (w % CS) * CTW;
            V4 acc[CTW];
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct) acc[ct] = V4{0, 0, 0, 0};
            tile_product<CTW>(Iq + static_cast<int64_t>(rt * 16 + li) * np, np, Rs + cg * 16, LR, li, lk, acc, false);
#pragma unroll
            for (int ct = 0; ct < CTW; ++ct)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Xs[(rt * 16 + drow<FT>(lk, r)) * CW + (cg + ct) * 16 + li] += static_cast<double>(acc[ct][r]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomic_max_pos(&patch_res[a], s_worst);
        if (extra) atomicAdd(extra_rounds, extra);
    }
    // flux coefficients per (spatial cell, local vertex, flux dof) and column:
    // one (cell, flux dof) row per wave step, lanes along the contiguous
    // columns.  Lane l first fetches the map of the wave's l-th row, so the
    // rows' topology loads are in flight together.
    const int cb = topo.cell_ptr[a], nc = topo.cell_ptr[a + 1] - cb;
    const int nrows = nc * Nf;
    for (int base = wave; base < nrows; base += nw * 64) {
        const int my = base + nw * lane;
        int my_fi = -1;
        double my_sg = 0.0;
        long long my_off = 0;
        if (my < nrows) {
            const int ci = my / Nf, i = my - ci * Nf;
            my_fi = topo.fmap[static_cast<int64_t>(cb + ci) * Nf + i];
            my_sg = topo.fsign[static_cast<int64_t>(cb + ci) * Nf + i];
            const int64_t sc = topo.cells[cb + ci];
            my_off = ((sc * P.NV + topo.lv[cb + ci]) * Nf + i) * ncols + c0;
        }
        const int cnt = min(64, (nrows - base + nw - 1) / nw);
        for (int t = 0; t < cnt; ++t) {
            const int fi = __shfl(my_fi, t, 64);
            const double sg = __shfl(my_sg, t, 64);
            double* out = sig + __shfl(my_off, t, 64);
            for (int c = lane; c < ncw; c += 64) __builtin_nontemporal_store((fi >= 0) ? sg * Xs[fi * CW + c] : 0.0, &out[c]);
        }
    }
}

// Per space-time cell terms, in exact coefficient form:
//   flux^2  = sum_k h |L_k|^2 d_k^T A_c d_k,  d = s + coef(M grad theta) (mismatch in the flux basis)
//   resid^2 = sum_k det h |L_k|^2 |rho_k|^2,   rho = projection of ell - d_t u - div s
//   recon^2 = h sum_{l,l'} Mt_{ll'} e_l^T K_c e_l',  e = lambda - lambda_tilde
//   lam^2   = same with lambda,  trace^2 = int_{t=0} (u0 - u)^2 on the first layer.
// One workgroup per spatial cell c, one thread per (slab, time mode) column of
// a chunk of slabs.  The cell's flux mass A_c, gradient projection IG and
// stiffness K_c are formed once in LDS and shared by every slab; the flux
// coefficients sig[(c, v, i)][slab * NT + mode] are read coalesced.
template <int DIM, int PP>
__global__ void __launch_bounds__(128)
    k_diag(EstParams P, EstTables T, const double* verts, const int* cells, const double* Mcell, const int* cell_free,
           const double* times, const double* lam, const double* lt, const double* x, const double* sig,
           const double* u0b, int nb, double* flux, double* resid, double* recon, double* lam2, double* trace)
{
    using L = LocalSizes<DIM, PP>;
    constexpr int NS = L::NS, NT = L::NT, NF = L::NF, NQ = L::NQ, NV = L::NV;
    constexpr int KC = 128 / NT;  // slabs per chunk
    __shared__ double Ac[NF * NF], IG[NF * NS], Kc[NS * NS];
    __shared__ double Wt[NT * NT], Wdt[NT * NT], Mtt[NT * NT], Ln[NT], Ph[NQ * NS], Bh[NQ * NF], qb[NQ];
    __shared__ double st[KC][NS][NT][4];  // theta, u, lambda, lambda - lambda_tilde
    __shared__ double colf[KC * NT], colr[KC * NT];
    const int c = blockIdx.x;
    const int tid = threadIdx.x;
    const int ncols = P.m * NT;
    const Metric2 g = metric_of(DIM, c, verts, cells, Mcell);
    for (int e = tid; e < NF * NF; e += blockDim.x) {
        const int i = e / NF, j = e % NF;
        double a = 0.0;
        for (int ab = 0; ab < DIM * DIM; ++ab)
            a += g.H[(ab / DIM) * 2 + (ab % DIM)] * T.RTab[(static_cast<int64_t>(ab) * NF + i) * NF + j];
        Ac[e] = a;
    }
    for (int e = tid; e < NF * NS; e += blockDim.x) {
        const int i = e / NS, mm = e % NS;
        double a = 0.0;
        for (int ab = 0; ab < DIM * DIM; ++ab)
            a += g.G[(ab / DIM) * 2 + (ab % DIM)] * T.Igrad[(static_cast<int64_t>(i) * DIM * DIM + ab) * NS + mm];
        IG[e] = a;
    }
    for (int e = tid; e < NS * NS; e += blockDim.x) {
        const int i = e / NS, j = e % NS;
        double a = 0.0;
        for (int ab = 0; ab < DIM * DIM; ++ab)
            a += g.G[(ab / DIM) * 2 + (ab % DIM)] * T.Khat[(static_cast<int64_t>(ab) * NS + i) * NS + j];
        Kc[e] = a;
    }
    for (int e = tid; e < NT * NT; e += blockDim.x) {
        Wt[e] = T.W[e];
        Wdt[e] = T.Wd[e];
        Mtt[e] = T.Mt[e];
    }
    for (int e = tid; e < NT; e += blockDim.x) Ln[e] = T.Lnorm[e];
    for (int e = tid; e < NQ * NS; e += blockDim.x) Ph[e] = T.Phi[e];
    for (int e = tid; e < NQ * NF; e += blockDim.x) Bh[e] = T.Bhat[e];
    for (int e = tid; e < NQ; e += blockDim.x) qb[e] = T.qbar[e];
    for (int k0 = 0; k0 < P.m; k0 += KC) {
        const int nsl = min(KC, P.m - k0);
        __syncthreads();
        for (int e = tid; e < nsl * NS * NT; e += blockDim.x) {
            const int sl = e / (NS * NT), mm = (e / NT) % NS, l = e % NT;
            const int k = k0 + sl;
            const int s = cell_free[c * NS + mm];
            double th = 0.0, uu = 0.0, la = 0.0, el = 0.0;
            if (s >= 0) {
                const int64_t yi = (static_cast<int64_t>(k) * NT + l) * P.nf + s;
                uu = x[P.nY + (static_cast<int64_t>(k) * PP + l) * P.nf + s];
                th = lt[yi] + uu;
                la = lam[yi];
                el = lam[yi] - lt[yi];
            }
            st[sl][mm][l][0] = th;
            st[sl][mm][l][1] = uu;
            st[sl][mm][l][2] = la;
            st[sl][mm][l][3] = el;
        }
        __syncthreads();
        if (tid < nsl * NT) {
            const int sl = tid / NT, kk = tid % NT;
            const int k = k0 + sl;
            const int col = k * NT + kk;
            const double h = times[k + 1] - times[k];
            double thL[NS];
#pragma unroll
            for (int mm = 0; mm < NS; ++mm) {
                double a = 0.0;
#pragma unroll
                for (int l = 0; l < NT; ++l) a += st[sl][mm][l][0] * Wt[l * NT + kk];
                thL[mm] = a / Ln[kk];
            }
            // d = s + coef(M grad theta): flux mismatch in the local flux basis
            double sv[NF], d[NF];
#pragma unroll
            for (int i = 0; i < NF; ++i) {
                double a = 0.0;
#pragma unroll
                for (int v = 0; v < NV; ++v) a += sig[((static_cast<int64_t>(c) * NV + v) * NF + i) * ncols + col];
                sv[i] = a;
                double cg = 0.0;
#pragma unroll
                for (int mm = 0; mm < NS; ++mm) cg += IG[i * NS + mm] * thL[mm];
                d[i] = a + cg;
            }
            double q2 = 0.0;
#pragma unroll
            for (int i = 0; i < NF; ++i) {
                double row = 0.0;
#pragma unroll
                for (int j = 0; j < NF; ++j) row += Ac[i * NF + j] * d[j];
                q2 += d[i] * row;
            }
            // divergence residual ell - d_t u - div s in the orthonormal scalar basis
            double r2 = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                double rho = (kk == 0) ? P.ell * qb[q] : 0.0;
                double dtu = 0.0;
#pragma unroll
                for (int mm = 0; mm < NS; ++mm) {
                    double uw = 0.0;
#pragma unroll
                    for (int l = 0; l < NT; ++l) uw += st[sl][mm][l][1] * Wdt[l * NT + kk];
                    dtu += Ph[q * NS + mm] * uw;
                }
                rho -= dtu / (Ln[kk] * h);
                double dv = 0.0;
#pragma unroll
                for (int j = 0; j < NF; ++j) dv += Bh[q * NF + j] * sv[j];
                rho -= dv / g.det;
                r2 += rho * rho;
            }
            colf[tid] = h * Ln[kk] * q2;
            colr[tid] = g.det * h * Ln[kk] * r2;
        }
        __syncthreads();
        if (tid < nsl) {
            const int sl = tid, k = k0 + sl;
            const double h = times[k + 1] - times[k];
            // Y-norm terms with K_c and the Lagrange time mass matrix
            double lq = 0.0, eq = 0.0;
#pragma unroll
            for (int i = 0; i < NS; ++i)
#pragma unroll
                for (int j = 0; j < NS; ++j) {
                    double tl = 0.0, te = 0.0;
#pragma unroll
                    for (int l = 0; l < NT; ++l)
#pragma unroll
                        for (int l2 = 0; l2 < NT; ++l2) {
                            tl += st[sl][i][l][2] * Mtt[l * NT + l2] * st[sl][j][l2][2];
                            te += st[sl][i][l][3] * Mtt[l * NT + l2] * st[sl][j][l2][3];
                        }
                    lq += Kc[i * NS + j] * tl;
                    eq += Kc[i * NS + j] * te;
                }
            double fsum = 0.0, rsum = 0.0;
#pragma unroll
            for (int kk = 0; kk < NT; ++kk) {
                fsum += colf[sl * NT + kk];
                rsum += colr[sl * NT + kk];
            }
            double tr = 0.0;
            if (k == 0)
                for (int qq = 0; qq < nb; ++qq) {
                    double uv = 0.0;
#pragma unroll
                    for (int mm = 0; mm < NS; ++mm) uv += st[0][mm][0][1] * T.bphi[qq * NS + mm];
                    const double dd = u0b[static_cast<int64_t>(c) * nb + qq] - uv;
                    tr += T.bw[qq] * g.det * dd * dd;
                }
            const int64_t cell = static_cast<int64_t>(k) * P.ns + c;
            flux[cell] = fsum;
            resid[cell] = rsum;
            recon[cell] = h * eq;
            lam2[cell] = h * lq;
            trace[cell] = tr;
        }
    }
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

// Column blocking of the patch solve at padded order np: blocks of 16 CT
// columns (CT <= 6, whole slabs) whose X / R tiles fit the LDS.  The best
// width depends on the order and the precision (two workgroups per CU and more
// column groups against fewer re-reads of the matrices: at 2+1D p=2, np = 128,
// CT = 3 beats CT = 6 in FP32; at p=3, np = 224, CT = 1 beats CT = 3), so the
// estimator times the candidates once per system (solve_candidates) and passes
// the winner as `forced`.  Without one, the widest block that fits is used.
// When even 16 columns do not fit, the tiles go to a global workspace, one
// column tile per workgroup.
struct SolveShape {
    int CT, cpw, ncg;
    bool gws;
    size_t lds_bytes;   // dynamic LDS per workgroup
    int64_t ws_stride;  // workspace doubles per workgroup (gws only)
};

namespace {
constexpr size_t SOLVE_LDS_CAP = 160 * 1024 - 4096;

// X tile [np][16 ct] doubles, R tile [np][LR] (LR = 16 ct + 2 in FP32, 16 ct in FP64)
size_t solve_tiles(int np, int ct, size_t ft_bytes)
{
    const size_t cw = 16 * ct, lr = (ft_bytes == sizeof(float)) ? cw + 2 : cw;
    return static_cast<size_t>(np) * (cw * sizeof(double) + lr * ft_bytes);
}

int solve_ct_max(int ncols)
{
    int ctmax = 6;
    if (const char* e = std::getenv("APF_SOLVE_CT_MAX")) ctmax = std::min(ctmax, std::max(1, std::atoi(e)));
    return std::min(ctmax, (ncols + 15) / 16);
}

int cols_per_group(int ct, int Nt) { return std::max(Nt, (16 * ct / Nt) * Nt); }
}  // namespace

SolveShape solve_shape(int np, int ncols, int Nt, size_t ft_bytes, int forced = 0)
{
    SolveShape sh{};
    sh.gws = solve_tiles(np, 1, ft_bytes) > SOLVE_LDS_CAP;
    int CT = solve_ct_max(ncols);
    if (sh.gws) {
        CT = 1;
    } else if (forced > 0 && solve_tiles(np, forced, ft_bytes) <= SOLVE_LDS_CAP) {
        CT = forced;
    } else {
        while (CT > 1 && solve_tiles(np, CT, ft_bytes) > SOLVE_LDS_CAP) --CT;
    }
    sh.CT = CT;
    sh.cpw = cols_per_group(CT, Nt);
    if (sh.cpw > 16 * CT) throw std::runtime_error("patch solve: a slab's modes exceed the column block");
    sh.ncg = (ncols + sh.cpw - 1) / sh.cpw;
    sh.lds_bytes = sh.gws ? 0 : solve_tiles(np, CT, ft_bytes);
    sh.ws_stride = sh.gws ? static_cast<int64_t>((solve_tiles(np, CT, ft_bytes) + sizeof(double) - 1) / sizeof(double))
                          : 0;
    return sh;
}

// Widths worth timing: tiles within the LDS and column groups that split the
// columns evenly (the last group at least half full).
std::vector<int> solve_candidates(int np, int ncols, int Nt, size_t ft_bytes)
{
    std::vector<int> out;
    if (solve_tiles(np, 1, ft_bytes) > SOLVE_LDS_CAP) return out;
    for (int ct = solve_ct_max(ncols); ct >= 1; --ct) {
        if (solve_tiles(np, ct, ft_bytes) > SOLVE_LDS_CAP) continue;
        const int cpw = cols_per_group(ct, Nt), rest = ncols % cpw;
        if (rest == 0 || 2 * rest >= cpw) out.push_back(ct);
    }
    return out;
}

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
    // processing order: largest patches first; slots padded to multiples of
    // 16 (the MFMA tile), one size bucket per padded order
    std::vector<int> pn(P.n_patches);
    for (int a = 0; a < P.n_patches; ++a) pn[a] = P.n_flux[a] + P.n_scalar[a] + (P.boundary[a] ? 0 : 1);
    std::vector<int> order(P.n_patches);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pn[a] > pn[b]; });
    B.order = upi(order);
    std::vector<int> sn(P.n_patches), snp(P.n_patches);
    std::vector<int64_t> off2(P.n_patches), off1(P.n_patches);
    int64_t o2 = 0, o1 = 0;
    B.buckets.clear();
    for (int q = 0; q < P.n_patches; ++q) {
        const int a = order[q];
        sn[q] = pn[a];
        snp[q] = (pn[a] + 15) / 16 * 16;
        off2[q] = o2;
        off1[q] = o1;
        o2 += static_cast<int64_t>(snp[q]) * snp[q];
        o1 += snp[q];
        const int nc = P.cell_ptr[a + 1] - P.cell_ptr[a];
        if (B.buckets.empty() || B.buckets.back().np != snp[q]) B.buckets.push_back({q, 0, snp[q], 0, 0, false});
        B.buckets.back().count++;
        B.buckets.back().nc_max = std::max(B.buckets.back().nc_max, nc);
        B.buckets.back().blk_max = std::max({B.buckets.back().blk_max, P.n_flux[a], pn[a] - P.n_flux[a]});
    }
    // buckets beyond the single-workgroup factor kernels (order > 128 with a
    // diagonal block > 128) and within the 512 x 4 rows of the panel kernel
    std::vector<int> bq;
    std::vector<int64_t> boff;
    int64_t bo = 0;
    for (auto& bk : B.buckets) {
        bk.big = bk.np > 128 && bk.blk_max > 128 && bk.np <= 2048;
        if (!bk.big) continue;
        for (int q = bk.q0; q < bk.q0 + bk.count; ++q) {
            bq.push_back(q);
            boff.push_back(bo);
            bo += static_cast<int64_t>(snp[q]) * snp[q] + 2 * snp[q];
        }
        B.big_npmax = std::max(B.big_npmax, bk.np);
    }
    B.big_count = static_cast<int>(bq.size());
    B.s_n = upi(sn);
    B.s_np = upi(snp);
    auto up64 = [&](const std::vector<int64_t>& v) {
        int64_t* p = dalloc<int64_t>(sys, v.size());
        if (!v.empty()) HIP_TRY(hipMemcpy(p, v.data(), v.size() * sizeof(int64_t), hipMemcpyHostToDevice));
        return p;
    };
    B.s_off2 = up64(off2);
    B.s_off1 = up64(off1);
    if (B.big_count) {
        B.big_q = upi(bq);
        B.big_off = up64(boff);
        B.big_ws = dalloc<double>(sys, static_cast<size_t>(bo));
    }
    B.nmax = P.n_patches ? snp[0] : 0;
    B.Kd = dalloc<double>(sys, static_cast<size_t>(o2));
    B.Ki = dalloc<double>(sys, static_cast<size_t>(o2));
    B.Ki32 = dalloc<float>(sys, static_cast<size_t>(o2));
    B.RHS = dalloc<double>(sys, static_cast<size_t>(o1) * S.mesh.m() * T.Nt);
    HIP_TRY(hipMemset(B.Kd, 0, static_cast<size_t>(o2) * sizeof(double)));
    HIP_TRY(hipMemset(B.Ki, 0, static_cast<size_t>(o2) * sizeof(double)));
    HIP_TRY(hipMemset(B.Ki32, 0, static_cast<size_t>(o2) * sizeof(float)));
    HIP_TRY(hipMemset(B.RHS, 0, static_cast<size_t>(o1) * S.mesh.m() * T.Nt * sizeof(double)));
    // one workspace for the buckets whose solve tiles exceed the LDS (buckets
    // run one after another on the stream), sized for either precision
    int64_t ws = 0;
    for (const auto& bk : B.buckets)
        for (size_t ft : {sizeof(double), sizeof(float)}) {
            const SolveShape sh = solve_shape(bk.np, S.mesh.m() * T.Nt, T.Nt, ft);
            if (sh.gws) ws = std::max(ws, static_cast<int64_t>(bk.count) * sh.ncg * sh.ws_stride);
        }
    B.solve_ws = ws ? dalloc<double>(sys, static_cast<size_t>(ws)) : nullptr;
    B.solve_ws_doubles = ws;
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
    B.extra_rounds = dalloc<int>(sys, 1);
    B.tune_extra = dalloc<int>(sys, 1);
}

}  // namespace apf

extern "C" int apf_estimate(apf_system* sys, int prec, int refine_steps, apf_estimate_out* out, double* flux_by_cell,
                            double* resid_by_cell, double* recon_by_cell)
{
    return apf_estimate_ex(sys, prec, refine_steps, APF_SLAB_FROM_SOLUTION, out, flux_by_cell, resid_by_cell,
                           recon_by_cell);
}

extern "C" int apf_estimate_ex(apf_system* sys, int prec, int refine_steps, int slab_mode, apf_estimate_out* out,
                               double* flux_by_cell, double* resid_by_cell, double* recon_by_cell)
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
    if (slab_mode != APF_SLAB_FROM_SOLUTION && slab_mode != APF_SLAB_SOLVE)
        return fail(APF_BAD_ARG, "apf_estimate: unknown slab reconstruction mode");
    const Space& S = *sys->space;
    const RefTables& T = *sys->tables;
    if (S.p > 3) return fail(APF_BAD_ARG, "apf_estimate: the estimator supports p <= 3");
    EstimatorBuffers& B = sys->est;
    try {
        HIP_TRY(hipSetDevice(sys->ctx->device));
        hipStream_t st = sys->stream;
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
        P.refine_steps = (prec == APF_FP32_REFINE) ? refine_steps : 0;
        EstTables Tb{B.RTab, B.Bhat, B.qbar, B.R, B.Pv, B.Q, B.E, B.Phi, B.Khat, B.dbary, B.Wt, B.Wdt, B.Lnorm_, B.Mt,
                     B.Igrad, sys->d_bw, sys->d_bphi};
        PatchTopo topo{B.p_cell_ptr, B.p_cells, B.p_lv, B.p_fmap, B.p_fsign, B.p_nflux, B.p_nscal, B.p_boundary, B.order};
        // slab reconstruction: on a tensor mesh the slab space is Y restricted to
        // the slab, so lambda_tilde solves the lambda rows of the main system
        // slab by slab; with the main solve converged it equals lambda.
        out->slab_iterations = 0;
        out->slab_relative_residual = -1.0;
        out->slab_stop = APF_SLAB_STOP_NONE;
        hipEvent_t et0, et1;  // device span of the whole estimate
        HIP_TRY(hipEventCreate(&et0));
        HIP_TRY(hipEventCreate(&et1));
        HIP_TRY(hipEventRecord(et0, st));
        auto slab_phase = [&]() {
            const auto t_slab = std::chrono::steady_clock::now();
            if (slab_mode == APF_SLAB_SOLVE)
                slab_reconstruct(sys, B.lt, 1e-12, 100, &out->slab_iterations, &out->slab_relative_residual,
                                 &out->slab_stop);
            else
                HIP_TRY(hipMemcpyAsync(B.lt, sys->d_x, S.nY * sizeof(double), hipMemcpyDeviceToDevice, st));
            out->seconds_slab = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_slab).count();
        };
        int phase = 0;  // launch(): 0 all, 1 factor only, 2 right-hand sides and solves
        HIP_TRY(hipMemsetAsync(B.patch_res, 0, S.patches.n_patches * sizeof(double), st));
        HIP_TRY(hipMemsetAsync(B.extra_rounds, 0, sizeof(int), st));
        const size_t lds_cap = 160 * 1024;
        PatchSlots ps{B.s_n, B.s_np, B.s_off2, B.s_off1};
        const int ncols = P.m * P.Nt;
        auto set_lds = [&](const void* fn, size_t bytes) {
            HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
        };
        static const bool fp64_blocked = [] {
            const char* e = std::getenv("APF_FP64_BLOCKED");
            return !(e && std::atoi(e) == 0);
        }();
        auto launch = [&](auto tag) {
            using FT = decltype(tag);
            // the factor of one bucket in precision F2 (FP32: blocked Gauss-Jordan with MFMA
            // panel updates up to order 128; otherwise one register-resident Gauss-Jordan
            // where the matrix fits the register file, else block elimination with two)
            auto factor_bucket = [&](const EstimatorBuffers::Bucket& bk, auto ftag) {
                using F2 = decltype(ftag);
                F2* K2 = std::is_same<F2, float>::value ? reinterpret_cast<F2*>(B.Ki32) : reinterpret_cast<F2*>(B.Ki);
                // blocked Gauss-Jordan with MFMA panel updates up to order 128 (FP64 too
                // unless APF_FP64_BLOCKED=0, which takes the register-resident one)
                if (std::is_same<F2, float>::value || fp64_blocked) {
                    if (bk.np <= 64) {
                        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor_blocked<F2, 1>), dim3(bk.count), dim3(256),
                                           0, st, P, Tb, topo, ps, bk.q0, sys->d_verts, sys->d_cells, sys->d_Mcell,
                                           B.Kd, K2);
                        return;
                    }
                    if (bk.np <= 128) {
                        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor_blocked<F2, 2>), dim3(bk.count), dim3(256),
                                           0, st, P, Tb, topo, ps, bk.q0, sys->d_verts, sys->d_cells, sys->d_Mcell,
                                           B.Kd, K2);
                        return;
                    }
                }
                if (bk.np <= 64) {
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor_full<F2, 4, 16, 1>), dim3(bk.count), dim3(256), 0,
                                       st, P, Tb, topo, ps, bk.q0, sys->d_verts, sys->d_cells, sys->d_Mcell, B.Kd, K2);
                    return;
                }
                if (bk.np <= 128) {
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor_full<F2, 4, 32, 2>), dim3(bk.count), dim3(256), 0,
                                       st, P, Tb, topo, ps, bk.q0, sys->d_verts, sys->d_cells, sys->d_Mcell, B.Kd, K2);
                    return;
                }
                if (bk.blk_max <= 128) {
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor_schur<F2, 8, 16, 2>), dim3(bk.count), dim3(512), 0,
                                       st, P, Tb, topo, ps, bk.q0, sys->d_verts, sys->d_cells, sys->d_Mcell, B.Kd, B.Ki,
                                       B.Ki32);
                    return;
                }
                const size_t npb = static_cast<size_t>(bk.np);
                const size_t in_lds_bytes = npb * npb * sizeof(F2) + npb * sizeof(int);
                const int in_lds = in_lds_bytes <= lds_cap ? 1 : 0;
                const size_t lf = in_lds ? in_lds_bytes : npb * sizeof(int);
                set_lds(reinterpret_cast<const void*>(&k_patch_factor<F2>), lf);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_patch_factor<F2>), dim3(bk.count), dim3(PATCH_THREADS), lf, st, P,
                                   Tb, topo, ps, bk.q0, in_lds, sys->d_verts, sys->d_cells, sys->d_Mcell, B.Kd, K2);
            };
            // mixed path: a bucket whose FP64 factor and single solve measured faster
            auto use_fp64 = [&](const EstimatorBuffers::Bucket& bk) {
                if (!std::is_same<FT, float>::value || bk.big) return false;
                const auto it = sys->mixed_fp64.find(bk.np);
                return it != sys->mixed_fp64.end() && it->second;
            };
            if (phase != 2)
                for (const auto& bk : B.buckets) {
                    if (bk.big) continue;
                    if (use_fp64(bk)) factor_bucket(bk, double{});
                    else factor_bucket(bk, FT{});
                }
            // the large patches of all big buckets in precision F2: blocked Gauss-Jordan,
            // one panel per pair of launches
            auto big_factor = [&](auto ftag) {
                using F2 = decltype(ftag);
                F2* K2 = std::is_same<F2, float>::value ? reinterpret_cast<F2*>(B.Ki32) : reinterpret_cast<F2*>(B.Ki);
                const BigPatches bp{B.big_q, B.big_off, B.big_ws};
                const int G = B.big_count, npm = B.big_npmax;
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_prepare<F2>), dim3(G), dim3(PATCH_THREADS), 0, st, P, Tb,
                                   topo, ps, bp, sys->d_verts, sys->d_cells, sys->d_Mcell, B.Kd);
                const int nstrips = (npm + 63) / 64, rpt = (npm + 511) / 512;
                for (int c0 = 0; c0 < npm; c0 += 16) {
                    switch (rpt) {
                        case 1: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_panel<F2, 1>), dim3(G), dim3(512), 0, st, ps, bp, c0); break;
                        case 2: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_panel<F2, 2>), dim3(G), dim3(512), 0, st, ps, bp, c0); break;
                        case 3: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_panel<F2, 3>), dim3(G), dim3(512), 0, st, ps, bp, c0); break;
                        default: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_panel<F2, 4>), dim3(G), dim3(512), 0, st, ps, bp, c0); break;
                    }
                    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_update<F2>), dim3(G * nstrips), dim3(256), 0, st, ps, bp,
                                       c0, nstrips);
                }
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_big_finish<F2>), dim3(G * (npm / 16)), dim3(256), 0, st, ps, bp,
                                   npm / 16, K2);
            };
            // mixed path: the big buckets as one group (their factor runs together)
            auto big_fp64 = [&]() {
                if (!std::is_same<FT, float>::value) return false;
                const auto it = sys->mixed_fp64.find(-1);
                return it != sys->mixed_fp64.end() && it->second;
            };
            if (B.big_count && phase != 2) {
                if (big_fp64()) big_factor(double{});
                else big_factor(FT{});
            }
            if (phase != 2) HIP_TRY(hipEventRecord(sys->evp[1], st));
            if (phase == 1) return;
            for (const auto& bk : B.buckets) {
                auto lr = [&](int g) {
                    return (static_cast<size_t>(bk.np) * g * P.Nt +
                            static_cast<size_t>(bk.nc_max) * g * P.Ns * P.Nt * 2 + bk.nc_max * 4) * sizeof(double) +
                           static_cast<size_t>(bk.nc_max) * (1 + 2 * P.Nf) * sizeof(int);
                };
                // one pass of the bucket with sg slabs per workgroup
                auto run = [&](int sg) {
                    if (lr(sg) > lds_cap) throw std::runtime_error("patch right-hand side block exceeds LDS");
                    const int nsg = (P.m + sg - 1) / sg;
                    auto rhs = [&](auto kern) {
                        set_lds(reinterpret_cast<const void*>(kern), lr(sg));
                        hipLaunchKernelGGL(kern, dim3(bk.count * nsg), dim3(RHS_THREADS), lr(sg), st, P, Tb, topo, ps,
                                           bk.q0, sg, nsg, bk.nc_max, sys->d_verts, sys->d_cells, sys->d_Mcell,
                                           sys->d_cell_free, sys->d_times, B.lt, sys->d_x, B.RHS);
                    };
                    switch (S.dim * 10 + S.p) {
                        case 11: rhs(k_patch_rhs<1, 1>); break;
                        case 12: rhs(k_patch_rhs<1, 2>); break;
                        case 13: rhs(k_patch_rhs<1, 3>); break;
                        case 21: rhs(k_patch_rhs<2, 1>); break;
                        case 22: rhs(k_patch_rhs<2, 2>); break;
                        case 23: rhs(k_patch_rhs<2, 3>); break;
                        default: throw std::runtime_error("patch right-hand sides: unsupported (dim, p)");
                    }
                };
                // slabs per workgroup (1, 2, 4, .., 32 within 64 KiB of LDS; small patches
                // amortise the workgroup's table and map loads over more slabs): timed once
                // per system and padded order, each candidate twice (the passes only write RHS)
                auto it = sys->rhs_sg.find(bk.np);
                if (it == sys->rhs_sg.end()) {
                    std::vector<int> cand;
                    for (int g = 32; g >= 1; g /= 2)
                        if (g <= P.m && (lr(g) <= 64 * 1024 || g == 1)) cand.push_back(g);
                    const char* tune_env = std::getenv("APF_SOLVE_TUNE");
                    int best = cand.empty() ? 1 : cand.front();
                    if (cand.size() > 1 && !(tune_env && std::atoi(tune_env) == 0)) {
                        hipEvent_t ta, tb;
                        HIP_TRY(hipEventCreate(&ta));
                        HIP_TRY(hipEventCreate(&tb));
                        float best_ms = 0.0f;
                        best = 0;
                        for (int g : cand) {
                            float ms = 0.0f;
                            for (int rep = 0; rep < 2; ++rep) {
                                HIP_TRY(hipEventRecord(ta, st));
                                run(g);
                                HIP_TRY(hipEventRecord(tb, st));
                                HIP_TRY(hipEventSynchronize(tb));
                                float t = 0.0f;
                                HIP_TRY(hipEventElapsedTime(&t, ta, tb));
                                ms = (rep == 0) ? t : std::min(ms, t);
                            }
                            if (best == 0 || ms < best_ms) {
                                best = g;
                                best_ms = ms;
                            }
                        }
                        (void)hipEventDestroy(ta);
                        (void)hipEventDestroy(tb);
                    }
                    it = sys->rhs_sg.emplace(bk.np, best).first;
                }
                run(it->second);
            }
            HIP_TRY(hipEventRecord(sys->evp[2], st));
            // one solve of bucket bk in precision ST (inverse Ki32 or Ki); `counter` receives
            // the refinement rounds beyond refine_steps
            auto solve_bucket = [&](const EstimatorBuffers::Bucket& bk, auto stag, int* counter) {
                using ST = decltype(stag);
                ST* Ks = std::is_same<ST, float>::value ? reinterpret_cast<ST*>(B.Ki32) : reinterpret_cast<ST*>(B.Ki);
                EstParams Ps = P;
                if (!std::is_same<ST, FT>::value) Ps.refine_steps = 0;
                // one solve of the bucket with column blocking sh; `extra` counts refinement rounds
                auto run = [&](const SolveShape& sh, int* extra) {
                    const unsigned grid = static_cast<unsigned>((bk.count + 7) / 8 * 8 * sh.ncg);
                    if (sh.gws && static_cast<int64_t>(bk.count) * sh.ncg * sh.ws_stride > B.solve_ws_doubles)
                        throw std::runtime_error("patch solve: workspace smaller than the column blocking needs");
                    auto solve = [&](auto kern) {
                        if (!sh.gws) set_lds(reinterpret_cast<const void*>(kern), sh.lds_bytes);
                        // one wave per item (row tile, column-tile group) up to the kernel's width:
                        // the kernel loops over its waves, and small patches then leave none idle
                        const int items = (bk.np / 16) * ((sh.CT % 2 == 0) ? 2 : 1);
                        const int threads = std::min(solve_threads(sh.CT), std::max(1, items) * 64);
                        hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), sh.lds_bytes, st, Ps, topo, ps,
                                           bk.q0, bk.count, sh.ncg, sh.cpw, B.Kd, Ks, B.RHS, B.sig, B.patch_res, extra,
                                           B.solve_ws, sh.ws_stride);
                    };
                    if (sh.gws) {
                        switch (sh.CT) {
                            case 1: solve(k_patch_solve<ST, 1, true>); break;
                            case 2: solve(k_patch_solve<ST, 2, true>); break;
                            case 3: solve(k_patch_solve<ST, 3, true>); break;
                            case 4: solve(k_patch_solve<ST, 4, true>); break;
                            case 5: solve(k_patch_solve<ST, 5, true>); break;
                            default: solve(k_patch_solve<ST, 6, true>); break;
                        }
                        return;
                    }
                    switch (sh.CT) {
                        case 1: solve(k_patch_solve<ST, 1, false>); break;
                        case 2: solve(k_patch_solve<ST, 2, false>); break;
                        case 3: solve(k_patch_solve<ST, 3, false>); break;
                        case 4: solve(k_patch_solve<ST, 4, false>); break;
                        case 5: solve(k_patch_solve<ST, 5, false>); break;
                        default: solve(k_patch_solve<ST, 6, false>); break;
                    }
                };
                // the column blocking of this (order, precision): timed once per system
                // (each candidate twice, the faster run counts); the trial solves
                // compute the same result and count their rounds elsewhere
                const std::pair<int, int> key{bk.np, static_cast<int>(sizeof(ST))};
                auto it = sys->solve_ct.find(key);
                if (it == sys->solve_ct.end()) {
                    const std::vector<int> cand = solve_candidates(bk.np, ncols, P.Nt, sizeof(ST));
                    const char* tune_env = std::getenv("APF_SOLVE_TUNE");
                    int best = 0;
                    if (cand.size() > 1 && !(tune_env && std::atoi(tune_env) == 0)) {
                        hipEvent_t ta, tb;
                        HIP_TRY(hipEventCreate(&ta));
                        HIP_TRY(hipEventCreate(&tb));
                        float best_ms = 0.0f;
                        for (int ct : cand) {
                            const SolveShape shc = solve_shape(bk.np, ncols, P.Nt, sizeof(ST), ct);
                            float ms = 0.0f;
                            for (int rep = 0; rep < 2; ++rep) {
                                HIP_TRY(hipEventRecord(ta, st));
                                run(shc, B.tune_extra);
                                HIP_TRY(hipEventRecord(tb, st));
                                HIP_TRY(hipEventSynchronize(tb));
                                float t = 0.0f;
                                HIP_TRY(hipEventElapsedTime(&t, ta, tb));
                                ms = (rep == 0) ? t : std::min(ms, t);
                            }
                            if (best == 0 || ms < best_ms) {
                                best = ct;
                                best_ms = ms;
                            }
                        }
                        (void)hipEventDestroy(ta);
                        (void)hipEventDestroy(tb);
                    }
                    it = sys->solve_ct.emplace(key, best).first;
                }
                run(solve_shape(bk.np, ncols, P.Nt, sizeof(ST), it->second), counter);
            };
            // first mixed call for a bucket (or the group of big buckets): time the FP32
            // factor with the refined solve against the FP64 factor with one solve, each
            // after an untimed pass so that neither includes its tile trials (the trial
            // solves count their rounds apart); true if the FP64 pair is faster.  The FP32
            // factor of the first pass already ran in the factor phase.
            auto faster_fp64 = [&](auto fact32, auto fact64, auto solves) {
                solves(float{}, B.tune_extra);
                fact64();
                solves(double{}, B.tune_extra);
                hipEvent_t e[3];
                for (auto& ev : e) HIP_TRY(hipEventCreate(&ev));
                HIP_TRY(hipEventRecord(e[0], st));
                fact32();
                solves(float{}, B.tune_extra);
                HIP_TRY(hipEventRecord(e[1], st));
                fact64();
                solves(double{}, B.tune_extra);
                HIP_TRY(hipEventRecord(e[2], st));
                HIP_TRY(hipEventSynchronize(e[2]));
                float t32 = 0.0f, t64 = 0.0f;
                HIP_TRY(hipEventElapsedTime(&t32, e[0], e[1]));
                HIP_TRY(hipEventElapsedTime(&t64, e[1], e[2]));
                for (auto& ev : e) (void)hipEventDestroy(ev);
                return t64 < t32;
            };
            const char* tune_env = std::getenv("APF_SOLVE_TUNE");
            const bool tune = std::is_same<FT, float>::value && !(tune_env && std::atoi(tune_env) == 0);
            const bool big_tune = tune && B.big_count && !sys->mixed_fp64.count(-1);
            for (const auto& bk : B.buckets) {
                if (!std::is_same<FT, float>::value) {
                    solve_bucket(bk, FT{}, B.extra_rounds);
                    continue;
                }
                if (bk.big) {
                    // solved after the loop while the group's precision is being chosen
                    if (!big_tune) {
                        if (big_fp64()) solve_bucket(bk, double{}, B.extra_rounds);
                        else solve_bucket(bk, float{}, B.extra_rounds);
                    }
                    continue;
                }
                // a bucket whose inverse is FP64 in both precisions (block elimination
                // writes it in FP64; its FP32 copy is only a rounding) is solved once in
                // FP64: FP32 products and refinement rounds would only add work
                const bool fp64_inverse = bk.np > 128 && bk.blk_max <= 128;
                if (fp64_inverse || use_fp64(bk)) {
                    solve_bucket(bk, double{}, B.extra_rounds);
                    continue;
                }
                if (!tune || sys->mixed_fp64.count(bk.np)) {
                    solve_bucket(bk, float{}, B.extra_rounds);
                    continue;
                }
                // the chosen pair runs last, so the outputs and the round count are those
                // of a later call
                auto solves = [&](auto stag, int* counter) { solve_bucket(bk, stag, counter); };
                const bool fp64 = faster_fp64([&] { factor_bucket(bk, float{}); }, [&] { factor_bucket(bk, double{}); },
                                              solves);
                sys->mixed_fp64[bk.np] = fp64;
                if (fp64) {
                    factor_bucket(bk, double{});
                    solves(double{}, B.extra_rounds);
                } else {
                    factor_bucket(bk, float{});
                    solves(float{}, B.extra_rounds);
                }
            }
            if (big_tune) {
                auto big_solves = [&](auto stag, int* counter) {
                    for (const auto& b2 : B.buckets)
                        if (b2.big) solve_bucket(b2, stag, counter);
                };
                const bool fp64 = faster_fp64([&] { big_factor(float{}); }, [&] { big_factor(double{}); }, big_solves);
                sys->mixed_fp64[-1] = fp64;
                if (fp64) {
                    big_factor(double{});
                    big_solves(double{}, B.extra_rounds);
                } else {
                    big_factor(float{});
                    big_solves(float{}, B.extra_rounds);
                }
            }
        };
        // The patch matrices and their factors do not read lambda_tilde: after a system's
        // first estimate in this precision (which times its tile and precision candidates),
        // the factor phase can run on a second stream while the slab reconstruction, a chain
        // of small latency-bound launches, runs on the system stream.  Measured at 2+1D p=2
        // (the slab solve in 2-3 CG steps with FP32 V-cycles): mixed 3.16 -> 2.92 ms, FP64
        // 3.62 -> 3.06 ms; the slab solve stretches from 1.02 to 1.41 ms while the factor's
        // workgroups hold the CUs' LDS.  (With the slower 6-step slab solve the mixed estimate
        // lost: 4.15 -> 5.22 ms.)  APF_EST_OVERLAP=0 / 1 forces either way.
        bool overlap = slab_mode == APF_SLAB_SOLVE && sys->est_done.count(prec) > 0;
        if (const char* e = std::getenv("APF_EST_OVERLAP"))
            overlap = slab_mode == APF_SLAB_SOLVE && sys->est_done.count(prec) > 0 && std::atoi(e) != 0;
        if (overlap) {
            slab_prepare(sys);  // host set-up and allocations before the streams fork
            hipEvent_t fork, fdone;
            HIP_TRY(hipEventCreate(&fork));
            HIP_TRY(hipEventCreate(&fdone));
            HIP_TRY(hipEventRecord(fork, st));
            HIP_TRY(hipStreamWaitEvent(sys->stream2, fork, 0));
            const hipStream_t main_st = st;
            st = sys->stream2;
            HIP_TRY(hipEventRecord(sys->ev0, st));
            phase = 1;
            if (prec == APF_FP64) launch(double{});
            else launch(float{});
            HIP_TRY(hipEventRecord(fdone, st));
            st = main_st;
            slab_phase();
            HIP_TRY(hipStreamWaitEvent(st, fdone, 0));
            phase = 2;
            if (prec == APF_FP64) launch(double{});
            else launch(float{});
            (void)hipEventDestroy(fork);
            (void)hipEventDestroy(fdone);
        } else {
            slab_phase();
            HIP_TRY(hipEventRecord(sys->ev0, st));
            if (prec == APF_FP64) launch(double{});
            else launch(float{});
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(sys->ev1, st));
        const int64_t ncell = static_cast<int64_t>(P.m) * P.ns;
        hipEvent_t e2, e3;
        HIP_TRY(hipEventCreate(&e2));
        HIP_TRY(hipEventCreate(&e3));
        double h[6];
        int extra_rounds = 0;
        float ms2 = 0.f;
        // diagnostics, sums and the acceptance maximum (run again after a full-pivot retry)
        auto finish = [&]() {
            HIP_TRY(hipEventRecord(e2, st));
            auto diag = [&](auto kern) {
                hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(P.ns)), dim3(128), 0, st, P, Tb, sys->d_verts,
                                   sys->d_cells, sys->d_Mcell, sys->d_cell_free, sys->d_times, sys->d_x, B.lt, sys->d_x,
                                   B.sig, sys->d_u0b, T.nb, B.flux, B.resid, B.recon, B.lam2, B.trace);
            };
            const int key = S.dim * 10 + S.p;
            switch (key) {
                case 11: diag(k_diag<1, 1>); break;
                case 12: diag(k_diag<1, 2>); break;
                case 13: diag(k_diag<1, 3>); break;
                case 21: diag(k_diag<2, 1>); break;
                case 22: diag(k_diag<2, 2>); break;
                case 23: diag(k_diag<2, 3>); break;
                default: throw std::runtime_error("estimator diagnostics: unsupported (dim, p)");
            }
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipEventRecord(e3, st));
            HIP_TRY(hipEventRecord(et1, st));
            const int nparts = 256;
            hipLaunchKernelGGL(k_sum5, dim3(nparts), dim3(256), 0, st, ncell, B.flux, B.resid, B.recon, B.lam2, B.trace,
                               sys->d_part);
            hipLaunchKernelGGL(k_sum5_final, dim3(1), dim3(64), 0, st, sys->d_part, nparts, B.sums);
            hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, st, S.patches.n_patches, B.patch_res, B.sums + 5);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipMemcpyAsync(h, B.sums, 6 * sizeof(double), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipMemcpyAsync(&extra_rounds, B.extra_rounds, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            HIP_TRY(hipEventElapsedTime(&ms2, e2, e3));
        };
        finish();
        float ms1 = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms1, sys->ev0, sys->ev1));
        HIP_TRY(hipEventElapsedTime(&sys->t_factor_ms, sys->ev0, sys->evp[1]));
        HIP_TRY(hipEventElapsedTime(&sys->t_rhs_ms, sys->evp[1], sys->evp[2]));
        HIP_TRY(hipEventElapsedTime(&sys->t_solve_ms, sys->evp[2], sys->ev1));
        // patches that missed the acceptance: solve again with complete pivoting
        // (APF_PATCH_FULLPIV=1 sends every patch through this path)
        bool force = false;
        if (const char* e = std::getenv("APF_PATCH_FULLPIV")) force = std::atoi(e) != 0;
        int fullpiv = 0;
        double fullpiv_s = 0.0;
        if (force || !(h[5] <= 1e-10)) {
            const auto t_fp = std::chrono::steady_clock::now();
            const int npat = S.patches.n_patches;
            if (!B.fp_list) {
                B.fp_list = dalloc<int>(sys, static_cast<size_t>(npat) + 1);
                B.fp_count = dalloc<int>(sys, 1);
            }
            HIP_TRY(hipMemsetAsync(B.fp_count, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_fullpiv_list, dim3(static_cast<unsigned>(std::min(1024, (npat + 255) / 256 + 1))),
                               dim3(256), 0, st, npat, B.order, B.patch_res, force ? 1 : 0, B.fp_list, B.fp_count);
            HIP_TRY(hipMemcpyAsync(&fullpiv, B.fp_count, sizeof(int), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            if (fullpiv > 0) {
                const int ncols = P.m * P.Nt, nm = B.nmax;
                const int64_t per = static_cast<int64_t>(nm) * (nm + ncols) + static_cast<int64_t>(nm) * ncols +
                                    2 * static_cast<int64_t>(ncols);
                const int batch = static_cast<int>(
                    std::max<int64_t>(1, std::min<int64_t>(fullpiv, (int64_t(1) << 29) / (per * 8))));
                if (static_cast<int64_t>(batch) * per > B.fp_ws_doubles) {
                    B.fp_ws = dalloc<double>(sys, static_cast<size_t>(batch) * per);
                    B.fp_perm = dalloc<int>(sys, static_cast<size_t>(batch) * nm);
                    B.fp_ws_doubles = static_cast<int64_t>(batch) * per;
                }
                for (int b0 = 0; b0 < fullpiv; b0 += batch)
                    hipLaunchKernelGGL(k_patch_fullpiv, dim3(static_cast<unsigned>(std::min(batch, fullpiv - b0))),
                                       dim3(FP_THREADS), 0, st, P, topo, ps, B.fp_list + b0, B.Kd, B.RHS, B.fp_ws, per,
                                       B.fp_perm, nm, B.sig, B.patch_res);
                HIP_TRY(hipGetLastError());
                finish();
            }
            fullpiv_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_fp).count();
        }
        float mst = 0.f;
        HIP_TRY(hipEventElapsedTime(&mst, et0, et1));
        (void)hipEventDestroy(e2);
        (void)hipEventDestroy(e3);
        (void)hipEventDestroy(et0);
        (void)hipEventDestroy(et1);
        sys->est_done.insert(prec);
        out->seconds_total = mst * 1e-3;  // et1 is recorded again after a full-pivot retry
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
        out->extra_refinements = extra_rounds;
        out->fullpiv_patches = fullpiv;
        out->seconds_patches = ms1 * 1e-3 + fullpiv_s;
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
    } catch (const std::exception& e) {
        return fail(APF_BAD_ARG, std::string("apf_estimate: ") + e.what());
    }
}
