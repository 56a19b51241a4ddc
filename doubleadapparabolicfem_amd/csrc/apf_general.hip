// Element-local quadrature assembly of the saddle-point system on space-time meshes that
// are not tensor products: cells given one by one with their cell restriction rows
// (hanging-node constraints, Dirichlet rows), Y cells possibly finer than their X ancestor.
//
// Restates assemble_main_y_dense_local_blocks_2d
// (src/finite_element/assembly/main_system/dense_local_main_blocks_2d.hpp:176-361) and the
// trace blocks (:654-, mat_C.hpp:29, vec_g.hpp:40), with the scatter of scatter.hpp:66-115
// (K += C_test^T A_c C_trial, local entries |a_ij| <= 1e-15 skipped) into the CSR pattern of
// K = [A_y B^T; B C] (saddle_point_system.hpp:105-118).  Per Y cell, with
// phi the Y basis and psi the X basis evaluated at physical_to_reference(x_cell, x_q):
//   A_y(i,j) = sum_q w_q |J| grad phi_i . M(x_q) grad phi_j
//   B(i,j)   = sum_q w_q |J| (d_t psi_i phi_j + grad psi_i . M(x_q) grad phi_j)   (X row i, Y column j)
//   f_i      = sum_q w_q |J| ell phi_i
// on the prism rule (Duffy rule of degree 4p+1 x Gauss with 2p+1 points; 1+1D: Gauss x
// Gauss, runner_case_execution.hpp:29-37).  The rule and both bases are tensor products
// of a spatial and a temporal part, and M depends on x only, so the q-loop is evaluated
// in sum-factorised form: per cell and per spatial point the physical gradients, M(x_q)
// and the X basis at the mapped point; spatial matrices summed over the spatial points,
// time matrices over the time points; entries as sums of their products.  Same sums,
// reordered.  Tables of the reference cell are staged in LDS, per-point data too.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/apf.h"
#include "apf_internal.hpp"
#include "apf_tables.hpp"

extern thread_local std::string g_last_error;

using apf::HipFailure;

namespace {

constexpr double kZeroTol = 1e-15;  // scatter.hpp:66-115 default zero_tol
constexpr int kGeo = 8;             // per cell: 2D x0 y0 x1 y1 x2 y2 t0 t1; 1D x0 x1 t0 t1

int fail(apf_ctx* ctx, int code, const std::string& m)
{
    g_last_error = m;
    if (ctx) ctx->error = m;
    return code;
}

template <int DIM, int P>
struct GS {
    static constexpr int NS = DIM == 1 ? P + 1 : (P + 1) * (P + 2) / 2;
    static constexpr int NT = P + 1;
    static constexpr int NL = NS * NT;
    static constexpr int NQS = DIM == 1 ? 2 * P + 1 : (2 * P + 2) * (2 * P + 2);
    static constexpr int NQT = 2 * P + 1;
    static constexpr int MD = DIM * DIM;
    static constexpr int CB = (256 / (NS * NS)) < 16 ? (256 / (NS * NS)) : 16;  // Y cells per workgroup
    static constexpr int CBB = 256 / (NS * NS) < 32 ? 256 / (NS * NS) : 32;    // bottom cells per workgroup
};

struct GenDev {
    int64_t nyc, nxc, nY, n;
    const double* ygeo;
    const int64_t* yptr;
    const int* ydof;
    const double* yw;
    const int* anc;
    const double* xgeo;
    const int64_t* xptr;
    const int* xdof;
    const double* xw;
    const int* bottom;
    int64_t nbot;
    const int64_t* rowptr;
    const int* col;
    double* val;
    double* b;
    int* err;
    // reference tables
    const double* sq_x;    // [NQS][DIM]
    const double* sq_w;    // [NQS]
    const double* lamY;    // [NQS][NS]
    const double* glamY;   // [NQS][NS][DIM]
    const double* tq_x;    // [NQT]
    const double* tq_w;    // [NQT]
    const double* ellY;    // [NQT][NT]
    // coefficients
    int grid_n;
    const double* Mgrid;   // [grid_n^DIM][DIM][DIM]
    double ell;
    const double* u0b;     // [nbot][NQS]
};

// Equidistant Lagrange basis on [0, 1]: values and d/ds (segment_nodes i/p).
template <int P>
__device__ inline void seg_lagrange(double s, double* v, double* d)
{
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        double val = 1.0, der = 0.0;
#pragma unroll
        for (int j = 0; j <= P; ++j) {
            if (j == i) continue;
            const double inv = static_cast<double>(P) / (i - j);
            const double f = (s - static_cast<double>(j) / P) * inv;
            der = der * f + val * inv;
            val *= f;
        }
        v[i] = val;
        d[i] = der;
    }
}

template <int P>
__device__ inline void bary_factor(int a, double s, double& val, double& der)
{
    val = 1.0;
    der = 0.0;
    for (int m = 0; m < a; ++m) {
        const double r = 1.0 / (m + 1.0);
        const double f = (P * s - m) * r;
        der = der * f + val * (P * r);
        val *= f;
    }
}

// P_p Lagrange basis on the reference cell at xi: values and reference gradients
// (triangle: products of barycentric factors, node order c outer / b inner).
template <int DIM, int P>
__device__ inline void spatial_lagrange(const double* xi, double* v, double* g)
{
    if constexpr (DIM == 1) {
        double d[P + 1];
        seg_lagrange<P>(xi[0], v, d);
#pragma unroll
        for (int i = 0; i <= P; ++i) g[i] = d[i];
    } else {
        const double l0 = 1.0 - xi[0] - xi[1];
        int k = 0;
#pragma unroll
        for (int c = 0; c <= P; ++c)
#pragma unroll
            for (int b = 0; b <= P - c; ++b) {
                const int a = P - b - c;
                double fa, da, fb, db, fc, dc;
                bary_factor<P>(a, l0, fa, da);
                bary_factor<P>(b, xi[0], fb, db);
                bary_factor<P>(c, xi[1], fc, dc);
                v[k] = fa * fb * fc;
                g[2 * k] = -da * fb * fc + fa * db * fc;
                g[2 * k + 1] = -da * fb * fc + fa * fb * dc;
                ++k;
            }
    }
}

// Affine map of a cell: x = v0 + J xi.  Returns |det J| and J^-1 (row-major).
template <int DIM>
__device__ inline double cell_map(const double* geo, double* v0, double* Jinv)
{
    if constexpr (DIM == 1) {
        v0[0] = geo[0];
        const double J = geo[1] - geo[0];
        Jinv[0] = 1.0 / J;
        return fabs(J);
    } else {
        v0[0] = geo[0];
        v0[1] = geo[1];
        const double a = geo[2] - geo[0], b = geo[4] - geo[0], c = geo[3] - geo[1], d = geo[5] - geo[1];
        const double det = a * d - b * c;
        Jinv[0] = d / det;
        Jinv[1] = -b / det;
        Jinv[2] = -c / det;
        Jinv[3] = a / det;
        return fabs(det);
    }
}

template <int DIM>
__device__ inline void M_at(const GenDev& d, const double* x, double* M)
{
    if (d.grid_n <= 0) {
        if constexpr (DIM == 1) {
            M[0] = 1.0;
        } else {
            M[0] = 1.0; M[1] = 0.0; M[2] = 0.0; M[3] = 1.0;
        }
        return;
    }
    const int gn = d.grid_n;
    int ix = static_cast<int>(x[0] * gn);
    ix = ix < 0 ? 0 : (ix >= gn ? gn - 1 : ix);
    int cell = ix;
    if constexpr (DIM == 2) {
        int iy = static_cast<int>(x[1] * gn);
        iy = iy < 0 ? 0 : (iy >= gn ? gn - 1 : iy);
        cell = iy * gn + ix;
    }
#pragma unroll
    for (int e = 0; e < DIM * DIM; ++e) M[e] = d.Mgrid[cell * DIM * DIM + e];
}

// slot of column c in CSR row r (columns sorted); -1 if the pattern lacks it
__device__ inline int64_t find_slot(const int64_t* __restrict__ rowptr, const int* __restrict__ col, int64_t r,
                                    int c)
{
    int64_t lo = rowptr[r], hi = rowptr[r + 1] - 1;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    return (lo <= hi && col[lo] == c) ? lo : -1;
}

__device__ inline void add_entry(const GenDev& d, int64_t r, int c, double v)
{
    const int64_t s = find_slot(d.rowptr, d.col, r, c);
    if (s < 0) {
        atomicAdd(d.err, 1);
        return;
    }
    atomicAdd(&d.val[s], v);
}

template <int DIM, int P>
__global__ void __launch_bounds__(256) k_gen_cells(GenDev d)
{
    using S = GS<DIM, P>;
    constexpr int NS = S::NS, NT = S::NT, NL = S::NL, NQS = S::NQS, NQT = S::NQT, MD = S::MD, CB = S::CB;
    // reference tables
    __shared__ double t_lam[NQS * NS], t_glam[NQS * NS * DIM], t_w[NQS], t_x[NQS * DIM];
    __shared__ double t_tw[NQT], t_tx[NQT], t_ell[NQT * NT];
    // per cell: Y map (v0, J^-1, |det|, t0, h), X map (v0, J^-1, t0, h)
    __shared__ double g_v0[CB][DIM], g_Ji[CB][MD], g_det[CB], g_t0[CB], g_h[CB];
    __shared__ double gx_v0[CB][DIM], gx_Ji[CB][MD], gx_t0[CB], gx_h[CB];
    // per (cell, spatial point)
    __shared__ double q_M[CB * NQS * MD], q_mu[CB * NQS], q_lamX[CB * NQS * NS], q_gX[CB * NQS * NS * DIM];
    // per cell: spatial and time matrices
    __shared__ double KY[CB * NS * NS], KXY[CB * NS * NS], MXY[CB * NS * NS], mY[CB * NS];
    __shared__ double MtY[CB * NT * NT], MtXY[CB * NT * NT], DtXY[CB * NT * NT], mt[CB * NT];

    const int tid = threadIdx.x;
    const int64_t cbase = static_cast<int64_t>(blockIdx.x) * CB;
    const int ncb = static_cast<int>(min(static_cast<int64_t>(CB), d.nyc - cbase));
    for (int e = tid; e < NQS * NS; e += blockDim.x) t_lam[e] = d.lamY[e];
    for (int e = tid; e < NQS * NS * DIM; e += blockDim.x) t_glam[e] = d.glamY[e];
    for (int e = tid; e < NQS; e += blockDim.x) t_w[e] = d.sq_w[e];
    for (int e = tid; e < NQS * DIM; e += blockDim.x) t_x[e] = d.sq_x[e];
    for (int e = tid; e < NQT; e += blockDim.x) {
        t_tw[e] = d.tq_w[e];
        t_tx[e] = d.tq_x[e];
    }
    for (int e = tid; e < NQT * NT; e += blockDim.x) t_ell[e] = d.ellY[e];
    if (tid < ncb) {
        const int64_t c = cbase + tid;
        const double* gy = d.ygeo + c * kGeo;
        g_det[tid] = cell_map<DIM>(gy, g_v0[tid], g_Ji[tid]);
        g_t0[tid] = gy[kGeo - 2];
        g_h[tid] = gy[kGeo - 1] - gy[kGeo - 2];
        const double* gxp = d.xgeo + static_cast<int64_t>(d.anc[c]) * kGeo;
        (void)cell_map<DIM>(gxp, gx_v0[tid], gx_Ji[tid]);
        gx_t0[tid] = gxp[kGeo - 2];
        gx_h[tid] = gxp[kGeo - 1] - gxp[kGeo - 2];
    }
    __syncthreads();
    // ---- per (cell, spatial point): x_q, M(x_q), w |det|, X basis at the mapped point
    for (int e = tid; e < ncb * NQS; e += blockDim.x) {
        const int cl = e / NQS, q = e % NQS;
        double xq[DIM], xi[DIM];
        const double* gy = d.ygeo + (cbase + cl) * kGeo;
        if constexpr (DIM == 1) {
            xq[0] = gy[0] + (gy[1] - gy[0]) * t_x[q];
            xi[0] = gx_Ji[cl][0] * (xq[0] - gx_v0[cl][0]);
        } else {
            const double a = t_x[2 * q], bb = t_x[2 * q + 1];
            xq[0] = gy[0] + (gy[2] - gy[0]) * a + (gy[4] - gy[0]) * bb;
            xq[1] = gy[1] + (gy[3] - gy[1]) * a + (gy[5] - gy[1]) * bb;
            const double dx = xq[0] - gx_v0[cl][0], dy = xq[1] - gx_v0[cl][1];
            xi[0] = gx_Ji[cl][0] * dx + gx_Ji[cl][1] * dy;
            xi[1] = gx_Ji[cl][2] * dx + gx_Ji[cl][3] * dy;
        }
        M_at<DIM>(d, xq, &q_M[e * MD]);
        q_mu[e] = t_w[q] * g_det[cl];
        double v[NS], g[NS * DIM];
        spatial_lagrange<DIM, P>(xi, v, g);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            q_lamX[e * NS + s] = v[s];
            // physical gradient J^-T grad_ref
            if constexpr (DIM == 1) {
                q_gX[e * NS + s] = gx_Ji[cl][0] * g[s];
            } else {
                q_gX[(e * NS + s) * 2] = gx_Ji[cl][0] * g[2 * s] + gx_Ji[cl][2] * g[2 * s + 1];
                q_gX[(e * NS + s) * 2 + 1] = gx_Ji[cl][1] * g[2 * s] + gx_Ji[cl][3] * g[2 * s + 1];
            }
        }
    }
    __syncthreads();
    // ---- spatial matrices (cell, s, s') over the spatial points
    for (int e = tid; e < ncb * NS * NS; e += blockDim.x) {
        const int cl = e / (NS * NS), s = (e / NS) % NS, sp = e % NS;
        const double* Ji = g_Ji[cl];
        double ky = 0.0, kxy = 0.0, mxy = 0.0, my = 0.0;
        for (int q = 0; q < NQS; ++q) {
            const int pe = cl * NQS + q;
            const double mu = q_mu[pe];
            const double* M = &q_M[pe * MD];
            if constexpr (DIM == 1) {
                const double gs = Ji[0] * t_glam[q * NS + s], gsp = Ji[0] * t_glam[q * NS + sp];
                const double mg = M[0] * gsp;
                ky += mu * gs * mg;
                kxy += mu * q_gX[pe * NS + s] * mg;
            } else {
                const double* gr = &t_glam[q * NS * 2];
                const double gs0 = Ji[0] * gr[2 * s] + Ji[2] * gr[2 * s + 1];
                const double gs1 = Ji[1] * gr[2 * s] + Ji[3] * gr[2 * s + 1];
                const double gp0 = Ji[0] * gr[2 * sp] + Ji[2] * gr[2 * sp + 1];
                const double gp1 = Ji[1] * gr[2 * sp] + Ji[3] * gr[2 * sp + 1];
                const double mg0 = M[0] * gp0 + M[1] * gp1, mg1 = M[2] * gp0 + M[3] * gp1;
                ky += mu * (gs0 * mg0 + gs1 * mg1);
                kxy += mu * (q_gX[(pe * NS + s) * 2] * mg0 + q_gX[(pe * NS + s) * 2 + 1] * mg1);
            }
            mxy += mu * q_lamX[pe * NS + s] * t_lam[q * NS + sp];
            my += mu * t_lam[q * NS + s];
        }
        KY[e] = ky;
        KXY[e] = kxy;
        MXY[e] = mxy;
        if (sp == 0) mY[cl * NS + s] = my;
    }
    // ---- time matrices (cell, t, t') over the time points
    for (int e = tid; e < ncb * NT * NT; e += blockDim.x) {
        const int cl = e / (NT * NT), t = (e / NT) % NT, tp = e % NT;
        const double h = g_h[cl], hx = gx_h[cl];
        double my = 0.0, mxy = 0.0, dxy = 0.0, m1 = 0.0;
        for (int q = 0; q < NQT; ++q) {
            const double tau = t_tx[q];
            const double tx = (g_t0[cl] + h * tau - gx_t0[cl]) / hx;
            double lx[P + 1], dlx[P + 1];
            seg_lagrange<P>(tx, lx, dlx);
            const double w = h * t_tw[q];
            const double ly = t_ell[q * NT + tp];
            my += w * t_ell[q * NT + t] * ly;
            mxy += w * lx[t] * ly;
            dxy += w * (dlx[t] / hx) * ly;
            m1 += w * t_ell[q * NT + t];
        }
        MtY[e] = my;
        MtXY[e] = mxy;
        DtXY[e] = dxy;
        if (tp == 0) mt[cl * NT + t] = m1;
    }
    __syncthreads();
    // ---- entries and scatter through the restriction rows
    for (int e = tid; e < ncb * NL * NL; e += blockDim.x) {
        const int cl = e / (NL * NL), i = (e / NL) % NL, j = e % NL;
        const int s = i / NT, t = i % NT, sp = j / NT, tp = j % NT;
        const int64_t c = cbase + cl;
        const int m0 = cl * NS * NS + s * NS + sp, t0 = cl * NT * NT + t * NT + tp;
        const int64_t yrj0 = d.yptr[c * NL + j], yrj1 = d.yptr[c * NL + j + 1];
        // A_y(i, j): Y row i, Y column j
        const double a = KY[m0] * MtY[t0];
        if (fabs(a) > kZeroTol) {
            for (int64_t ei = d.yptr[c * NL + i]; ei < d.yptr[c * NL + i + 1]; ++ei) {
                const int gi = d.ydof[ei];
                const double wi = d.yw[ei];
                for (int64_t ej = yrj0; ej < yrj1; ++ej) add_entry(d, gi, d.ydof[ej], wi * d.yw[ej] * a);
            }
        }
        // B(i, j): X row i of the ancestor, Y column j; and its transpose
        const double bv = MXY[m0] * DtXY[t0] + KXY[m0] * MtXY[t0];
        if (fabs(bv) > kZeroTol) {
            const int64_t xc = d.anc[c];
            for (int64_t ei = d.xptr[xc * NL + i]; ei < d.xptr[xc * NL + i + 1]; ++ei) {
                const int gx = d.xdof[ei];
                const double wi = d.xw[ei];
                for (int64_t ej = yrj0; ej < yrj1; ++ej) {
                    const int gy = d.ydof[ej];
                    const double v = wi * d.yw[ej] * bv;
                    add_entry(d, d.nY + gx, gy, v);
                    add_entry(d, gy, static_cast<int>(d.nY + gx), v);
                }
            }
        }
    }
    // f_i = ell int phi_i
    if (d.ell != 0.0) {
        for (int e = tid; e < ncb * NL; e += blockDim.x) {
            const int cl = e / NL, i = e % NL;
            const int64_t c = cbase + cl;
            const double fv = d.ell * mY[cl * NS + i / NT] * mt[cl * NT + i % NT];
            if (fabs(fv) <= kZeroTol) continue;
            for (int64_t ei = d.yptr[c * NL + i]; ei < d.yptr[c * NL + i + 1]; ++ei)
                atomicAdd(&d.b[d.ydof[ei]], d.yw[ei] * fv);
        }
    }
}

// C_signed = -int_{t=0} psi_i psi_j and g = -int_{t=0} u0 psi_i on the X cells with t0 = 0
// (local_trace_bilinear_forms.hpp:20-49, local_trace_linear_forms.hpp:22-52, g_scale = -1
// at solve_main_system.hpp:1833): only the local dofs at the bottom time node (i_t = 0).
template <int DIM, int P>
__global__ void __launch_bounds__(256) k_gen_bottom(GenDev d)
{
    using S = GS<DIM, P>;
    constexpr int NS = S::NS, NT = S::NT, NL = S::NL, NQS = S::NQS, MD = S::MD, CB = S::CBB;
    __shared__ double t_lam[NQS * NS], t_w[NQS];
    __shared__ double g_det[CB];
    const int tid = threadIdx.x;
    const int64_t bbase = static_cast<int64_t>(blockIdx.x) * CB;
    const int ncb = static_cast<int>(min(static_cast<int64_t>(CB), d.nbot - bbase));
    for (int e = tid; e < NQS * NS; e += blockDim.x) t_lam[e] = d.lamY[e];
    for (int e = tid; e < NQS; e += blockDim.x) t_w[e] = d.sq_w[e];
    if (tid < ncb) {
        double v0[DIM], Ji[MD];
        g_det[tid] = cell_map<DIM>(d.xgeo + static_cast<int64_t>(d.bottom[bbase + tid]) * kGeo, v0, Ji);
    }
    __syncthreads();
    for (int e = tid; e < ncb * NS * NS; e += blockDim.x) {
        const int cl = e / (NS * NS), s = (e / NS) % NS, sp = e % NS;
        const int64_t bc = bbase + cl;
        const int64_t xc = d.bottom[bc];
        double cv = 0.0, gv = 0.0;
        for (int q = 0; q < NQS; ++q) {
            const double mu = t_w[q] * g_det[cl];
            cv += mu * t_lam[q * NS + s] * t_lam[q * NS + sp];
            if (sp == 0) gv += mu * d.u0b[bc * NQS + q] * t_lam[q * NS + s];
        }
        cv = -cv;
        const int i = s * NT, j = sp * NT;  // bottom time node
        if (fabs(cv) > kZeroTol) {
            for (int64_t ei = d.xptr[xc * NL + i]; ei < d.xptr[xc * NL + i + 1]; ++ei)
                for (int64_t ej = d.xptr[xc * NL + j]; ej < d.xptr[xc * NL + j + 1]; ++ej)
                    add_entry(d, d.nY + d.xdof[ei], static_cast<int>(d.nY + d.xdof[ej]), d.xw[ei] * d.xw[ej] * cv);
        }
        if (sp == 0) {
            gv = -gv;
            if (fabs(gv) > kZeroTol)
                for (int64_t ei = d.xptr[xc * NL + i]; ei < d.xptr[xc * NL + i + 1]; ++ei)
                    atomicAdd(&d.b[d.nY + d.xdof[ei]], d.xw[ei] * gv);
        }
    }
}

// X-marking indicator parts (algorithm/detail/indicator_utilities.hpp:435-473):
// lambda_Y^2 of each Y cell, int_cell M grad_x lambda . grad_x lambda on the cell rule
// (compute_spatial_energy_squared_by_active_cell, :109-171), one thread per Y cell.
template <int DIM, int P>
__global__ void __launch_bounds__(256) k_gen_lambda_energy(GenDev d, const double* __restrict__ x,
                                                           double* __restrict__ out)
{
    using S = GS<DIM, P>;
    constexpr int NS = S::NS, NT = S::NT, NL = S::NL, NQS = S::NQS, NQT = S::NQT, MD = S::MD;
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= d.nyc) return;
    double lam[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        double v = 0.0;
        for (int64_t e = d.yptr[c * NL + l]; e < d.yptr[c * NL + l + 1]; ++e) v += d.yw[e] * x[d.ydof[e]];
        lam[l] = v;
    }
    const double* gy = d.ygeo + c * kGeo;
    double v0[DIM], Ji[MD];
    const double det = cell_map<DIM>(gy, v0, Ji);
    const double h = gy[kGeo - 1] - gy[kGeo - 2];
    double acc = 0.0;
    for (int q = 0; q < NQS; ++q) {
        double xq[DIM];
        if constexpr (DIM == 1) {
            xq[0] = gy[0] + (gy[1] - gy[0]) * d.sq_x[q];
        } else {
            const double a = d.sq_x[2 * q], bb = d.sq_x[2 * q + 1];
            xq[0] = gy[0] + (gy[2] - gy[0]) * a + (gy[4] - gy[0]) * bb;
            xq[1] = gy[1] + (gy[3] - gy[1]) * a + (gy[5] - gy[1]) * bb;
        }
        double M[MD];
        M_at<DIM>(d, xq, M);
        double gp[NS][DIM];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const double* gr = d.glamY + (q * NS + s) * DIM;
            if constexpr (DIM == 1) {
                gp[s][0] = Ji[0] * gr[0];
            } else {
                gp[s][0] = Ji[0] * gr[0] + Ji[2] * gr[1];
                gp[s][1] = Ji[1] * gr[0] + Ji[3] * gr[1];
            }
        }
        for (int qt = 0; qt < NQT; ++qt) {
            double g[DIM];
#pragma unroll
            for (int a = 0; a < DIM; ++a) g[a] = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                double ls = 0.0;
#pragma unroll
                for (int j = 0; j < NT; ++j) ls += lam[s * NT + j] * d.ellY[qt * NT + j];
#pragma unroll
                for (int a = 0; a < DIM; ++a) g[a] += ls * gp[s][a];
            }
            double e = 0.0;
#pragma unroll
            for (int a = 0; a < DIM; ++a) {
                double mg = 0.0;
#pragma unroll
                for (int b = 0; b < DIM; ++b) mg += M[a * DIM + b] * g[b];
                e += g[a] * mg;
            }
            acc += d.sq_w[q] * det * d.tq_w[qt] * h * e;
        }
    }
    out[c] = acc;
}

// initial-trace mismatch int_{t=0} (u0 - u_delta)^2 of each X cell with t0 = 0
// (compute_initial_trace_mismatch_squared_by_x_cell, indicator_utilities.hpp:333-396), one
// thread per such cell; out[] (one entry per X cell) is zero elsewhere.
template <int DIM, int P>
__global__ void __launch_bounds__(256) k_gen_trace(GenDev d, const double* __restrict__ x, double* __restrict__ out)
{
    using S = GS<DIM, P>;
    constexpr int NS = S::NS, NT = S::NT, NL = S::NL, NQS = S::NQS, MD = S::MD;
    const int64_t bc = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (bc >= d.nbot) return;
    const int64_t xc = d.bottom[bc];
    double us[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const int l = s * NT;  // bottom time node
        double v = 0.0;
        for (int64_t e = d.xptr[xc * NL + l]; e < d.xptr[xc * NL + l + 1]; ++e) v += d.xw[e] * x[d.nY + d.xdof[e]];
        us[s] = v;
    }
    double v0[DIM], Ji[MD];
    const double det = cell_map<DIM>(d.xgeo + xc * kGeo, v0, Ji);
    double acc = 0.0;
    for (int q = 0; q < NQS; ++q) {
        double u = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) u += us[s] * d.lamY[q * NS + s];
        const double diff = d.u0b[bc * NQS + q] - u;
        acc += diff * diff * det * d.sq_w[q];
    }
    out[xc] = acc;
}

// Parabolic graph-norm parts on each Y cell (assemble_parabolic_graph_norm_approximation,
// assembly/main_system/parabolic_graph_norm.hpp:22-65): B_dt(i, j) = int dpsi_i/dt phi_j
// (assemble_mat_B_dt, mat_B.hpp:250-273: the time-derivative part of B, with the X basis
// of the cell's ancestor at the mapped points) into BT = B_dt^T (rows Y dofs), and
// A_x(i, k) = int M grad_x psi_i . grad_x psi_k (assemble_mat_A on the X space, mat_A.hpp)
// into Hhat, integrated over the Y cells that tile each X cell.  Local values with
// |v| <= 1e-15 are skipped (scatter.hpp:92).
struct GraphDev {
    const int64_t* bt_rowptr;
    const int* bt_col;
    double* bt_val;
    const int64_t* h_rowptr;
    const int* h_col;
    double* h_val;
};

__device__ inline void add_csr(const int64_t* rowptr, const int* col, double* val, int* err, int64_t r, int c,
                               double v)
{
    const int64_t s = find_slot(rowptr, col, r, c);
    if (s < 0) {
        atomicAdd(err, 1);
        return;
    }
    atomicAdd(&val[s], v);
}

template <int DIM, int P>
__global__ void __launch_bounds__(256) k_gen_graph(GenDev d, GraphDev g)
{
    using S = GS<DIM, P>;
    constexpr int NS = S::NS, NT = S::NT, NL = S::NL, NQS = S::NQS, NQT = S::NQT, MD = S::MD, CB = S::CB;
    __shared__ double t_lam[NQS * NS], t_w[NQS], t_x[NQS * DIM];
    __shared__ double t_tw[NQT], t_tx[NQT], t_ell[NQT * NT];
    __shared__ double g_det[CB], g_t0[CB], g_h[CB];
    __shared__ double gx_v0[CB][DIM], gx_Ji[CB][MD], gx_t0[CB], gx_h[CB];
    __shared__ double q_M[CB * NQS * MD], q_mu[CB * NQS], q_lamX[CB * NQS * NS], q_gX[CB * NQS * NS * DIM];
    __shared__ double MXY[CB * NS * NS], KXX[CB * NS * NS], DtXY[CB * NT * NT], MtXX[CB * NT * NT];

    const int tid = threadIdx.x;
    const int64_t cbase = static_cast<int64_t>(blockIdx.x) * CB;
    const int ncb = static_cast<int>(min(static_cast<int64_t>(CB), d.nyc - cbase));
    for (int e = tid; e < NQS * NS; e += blockDim.x) t_lam[e] = d.lamY[e];
    for (int e = tid; e < NQS; e += blockDim.x) t_w[e] = d.sq_w[e];
    for (int e = tid; e < NQS * DIM; e += blockDim.x) t_x[e] = d.sq_x[e];
    for (int e = tid; e < NQT; e += blockDim.x) {
        t_tw[e] = d.tq_w[e];
        t_tx[e] = d.tq_x[e];
    }
    for (int e = tid; e < NQT * NT; e += blockDim.x) t_ell[e] = d.ellY[e];
    if (tid < ncb) {
        const int64_t c = cbase + tid;
        const double* gy = d.ygeo + c * kGeo;
        double v0[DIM], Ji[MD];
        g_det[tid] = cell_map<DIM>(gy, v0, Ji);
        g_t0[tid] = gy[kGeo - 2];
        g_h[tid] = gy[kGeo - 1] - gy[kGeo - 2];
        const double* gxp = d.xgeo + static_cast<int64_t>(d.anc[c]) * kGeo;
        (void)cell_map<DIM>(gxp, gx_v0[tid], gx_Ji[tid]);
        gx_t0[tid] = gxp[kGeo - 2];
        gx_h[tid] = gxp[kGeo - 1] - gxp[kGeo - 2];
    }
    __syncthreads();
    for (int e = tid; e < ncb * NQS; e += blockDim.x) {
        const int cl = e / NQS, q = e % NQS;
        double xq[DIM], xi[DIM];
        const double* gy = d.ygeo + (cbase + cl) * kGeo;
        if constexpr (DIM == 1) {
            xq[0] = gy[0] + (gy[1] - gy[0]) * t_x[q];
            xi[0] = gx_Ji[cl][0] * (xq[0] - gx_v0[cl][0]);
        } else {
            const double a = t_x[2 * q], bb = t_x[2 * q + 1];
            xq[0] = gy[0] + (gy[2] - gy[0]) * a + (gy[4] - gy[0]) * bb;
            xq[1] = gy[1] + (gy[3] - gy[1]) * a + (gy[5] - gy[1]) * bb;
            const double dx = xq[0] - gx_v0[cl][0], dy = xq[1] - gx_v0[cl][1];
            xi[0] = gx_Ji[cl][0] * dx + gx_Ji[cl][1] * dy;
            xi[1] = gx_Ji[cl][2] * dx + gx_Ji[cl][3] * dy;
        }
        M_at<DIM>(d, xq, &q_M[e * MD]);
        q_mu[e] = t_w[q] * g_det[cl];
        double v[NS], gr[NS * DIM];
        spatial_lagrange<DIM, P>(xi, v, gr);
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            q_lamX[e * NS + s] = v[s];
            if constexpr (DIM == 1) {
                q_gX[e * NS + s] = gx_Ji[cl][0] * gr[s];
            } else {
                q_gX[(e * NS + s) * 2] = gx_Ji[cl][0] * gr[2 * s] + gx_Ji[cl][2] * gr[2 * s + 1];
                q_gX[(e * NS + s) * 2 + 1] = gx_Ji[cl][1] * gr[2 * s] + gx_Ji[cl][3] * gr[2 * s + 1];
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < ncb * NS * NS; e += blockDim.x) {
        const int cl = e / (NS * NS), s = (e / NS) % NS, sp = e % NS;
        double kxx = 0.0, mxy = 0.0;
        for (int q = 0; q < NQS; ++q) {
            const int pe = cl * NQS + q;
            const double mu = q_mu[pe];
            const double* M = &q_M[pe * MD];
            if constexpr (DIM == 1) {
                kxx += mu * q_gX[pe * NS + s] * M[0] * q_gX[pe * NS + sp];
            } else {
                const double* ga = &q_gX[(pe * NS + s) * 2];
                const double* gb = &q_gX[(pe * NS + sp) * 2];
                kxx += mu * (ga[0] * (M[0] * gb[0] + M[1] * gb[1]) + ga[1] * (M[2] * gb[0] + M[3] * gb[1]));
            }
            mxy += mu * q_lamX[pe * NS + s] * t_lam[q * NS + sp];
        }
        KXX[e] = kxx;
        MXY[e] = mxy;
    }
    for (int e = tid; e < ncb * NT * NT; e += blockDim.x) {
        const int cl = e / (NT * NT), t = (e / NT) % NT, tp = e % NT;
        const double h = g_h[cl], hx = gx_h[cl];
        double dxy = 0.0, mxx = 0.0;
        for (int q = 0; q < NQT; ++q) {
            const double tx = (g_t0[cl] + h * t_tx[q] - gx_t0[cl]) / hx;
            double lx[P + 1], dlx[P + 1];
            seg_lagrange<P>(tx, lx, dlx);
            const double w = h * t_tw[q];
            dxy += w * (dlx[t] / hx) * t_ell[q * NT + tp];
            mxx += w * lx[t] * lx[tp];
        }
        DtXY[e] = dxy;
        MtXX[e] = mxx;
    }
    __syncthreads();
    for (int e = tid; e < ncb * NL * NL; e += blockDim.x) {
        const int cl = e / (NL * NL), i = (e / NL) % NL, j = e % NL;
        const int s = i / NT, t = i % NT, sp = j / NT, tp = j % NT;
        const int64_t c = cbase + cl, xc = d.anc[c];
        const int m0 = cl * NS * NS + s * NS + sp, t0 = cl * NT * NT + t * NT + tp;
        // B_dt(X local i of the ancestor, Y local j) into BT(Y, X)
        const double bv = MXY[m0] * DtXY[t0];
        if (fabs(bv) > kZeroTol) {
            for (int64_t ei = d.xptr[xc * NL + i]; ei < d.xptr[xc * NL + i + 1]; ++ei)
                for (int64_t ej = d.yptr[c * NL + j]; ej < d.yptr[c * NL + j + 1]; ++ej)
                    add_csr(g.bt_rowptr, g.bt_col, g.bt_val, d.err, d.ydof[ej], d.xdof[ei], d.xw[ei] * d.yw[ej] * bv);
        }
        // A_x(X local i, X local j) over this Y cell
        const double av = KXX[m0] * MtXX[t0];
        if (fabs(av) > kZeroTol) {
            for (int64_t ei = d.xptr[xc * NL + i]; ei < d.xptr[xc * NL + i + 1]; ++ei)
                for (int64_t ej = d.xptr[xc * NL + j]; ej < d.xptr[xc * NL + j + 1]; ++ej)
                    add_csr(g.h_rowptr, g.h_col, g.h_val, d.err, d.xdof[ei], d.xdof[ej], d.xw[ei] * d.xw[ej] * av);
        }
    }
}

// Hhat += -C_signed (K's X-X block) and, per Y dof j, the outer product of column j of B_dt
// scaled by 1 / A_y(j, j) (diagonal_schur_approximation_with_diagnostics,
// linear_algebra/preconditioners/schur_approximations.hpp:80-212).  A_y(j, j) <= 1e-14
// sets the error flag (the reference throws "A_y diagonal must be positive").
__global__ void k_hhat_schur(GenDev d, GraphDev g)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= d.nY) return;
    const int64_t s = find_slot(d.rowptr, d.col, j, static_cast<int>(j));
    const double diag = s >= 0 ? d.val[s] : 0.0;
    if (!(diag > 1.0e-14)) {
        atomicAdd(d.err + 1, 1);
        return;
    }
    const double inv = 1.0 / diag;
    for (int64_t a = g.bt_rowptr[j]; a < g.bt_rowptr[j + 1]; ++a) {
        const double va = g.bt_val[a];
        if (va == 0.0) continue;
        for (int64_t b = g.bt_rowptr[j]; b < g.bt_rowptr[j + 1]; ++b) {
            const double v = inv * va * g.bt_val[b];
            if (v != 0.0) add_csr(g.h_rowptr, g.h_col, g.h_val, d.err, g.bt_col[a], g.bt_col[b], v);
        }
    }
}

__global__ void k_hhat_trace(GenDev d, GraphDev g, int64_t nX)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= nX) return;
    const int64_t r = d.nY + i;
    for (int64_t e = d.rowptr[r]; e < d.rowptr[r + 1]; ++e) {
        const int c = d.col[e];
        if (c < d.nY || d.val[e] == 0.0) continue;
        add_csr(g.h_rowptr, g.h_col, g.h_val, d.err, i, static_cast<int>(c - d.nY), -d.val[e]);
    }
}

__global__ void k_zero(int64_t n, double* __restrict__ a)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        a[i] = 0.0;
}

__global__ void k_zero_int(int* a) { *a = 0; }

// ------------------------------------------------------------------ host side
struct CellSpaceHost {
    int dim, p;
    int64_t nc, ndofs;
    std::vector<double> geo;     // [nc][kGeo]
    std::vector<int64_t> ptr;
    std::vector<int> dof;
    std::vector<double> w;
};

int n_local(int dim, int p) { return (dim == 1 ? p + 1 : (p + 1) * (p + 2) / 2) * (p + 1); }

void read_space(const apf_cell_space_desc* D, const char* name, CellSpaceHost& S)
{
    const std::string nm(name);
    if (!D) throw std::invalid_argument(nm + ": null descriptor");
    if (D->dim != 1 && D->dim != 2) throw std::invalid_argument(nm + ": dim must be 1 or 2");
    if (D->p < 1 || D->p > 3) throw std::invalid_argument(nm + ": p must be 1..3 on non-tensor meshes");
    if (D->n_cells < 1 || D->n_vertices < 1 || !D->vertices || !D->cells || !D->times || !D->restr_ptr)
        throw std::invalid_argument(nm + ": empty mesh or null arrays");
    if (D->n_dofs < 1 || D->n_dofs > INT32_MAX) throw std::invalid_argument(nm + ": bad n_dofs");
    S.dim = D->dim;
    S.p = D->p;
    S.nc = D->n_cells;
    S.ndofs = D->n_dofs;
    const int nv = D->dim + 1, nl = n_local(D->dim, D->p);
    S.geo.assign(static_cast<size_t>(S.nc) * kGeo, 0.0);
    for (int64_t c = 0; c < S.nc; ++c) {
        for (int v = 0; v < nv; ++v) {
            const int vid = D->cells[c * nv + v];
            if (vid < 0 || vid >= D->n_vertices) throw std::invalid_argument(nm + ": vertex id out of range");
            for (int a = 0; a < D->dim; ++a) S.geo[c * kGeo + v * D->dim + a] = D->vertices[vid * D->dim + a];
        }
        const double t0 = D->times[2 * c], t1 = D->times[2 * c + 1];
        if (!(t1 > t0)) throw std::invalid_argument(nm + ": cell with t1 <= t0");
        S.geo[c * kGeo + kGeo - 2] = t0;
        S.geo[c * kGeo + kGeo - 1] = t1;
        double det;
        const double* g = &S.geo[c * kGeo];
        if (D->dim == 1) det = g[1] - g[0];
        else det = (g[2] - g[0]) * (g[5] - g[1]) - (g[4] - g[0]) * (g[3] - g[1]);
        if (!(std::fabs(det) > 0.0)) throw std::invalid_argument(nm + ": degenerate cell");
    }
    const int64_t nrows = S.nc * nl;
    S.ptr.assign(D->restr_ptr, D->restr_ptr + nrows + 1);
    if (S.ptr[0] != 0) throw std::invalid_argument(nm + ": restriction rows must start at 0");
    for (int64_t r = 0; r < nrows; ++r)
        if (S.ptr[r + 1] < S.ptr[r]) throw std::invalid_argument(nm + ": restriction row pointer decreases");
    const int64_t ne = S.ptr[nrows];
    if (ne > 0 && (!D->restr_dof || !D->restr_weight)) throw std::invalid_argument(nm + ": null restriction arrays");
    S.dof.assign(D->restr_dof, D->restr_dof + ne);
    S.w.assign(D->restr_weight, D->restr_weight + ne);
    for (int64_t r = 0; r < nrows; ++r)
        for (int64_t e = S.ptr[r]; e < S.ptr[r + 1]; ++e) {
            if (S.dof[e] < 0 || S.dof[e] >= S.ndofs) throw std::invalid_argument(nm + ": true dof out of range");
            if (e > S.ptr[r] && S.dof[e] <= S.dof[e - 1])
                throw std::invalid_argument(nm + ": restriction row not strictly ordered by true dof");
            if (!std::isfinite(S.w[e])) throw std::invalid_argument(nm + ": non-finite restriction weight");
        }
}

// sorted unique true dofs of a cell's rows
void cell_dofs(const CellSpaceHost& S, int64_t c, std::vector<int>& out)
{
    const int nl = n_local(S.dim, S.p);
    out.assign(S.dof.begin() + S.ptr[c * nl], S.dof.begin() + S.ptr[(c + 1) * nl]);
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
}

// Symbolic CSR pattern of K = [A_y B^T; B C]: the true dofs that meet in one Y cell (A_y),
// one Y cell and its X ancestor (B, B^T) and one bottom X cell (C); rows sorted.
void build_pattern(const CellSpaceHost& X, const CellSpaceHost& Y, const std::vector<int>& anc,
                   const std::vector<int>& bottom, std::vector<int64_t>& rowptr, std::vector<int>& col)
{
    const int64_t nY = Y.ndofs, n = Y.ndofs + X.ndofs;
    std::vector<int64_t> cnt(n + 1, 0);
    std::vector<int> sy, sx;
    // pass 1: upper bounds
    for (int64_t c = 0; c < Y.nc; ++c) {
        cell_dofs(Y, c, sy);
        cell_dofs(X, anc[c], sx);
        for (int g : sy) cnt[g + 1] += static_cast<int64_t>(sy.size() + sx.size());
        for (int g : sx) cnt[nY + g + 1] += static_cast<int64_t>(sy.size());
    }
    for (int b : bottom) {
        cell_dofs(X, b, sx);
        for (int g : sx) cnt[nY + g + 1] += static_cast<int64_t>(sx.size());
    }
    for (int64_t r = 0; r < n; ++r) cnt[r + 1] += cnt[r];
    std::vector<int> buf(cnt[n]);
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t c = 0; c < Y.nc; ++c) {
        cell_dofs(Y, c, sy);
        cell_dofs(X, anc[c], sx);
        for (int g : sy) {
            for (int h : sy) buf[pos[g]++] = h;
            for (int h : sx) buf[pos[g]++] = static_cast<int>(nY + h);
        }
        for (int g : sx)
            for (int h : sy) buf[pos[nY + g]++] = h;
    }
    for (int b : bottom) {
        cell_dofs(X, b, sx);
        for (int g : sx)
            for (int h : sx) buf[pos[nY + g]++] = static_cast<int>(nY + h);
    }
    rowptr.assign(n + 1, 0);
    col.clear();
    col.reserve(buf.size() / 2);
    for (int64_t r = 0; r < n; ++r) {
        auto b0 = buf.begin() + cnt[r], b1 = buf.begin() + pos[r];
        std::sort(b0, b1);
        auto e = std::unique(b0, b1);
        col.insert(col.end(), b0, e);
        rowptr[r + 1] = static_cast<int64_t>(col.size());
    }
}

struct Prepared {
    CellSpaceHost X, Y;
    std::vector<int> anc, bottom;
};

void prepare(const apf_cell_space_desc* Xd, const apf_cell_space_desc* Yd, const int* y_ancestor, Prepared& P)
{
    read_space(Xd, "X space", P.X);
    read_space(Yd, "Y space", P.Y);
    if (P.X.dim != P.Y.dim || P.X.p != P.Y.p) throw std::invalid_argument("X and Y differ in dim or p");
    if (!y_ancestor) throw std::invalid_argument("null y_ancestor");
    if (P.X.ndofs + P.Y.ndofs > INT32_MAX) throw std::invalid_argument("more than 2^31 unknowns");
    P.anc.assign(y_ancestor, y_ancestor + P.Y.nc);
    for (int64_t c = 0; c < P.Y.nc; ++c) {
        const int a = P.anc[c];
        if (a < 0 || a >= P.X.nc) throw std::invalid_argument("X ancestor out of range");
        const double* gy = &P.Y.geo[c * kGeo];
        const double* gx = &P.X.geo[static_cast<int64_t>(a) * kGeo];
        if (gy[kGeo - 2] < gx[kGeo - 2] - 1e-14 || gy[kGeo - 1] > gx[kGeo - 1] + 1e-14)
            throw std::invalid_argument("Y cell not inside its X ancestor in time");
    }
    for (int64_t c = 0; c < P.X.nc; ++c)
        if (P.X.geo[c * kGeo + kGeo - 2] == 0.0) P.bottom.push_back(static_cast<int>(c));
}

// Patterns of B_dt^T (Y rows) and Hhat (X rows).  Row j of B_dt^T holds the X dofs of the
// ancestors of the Y cells that contain Y dof j; Hhat's pattern is that of B_dt B_dt^T (it
// contains A_x's and C's: every X cell holds Y cells, whose dofs couple to all its X
// dofs), row a the union of the B_dt^T rows of the Y dofs in row a of B_dt.  Both are built
// with marker arrays, rows split over threads.
template <typename F>
void parallel_rows(int64_t n, F f)
{
    const int nt = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(16, n / 1024)));
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(f, t, n * t / nt, n * (t + 1) / nt);
    f(0, 0, n / nt);
    for (auto& q : th) q.join();
}

void union_rows(int64_t nrows, int64_t ncols, const std::function<void(int64_t, std::vector<int>&, std::vector<int>&, int)>& gather,
                std::vector<int64_t>& rp, std::vector<int>& col)
{
    const int nt = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(16, nrows / 1024)));
    std::vector<std::vector<int>> cols(nt);
    std::vector<std::vector<int64_t>> lens(nt);
    parallel_rows(nrows, [&](int t, int64_t r0, int64_t r1) {
        std::vector<int> mark(static_cast<size_t>(ncols), -1);
        auto& out = cols[t];
        auto& len = lens[t];
        for (int64_t r = r0; r < r1; ++r) {
            const size_t start = out.size();
            gather(r, mark, out, static_cast<int>(r));
            std::sort(out.begin() + start, out.end());
            len.push_back(static_cast<int64_t>(out.size() - start));
        }
    });
    rp.assign(1, 0);
    col.clear();
    size_t total = 0;
    for (const auto& c : cols) total += c.size();
    col.reserve(total);
    for (int t = 0; t < nt; ++t) {
        for (int64_t l : lens[t]) rp.push_back(rp.back() + l);
        col.insert(col.end(), cols[t].begin(), cols[t].end());
    }
}

void graph_patterns(const Prepared& P, std::vector<int64_t>& bt_rp, std::vector<int>& bt_col,
                    std::vector<int64_t>& h_rp, std::vector<int>& h_col)
{
    const int64_t nY = P.Y.ndofs, nX = P.X.ndofs, nyc = P.Y.nc, nxc = P.X.nc;
    // Y dof -> Y cells, and the dof lists of every cell
    std::vector<std::vector<int>> ycd(nyc), xcd(nxc);
    for (int64_t c = 0; c < nyc; ++c) cell_dofs(P.Y, c, ycd[c]);
    for (int64_t a = 0; a < nxc; ++a) cell_dofs(P.X, a, xcd[a]);
    std::vector<int64_t> yc_ptr(nY + 1, 0);
    for (int64_t c = 0; c < nyc; ++c)
        for (int j : ycd[c]) ++yc_ptr[j + 1];
    for (int64_t j = 0; j < nY; ++j) yc_ptr[j + 1] += yc_ptr[j];
    std::vector<int> yc_cell(yc_ptr.back());
    {
        std::vector<int64_t> pos(yc_ptr.begin(), yc_ptr.end() - 1);
        for (int64_t c = 0; c < nyc; ++c)
            for (int j : ycd[c]) yc_cell[pos[j]++] = static_cast<int>(c);
    }
    union_rows(nY, nX, [&](int64_t j, std::vector<int>& mark, std::vector<int>& out, int tag) {
        for (int64_t e = yc_ptr[j]; e < yc_ptr[j + 1]; ++e)
            for (int g : xcd[P.anc[yc_cell[e]]])
                if (mark[g] != tag) {
                    mark[g] = tag;
                    out.push_back(g);
                }
    }, bt_rp, bt_col);
    // B_dt rows (X dof -> Y dofs): the transpose of B_dt^T's pattern
    std::vector<int64_t> b_rp(nX + 1, 0);
    for (int g : bt_col) ++b_rp[g + 1];
    for (int64_t a = 0; a < nX; ++a) b_rp[a + 1] += b_rp[a];
    std::vector<int> b_col(bt_col.size());
    {
        std::vector<int64_t> pos(b_rp.begin(), b_rp.end() - 1);
        for (int64_t j = 0; j < nY; ++j)
            for (int64_t e = bt_rp[j]; e < bt_rp[j + 1]; ++e) b_col[pos[bt_col[e]]++] = static_cast<int>(j);
    }
    union_rows(nX, nX, [&](int64_t a, std::vector<int>& mark, std::vector<int>& out, int tag) {
        for (int64_t e = b_rp[a]; e < b_rp[a + 1]; ++e) {
            const int j = b_col[e];
            for (int64_t f = bt_rp[j]; f < bt_rp[j + 1]; ++f) {
                const int g = bt_col[f];
                if (mark[g] != tag) {
                    mark[g] = tag;
                    out.push_back(g);
                }
            }
        }
    }, h_rp, h_col);
}
}  // namespace

struct apf_general_system {
    apf_ctx* ctx = nullptr;
    Prepared host;                   // cells and restriction rows (the estimator builds its maps from them)
    int dim = 0, p = 0;
    int64_t nX = 0, nY = 0, n = 0, nnz = 0, nyc = 0, nxc = 0, nbot = 0;
    int nqs = 0;
    std::vector<void*> allocs;
    std::vector<double> bottom_pts;  // [nbot][nqs][dim]
    GenDev d{};
    int64_t* rowptr = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    double* b = nullptr;
    double* u0b = nullptr;
    double* Mgrid = nullptr;
    int64_t Mgrid_len = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool assembled = false;

    template <typename T>
    T* alloc(size_t count)
    {
        void* p_ = nullptr;
        HIP_TRY(hipMalloc(&p_, std::max<size_t>(1, count) * sizeof(T)));
        allocs.push_back(p_);
        return static_cast<T*>(p_);
    }
    template <typename T>
    T* upload(const std::vector<T>& v)
    {
        T* p_ = alloc<T>(v.size());
        if (!v.empty()) HIP_TRY(hipMemcpy(p_, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
        return p_;
    }
};

extern "C" {

int apf_general_pattern(const apf_cell_space_desc* X, const apf_cell_space_desc* Y, const int* y_ancestor,
                        int64_t* nnz, int64_t* rowptr, int* col)
{
    try {
        Prepared P;
        prepare(X, Y, y_ancestor, P);
        std::vector<int64_t> rp;
        std::vector<int> cl;
        build_pattern(P.X, P.Y, P.anc, P.bottom, rp, cl);
        if (nnz) *nnz = static_cast<int64_t>(cl.size());
        if (rowptr) std::copy(rp.begin(), rp.end(), rowptr);
        if (col) std::copy(cl.begin(), cl.end(), col);
        return APF_OK;
    } catch (const std::exception& e) {
        return fail(nullptr, APF_BAD_ARG, std::string("apf_general_pattern: ") + e.what());
    }
}

int apf_general_system_create(apf_ctx* ctx, const apf_cell_space_desc* Xd, const apf_cell_space_desc* Yd,
                              const int* y_ancestor, apf_general_system** out)
{
    if (!ctx || !out) return fail(ctx, APF_BAD_ARG, "apf_general_system_create: null argument");
    auto sys = std::make_unique<apf_general_system>();
    try {
        Prepared P;
        prepare(Xd, Yd, y_ancestor, P);
        HIP_TRY(hipSetDevice(ctx->device));
        sys->ctx = ctx;
        const int dim = sys->dim = P.X.dim, p = sys->p = P.X.p;
        sys->nX = P.X.ndofs;
        sys->nY = P.Y.ndofs;
        sys->n = sys->nX + sys->nY;
        sys->nyc = P.Y.nc;
        sys->nxc = P.X.nc;
        sys->nbot = static_cast<int64_t>(P.bottom.size());
        sys->host = P;
        std::vector<int64_t> rp;
        std::vector<int> cl;
        build_pattern(P.X, P.Y, P.anc, P.bottom, rp, cl);
        sys->nnz = static_cast<int64_t>(cl.size());
        // reference tables: spatial rule, Y spatial basis at its points, time rule and basis
        const apf::Rule1D tr = apf::gauss_legendre01(2 * p + 1);
        std::vector<double> sq_x, sq_w;
        if (dim == 1) {
            const apf::Rule1D r = apf::gauss_legendre01(2 * p + 1);
            sq_x = r.x;
            sq_w = r.w;
        } else {
            const apf::Rule2D r = apf::duffy(4 * p + 1);
            for (size_t q = 0; q < r.w.size(); ++q) {
                sq_x.push_back(r.x[q]);
                sq_x.push_back(r.y[q]);
            }
            sq_w = r.w;
        }
        const int nqs = sys->nqs = static_cast<int>(sq_w.size());
        const int Ns = dim == 1 ? p + 1 : (p + 1) * (p + 2) / 2, Nt = p + 1;
        std::vector<double> lam(static_cast<size_t>(nqs) * Ns), glam(static_cast<size_t>(nqs) * Ns * dim);
        std::vector<double> v(Ns), gx(Ns), gy(Ns);
        for (int q = 0; q < nqs; ++q) {
            if (dim == 1) {
                apf::segment_lagrange(p, sq_x[q], v.data(), gx.data());
                for (int s = 0; s < Ns; ++s) {
                    lam[q * Ns + s] = v[s];
                    glam[q * Ns + s] = gx[s];
                }
            } else {
                apf::triangle_lagrange(p, sq_x[2 * q], sq_x[2 * q + 1], v.data(), gx.data(), gy.data());
                for (int s = 0; s < Ns; ++s) {
                    lam[q * Ns + s] = v[s];
                    glam[(q * Ns + s) * 2] = gx[s];
                    glam[(q * Ns + s) * 2 + 1] = gy[s];
                }
            }
        }
        std::vector<double> ellY(tr.w.size() * Nt), dtmp(Nt);
        for (size_t q = 0; q < tr.w.size(); ++q) apf::segment_lagrange(p, tr.x[q], &ellY[q * Nt], dtmp.data());
        // bottom rule points of the X cells with t0 = 0 (where the caller evaluates u0)
        sys->bottom_pts.assign(static_cast<size_t>(sys->nbot) * nqs * dim, 0.0);
        for (int64_t bi = 0; bi < sys->nbot; ++bi) {
            const double* g = &P.X.geo[static_cast<int64_t>(P.bottom[bi]) * kGeo];
            for (int q = 0; q < nqs; ++q) {
                double* o = &sys->bottom_pts[(bi * nqs + q) * dim];
                if (dim == 1) {
                    o[0] = g[0] + (g[1] - g[0]) * sq_x[q];
                } else {
                    const double a = sq_x[2 * q], bb = sq_x[2 * q + 1];
                    o[0] = g[0] + (g[2] - g[0]) * a + (g[4] - g[0]) * bb;
                    o[1] = g[1] + (g[3] - g[1]) * a + (g[5] - g[1]) * bb;
                }
            }
        }
        GenDev& d = sys->d;
        d.nyc = P.Y.nc;
        d.nxc = P.X.nc;
        d.nY = sys->nY;
        d.n = sys->n;
        d.ygeo = sys->upload(P.Y.geo);
        d.yptr = sys->upload(P.Y.ptr);
        d.ydof = sys->upload(P.Y.dof);
        d.yw = sys->upload(P.Y.w);
        d.anc = sys->upload(P.anc);
        d.xgeo = sys->upload(P.X.geo);
        d.xptr = sys->upload(P.X.ptr);
        d.xdof = sys->upload(P.X.dof);
        d.xw = sys->upload(P.X.w);
        d.bottom = sys->upload(P.bottom);
        d.nbot = sys->nbot;
        sys->rowptr = sys->upload(rp);
        sys->col = sys->upload(cl);
        sys->val = sys->alloc<double>(static_cast<size_t>(sys->nnz));
        sys->b = sys->alloc<double>(static_cast<size_t>(sys->n));
        sys->u0b = sys->alloc<double>(static_cast<size_t>(sys->nbot) * nqs);
        d.rowptr = sys->rowptr;
        d.col = sys->col;
        d.val = sys->val;
        d.b = sys->b;
        d.err = sys->alloc<int>(1);
        d.sq_x = sys->upload(sq_x);
        d.sq_w = sys->upload(sq_w);
        d.lamY = sys->upload(lam);
        d.glamY = sys->upload(glam);
        d.tq_x = sys->upload(tr.x);
        d.tq_w = sys->upload(tr.w);
        d.ellY = sys->upload(ellY);
        d.u0b = sys->u0b;
        HIP_TRY(hipEventCreate(&sys->ev0));
        HIP_TRY(hipEventCreate(&sys->ev1));
        *out = sys.release();
        return APF_OK;
    } catch (const HipFailure& e) {
        for (void* q : sys->allocs) (void)hipFree(q);
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_system_create: ") + e.what());
    } catch (const std::exception& e) {
        for (void* q : sys->allocs) (void)hipFree(q);
        return fail(ctx, APF_BAD_ARG, std::string("apf_general_system_create: ") + e.what());
    }
}

void apf_general_system_destroy(apf_general_system* sys)
{
    if (!sys) return;
    (void)hipSetDevice(sys->ctx->device);
    (void)hipStreamSynchronize(sys->ctx->stream);
    for (void* q : sys->allocs) (void)hipFree(q);
    if (sys->ev0) (void)hipEventDestroy(sys->ev0);
    if (sys->ev1) (void)hipEventDestroy(sys->ev1);
    delete sys;
}

int apf_general_system_info(const apf_general_system* sys, int64_t* n, int64_t* n_lambda, int64_t* nnz,
                            int64_t* n_bottom_cells, int* n_bottom_points)
{
    if (!sys) return fail(nullptr, APF_BAD_ARG, "apf_general_system_info: null system");
    if (n) *n = sys->n;
    if (n_lambda) *n_lambda = sys->nY;
    if (nnz) *nnz = sys->nnz;
    if (n_bottom_cells) *n_bottom_cells = sys->nbot;
    if (n_bottom_points) *n_bottom_points = sys->nqs;
    return APF_OK;
}

int apf_general_bottom_points(const apf_general_system* sys, double* pts)
{
    if (!sys || !pts) return fail(nullptr, APF_BAD_ARG, "apf_general_bottom_points: null argument");
    std::copy(sys->bottom_pts.begin(), sys->bottom_pts.end(), pts);
    return APF_OK;
}

int apf_general_assemble(apf_general_system* sys, const apf_general_coeff* coeff, double* ms)
{
    if (!sys || !coeff) return fail(nullptr, APF_BAD_ARG, "apf_general_assemble: null argument");
    apf_ctx* ctx = sys->ctx;
    if (coeff->grid_n < 0 || (coeff->grid_n > 0 && !coeff->M_grid))
        return fail(ctx, APF_BAD_ARG, "apf_general_assemble: bad M grid");
    if (sys->nbot > 0 && !coeff->u0_bottom) return fail(ctx, APF_BAD_ARG, "apf_general_assemble: null u0_bottom");
    if (!std::isfinite(coeff->ell)) return fail(ctx, APF_BAD_ARG, "apf_general_assemble: non-finite ell");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const int dim = sys->dim, p = sys->p;
        GenDev d = sys->d;
        d.grid_n = coeff->grid_n;
        d.ell = coeff->ell;
        if (coeff->grid_n > 0) {
            int64_t len = 1;
            for (int a = 0; a < dim; ++a) len *= coeff->grid_n;
            len *= dim * dim;
            for (int64_t e = 0; e < len; ++e)
                if (!std::isfinite(coeff->M_grid[e])) return fail(ctx, APF_BAD_ARG, "apf_general_assemble: non-finite M");
            if (len > sys->Mgrid_len) {
                sys->Mgrid = sys->alloc<double>(static_cast<size_t>(len));
                sys->Mgrid_len = len;
            }
            HIP_TRY(hipMemcpyAsync(sys->Mgrid, coeff->M_grid, len * sizeof(double), hipMemcpyHostToDevice, st));
            d.Mgrid = sys->Mgrid;
        }
        if (sys->nbot > 0)
            HIP_TRY(hipMemcpyAsync(sys->u0b, coeff->u0_bottom, sys->nbot * sys->nqs * sizeof(double),
                                   hipMemcpyHostToDevice, st));
        HIP_TRY(hipEventRecord(sys->ev0, st));
        hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, sys->nnz, sys->val);
        hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, sys->n, sys->b);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, d.err);
        auto launch = [&](auto dimc, auto pc) {
            constexpr int D = decltype(dimc)::value, PP = decltype(pc)::value;
            using S = GS<D, PP>;
            const unsigned g1 = static_cast<unsigned>((sys->nyc + S::CB - 1) / S::CB);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gen_cells<D, PP>), dim3(g1), dim3(256), 0, st, d);
            if (sys->nbot > 0) {
                const unsigned g2 = static_cast<unsigned>((sys->nbot + S::CBB - 1) / S::CBB);
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gen_bottom<D, PP>), dim3(g2), dim3(256), 0, st, d);
            }
        };
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        if (dim == 1) {
            if (p == 1) launch(I1{}, I1{});
            else if (p == 2) launch(I1{}, I2{});
            else launch(I1{}, I3{});
        } else {
            if (p == 1) launch(I2{}, I1{});
            else if (p == 2) launch(I2{}, I2{});
            else launch(I2{}, I3{});
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(sys->ev1, st));
        int err = 0;
        HIP_TRY(hipMemcpyAsync(&err, d.err, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, sys->ev0, sys->ev1));
        if (ms) *ms = t;
        if (err) return fail(ctx, APF_BAD_ARG, "apf_general_assemble: " + std::to_string(err) +
                                                    " contributions outside the pattern");
        sys->assembled = true;
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_assemble: ") + e.what());
    } catch (const std::exception& e) {
        return fail(ctx, APF_BAD_ARG, std::string("apf_general_assemble: ") + e.what());
    }
}

int apf_general_download(apf_general_system* sys, int64_t* rowptr, int* col, double* values, double* b)
{
    if (!sys) return fail(nullptr, APF_BAD_ARG, "apf_general_download: null system");
    try {
        HIP_TRY(hipSetDevice(sys->ctx->device));
        HIP_TRY(hipStreamSynchronize(sys->ctx->stream));
        if (rowptr) HIP_TRY(hipMemcpy(rowptr, sys->rowptr, (sys->n + 1) * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (col) HIP_TRY(hipMemcpy(col, sys->col, sys->nnz * sizeof(int), hipMemcpyDeviceToHost));
        if (values) HIP_TRY(hipMemcpy(values, sys->val, sys->nnz * sizeof(double), hipMemcpyDeviceToHost));
        if (b) HIP_TRY(hipMemcpy(b, sys->b, sys->n * sizeof(double), hipMemcpyDeviceToHost));
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(sys->ctx, APF_HIP_ERROR, std::string("apf_general_download: ") + e.what());
    }
}

int apf_general_x_indicators(apf_general_system* sys, const double* x, const apf_general_coeff* coeff,
                             double* lambda_y2, double* trace2)
{
    if (!sys || !x || !coeff || !lambda_y2 || !trace2)
        return fail(sys ? sys->ctx : nullptr, APF_BAD_ARG, "apf_general_x_indicators: null argument");
    apf_ctx* ctx = sys->ctx;
    if (coeff->grid_n < 0 || (coeff->grid_n > 0 && !coeff->M_grid))
        return fail(ctx, APF_BAD_ARG, "apf_general_x_indicators: bad M grid");
    if (sys->nbot > 0 && !coeff->u0_bottom) return fail(ctx, APF_BAD_ARG, "apf_general_x_indicators: null u0_bottom");
    void* dx = nullptr;
    void* dout = nullptr;
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const int dim = sys->dim, p = sys->p;
        GenDev d = sys->d;
        d.grid_n = coeff->grid_n;
        if (coeff->grid_n > 0) {
            int64_t len = 1;
            for (int a = 0; a < dim; ++a) len *= coeff->grid_n;
            len *= dim * dim;
            if (len > sys->Mgrid_len) {
                sys->Mgrid = sys->alloc<double>(static_cast<size_t>(len));
                sys->Mgrid_len = len;
            }
            HIP_TRY(hipMemcpyAsync(sys->Mgrid, coeff->M_grid, len * sizeof(double), hipMemcpyHostToDevice, st));
            d.Mgrid = sys->Mgrid;
        }
        if (sys->nbot > 0)
            HIP_TRY(hipMemcpyAsync(sys->u0b, coeff->u0_bottom, sys->nbot * sys->nqs * sizeof(double),
                                   hipMemcpyHostToDevice, st));
        HIP_TRY(hipMalloc(&dx, std::max<int64_t>(1, sys->n) * sizeof(double)));
        HIP_TRY(hipMalloc(&dout, std::max<int64_t>(1, sys->nyc + sys->nxc) * sizeof(double)));
        HIP_TRY(hipMemcpyAsync(dx, x, sys->n * sizeof(double), hipMemcpyHostToDevice, st));
        double* oy = static_cast<double*>(dout);
        double* ox = oy + sys->nyc;
        hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, st, sys->nxc, ox);
        auto launch = [&](auto dimc, auto pc) {
            constexpr int D = decltype(dimc)::value, PP = decltype(pc)::value;
            const double* xs = static_cast<const double*>(dx);
            if (sys->nyc > 0)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gen_lambda_energy<D, PP>),
                                   dim3(static_cast<unsigned>((sys->nyc + 255) / 256)), dim3(256), 0, st, d, xs, oy);
            if (sys->nbot > 0)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gen_trace<D, PP>),
                                   dim3(static_cast<unsigned>((sys->nbot + 255) / 256)), dim3(256), 0, st, d, xs, ox);
        };
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        if (dim == 1) {
            if (p == 1) launch(I1{}, I1{});
            else if (p == 2) launch(I1{}, I2{});
            else launch(I1{}, I3{});
        } else {
            if (p == 1) launch(I2{}, I1{});
            else if (p == 2) launch(I2{}, I2{});
            else launch(I2{}, I3{});
        }
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(lambda_y2, oy, sys->nyc * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(trace2, ox, sys->nxc * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipFree(dx));
        HIP_TRY(hipFree(dout));
        return APF_OK;
    } catch (const HipFailure& e) {
        if (dx) (void)hipFree(dx);
        if (dout) (void)hipFree(dout);
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_x_indicators: ") + e.what());
    }
}

int apf_general_graph_norm(apf_general_system* sys, const apf_general_coeff* coeff, apf_csr_matrix** out)
{
    if (!sys || !coeff || !out) return fail(sys ? sys->ctx : nullptr, APF_BAD_ARG, "apf_general_graph_norm: null argument");
    apf_ctx* ctx = sys->ctx;
    if (!sys->assembled) return fail(ctx, APF_BAD_ARG, "apf_general_graph_norm: assemble the system first");
    if (coeff->grid_n < 0 || (coeff->grid_n > 0 && !coeff->M_grid))
        return fail(ctx, APF_BAD_ARG, "apf_general_graph_norm: bad M grid");
    std::vector<void*> tmp;
    auto M = std::make_unique<apf_csr_matrix>();
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        std::vector<int64_t> bt_rp, h_rp;
        std::vector<int> bt_col, h_col;
        const auto t_p0 = std::chrono::steady_clock::now();
        graph_patterns(sys->host, bt_rp, bt_col, h_rp, h_col);
        const auto t_p1 = std::chrono::steady_clock::now();
        auto dalloc = [&](size_t bytes) {
            void* q = nullptr;
            HIP_TRY(hipMalloc(&q, std::max<size_t>(bytes, 8)));
            return q;
        };
        GraphDev g{};
        void* p_btrp = dalloc(bt_rp.size() * sizeof(int64_t));
        tmp.push_back(p_btrp);
        void* p_btcol = dalloc(bt_col.size() * sizeof(int));
        tmp.push_back(p_btcol);
        void* p_btval = dalloc(bt_col.size() * sizeof(double));
        tmp.push_back(p_btval);
        void* p_err = dalloc(2 * sizeof(int));
        tmp.push_back(p_err);
        M->ctx = ctx;
        M->n = sys->nX;
        M->nnz = static_cast<int64_t>(h_col.size());
        M->rowptr = static_cast<int64_t*>(dalloc(h_rp.size() * sizeof(int64_t)));
        M->col = static_cast<int*>(dalloc(std::max<size_t>(1, h_col.size()) * sizeof(int)));
        M->val = static_cast<double*>(dalloc(std::max<size_t>(1, h_col.size()) * sizeof(double)));
        M->owned = true;
        HIP_TRY(hipMemcpyAsync(p_btrp, bt_rp.data(), bt_rp.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        if (!bt_col.empty())
            HIP_TRY(hipMemcpyAsync(p_btcol, bt_col.data(), bt_col.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(M->rowptr, h_rp.data(), h_rp.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
        if (!h_col.empty())
            HIP_TRY(hipMemcpyAsync(M->col, h_col.data(), h_col.size() * sizeof(int), hipMemcpyHostToDevice, st));
        g.bt_rowptr = static_cast<int64_t*>(p_btrp);
        g.bt_col = static_cast<int*>(p_btcol);
        g.bt_val = static_cast<double*>(p_btval);
        g.h_rowptr = M->rowptr;
        g.h_col = M->col;
        g.h_val = M->val;
        GenDev d = sys->d;
        d.err = static_cast<int*>(p_err);
        d.grid_n = coeff->grid_n;
        if (coeff->grid_n > 0) {
            int64_t len = 1;
            for (int a = 0; a < sys->dim; ++a) len *= coeff->grid_n;
            len *= sys->dim * sys->dim;
            if (len > sys->Mgrid_len) {
                sys->Mgrid = sys->alloc<double>(static_cast<size_t>(len));
                sys->Mgrid_len = len;
            }
            HIP_TRY(hipMemcpyAsync(sys->Mgrid, coeff->M_grid, len * sizeof(double), hipMemcpyHostToDevice, st));
            d.Mgrid = sys->Mgrid;
        }
        hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, st, static_cast<int64_t>(bt_col.size()), g.bt_val);
        hipLaunchKernelGGL(k_zero, dim3(256), dim3(256), 0, st, M->nnz, M->val);
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, static_cast<int*>(p_err));
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, static_cast<int*>(p_err) + 1);
        auto launch = [&](auto dimc, auto pc) {
            constexpr int D = decltype(dimc)::value, PP = decltype(pc)::value;
            using S = GS<D, PP>;
            if (sys->nyc > 0)
                hipLaunchKernelGGL(HIP_KERNEL_NAME(k_gen_graph<D, PP>),
                                   dim3(static_cast<unsigned>((sys->nyc + S::CB - 1) / S::CB)), dim3(256), 0, st, d, g);
        };
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        if (sys->dim == 1) {
            if (sys->p == 1) launch(I1{}, I1{});
            else if (sys->p == 2) launch(I1{}, I2{});
            else launch(I1{}, I3{});
        } else {
            if (sys->p == 1) launch(I2{}, I1{});
            else if (sys->p == 2) launch(I2{}, I2{});
            else launch(I2{}, I3{});
        }
        if (sys->nY > 0)
            hipLaunchKernelGGL(k_hhat_schur, dim3(static_cast<unsigned>((sys->nY + 255) / 256)), dim3(256), 0, st, d, g);
        if (sys->nX > 0)
            hipLaunchKernelGGL(k_hhat_trace, dim3(static_cast<unsigned>((sys->nX + 255) / 256)), dim3(256), 0, st, d, g,
                               sys->nX);
        HIP_TRY(hipGetLastError());
        int err[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(err, p_err, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (std::getenv("APF_HHAT_TIMING")) {
            const auto t_p2 = std::chrono::steady_clock::now();
            auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
            fprintf(stderr, "graph_norm: patterns %.1f ms (nnz BT %zu, Hhat %zu), device %.1f ms\n", ms(t_p0, t_p1),
                    bt_col.size(), h_col.size(), ms(t_p1, t_p2));
        }
        for (void* q : tmp) (void)hipFree(q);
        tmp.clear();
        if (err[1]) {
            apf_csr_matrix_destroy(M.release());
            return fail(ctx, APF_BAD_ARG, "diagonal_schur_approximation: A_y diagonal must be positive.");
        }
        if (err[0]) {
            apf_csr_matrix_destroy(M.release());
            return fail(ctx, APF_BAD_ARG, "apf_general_graph_norm: " + std::to_string(err[0]) +
                                              " contributions outside the pattern");
        }
        *out = M.release();
        return APF_OK;
    } catch (const HipFailure& e) {
        for (void* q : tmp) (void)hipFree(q);
        if (M->rowptr) (void)hipFree(M->rowptr);
        if (M->col) (void)hipFree(M->col);
        if (M->val) (void)hipFree(M->val);
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_graph_norm: ") + e.what());
    }
}

int apf_general_matrix(apf_general_system* sys, apf_csr_matrix** out)
{
    if (!sys || !out) return fail(nullptr, APF_BAD_ARG, "apf_general_matrix: null argument");
    auto M = std::make_unique<apf_csr_matrix>();
    M->ctx = sys->ctx;
    M->n = sys->n;
    M->nnz = sys->nnz;
    M->rowptr = sys->rowptr;
    M->col = sys->col;
    M->val = sys->val;
    M->owned = false;
    *out = M.release();
    return APF_OK;
}

}  // extern "C"

// ============================================================== estimator on slab groups
// The estimator works slab by slab (time slabs between the global time endpoints,
// time_slabs/time_slab_builder.hpp:53-120): the slab's Y space, its reconstruction
// lambda_tilde (A_slab lt = f_slab - (B^T u)_slab, time_slab_reconstruction.hpp:584-735),
// the (slab, vertex) patch problems and the per-cell indicators.  Slabs that share one
// spatial mesh form a tensor mesh (that mesh x the slabs' time nodes), on which Y is the
// tensor Y space and u_delta, the X function restricted to those slabs, lies in the tensor
// X space (every X mesh of those slabs is that mesh or a coarser nested one).  So each
// group is a tensor system of this library: lambda restricted to the group's Y dofs and
// the nodal interpolant of u_delta go into its solution vector, and the tensor estimator
// (with the slab solve) runs on it; totals and per-cell values are summed over the groups.
namespace {

__global__ void k_gather(int64_t n, const int* __restrict__ map, const double* __restrict__ src,
                         double* __restrict__ dst)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[map[i]];
}

__global__ void k_interp(int64_t n, const int64_t* __restrict__ ptr, const int* __restrict__ col,
                         const double* __restrict__ w, const double* __restrict__ src, double* __restrict__ dst)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double v = 0.0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; ++e) v += w[e] * src[col[e]];
        dst[i] = v;
    }
}

// X basis of cell `a` (local index i_s * (p+1) + i_t) at the physical point (x, t)
void x_basis_at(const CellSpaceHost& X, int64_t a, const double* x, double t, std::vector<double>& out)
{
    const int dim = X.dim, p = X.p, Ns = dim == 1 ? p + 1 : (p + 1) * (p + 2) / 2, Nt = p + 1;
    const double* g = &X.geo[a * kGeo];
    double xi0, xi1 = 0.0;
    if (dim == 1) {
        xi0 = (x[0] - g[0]) / (g[1] - g[0]);
    } else {
        const double A = g[2] - g[0], B = g[4] - g[0], C = g[3] - g[1], D = g[5] - g[1];
        const double det = A * D - B * C, dx = x[0] - g[0], dy = x[1] - g[1];
        xi0 = (D * dx - B * dy) / det;
        xi1 = (-C * dx + A * dy) / det;
    }
    std::vector<double> sv(Ns), tv(Nt), d1(Ns), d2(Ns), dt(Nt);
    if (dim == 1) apf::segment_lagrange(p, xi0, sv.data(), d1.data());
    else apf::triangle_lagrange(p, xi0, xi1, sv.data(), d1.data(), d2.data());
    apf::segment_lagrange(p, (t - g[kGeo - 2]) / (g[kGeo - 1] - g[kGeo - 2]), tv.data(), dt.data());
    out.assign(static_cast<size_t>(Ns) * Nt, 0.0);
    for (int is = 0; is < Ns; ++is)
        for (int it = 0; it < Nt; ++it) out[is * Nt + it] = sv[is] * tv[it];
}

// physical position of local spatial node i_s of a tensor-space cell
void node_position(const apf::TensorMesh& mesh, int p, int c, int is, double* x)
{
    if (mesh.dim == 1) {
        const double x0 = mesh.vertices[mesh.cells[2 * c]], x1 = mesh.vertices[mesh.cells[2 * c + 1]];
        x[0] = x0 + (x1 - x0) * is / p;
        return;
    }
    int k = 0;
    for (int cc = 0; cc <= p; ++cc)
        for (int b = 0; b <= p - cc; ++b, ++k)
            if (k == is) {
                const int a = p - b - cc;
                const int* V = &mesh.cells[3 * c];
                for (int d = 0; d < 2; ++d)
                    x[d] = (a * mesh.vertices[2 * V[0] + d] + b * mesh.vertices[2 * V[1] + d] +
                            cc * mesh.vertices[2 * V[2] + d]) / p;
                return;
            }
}

struct EstGroup {
    apf_space* space = nullptr;
    apf_system* sys = nullptr;
    int64_t first_y_cell = 0;
    int64_t nYg = 0, nXg = 0, ncells = 0;
    int* d_ymap = nullptr;       // group Y dof -> general Y dof
    bool sliced = false;         // Y cells cut into slabs: lambda interpolated, outputs summed
    std::vector<int> src;        // sliced: group cell (slab k, spatial cell c) -> source Y cell
    int64_t* d_lptr = nullptr;   // sliced: group Y dof -> interpolation row over general Y dofs
    int* d_lcol = nullptr;
    double* d_lw = nullptr;
    int64_t* d_iptr = nullptr;   // group X dof -> interpolation row over general X dofs
    int* d_icol = nullptr;
    double* d_iw = nullptr;
    bool assembled = false;
};

}  // namespace

struct apf_general_estimator {
    apf_general_system* gsys = nullptr;   // used by the calls, never by destroy (it may be gone)
    int device = 0;
    std::vector<EstGroup> groups;
    double* d_x = nullptr;           // the general solution [lambda; u]
    std::vector<void*> allocs;
};

extern "C" {

}  // extern "C"

namespace {

// Groups of slabs as tensor systems.  Plain mode (first_y_cell): the group's cells are the
// general system's Y cells first_y_cell[g] + k * ns + c.  Sliced mode (cell_ptr, cell_source):
// cell (k, c) of group g is a slice of the Y cell cell_source[cell_ptr[g] + k * ns + c].
int estimator_create(apf_general_system* gsys, int n_groups, apf_space* const* spaces, const int64_t* first_y_cell,
                     const int64_t* cell_ptr, const int* cell_source, apf_general_estimator** out)
{
    const bool sliced = cell_source != nullptr;
    if (!gsys || !spaces || !out || n_groups < 1 || (!sliced && !first_y_cell) || (sliced && !cell_ptr))
        return fail(gsys ? gsys->ctx : nullptr, APF_BAD_ARG, "apf_general_estimator_create: bad argument");
    apf_ctx* ctx = gsys->ctx;
    auto ge = std::make_unique<apf_general_estimator>();
    ge->gsys = gsys;
    ge->device = ctx->device;
    auto cleanup = [&]() {
        for (auto& g : ge->groups)
            if (g.sys) apf_system_destroy(g.sys);
        for (void* q : ge->allocs) (void)hipFree(q);
    };
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        const CellSpaceHost& X = gsys->host.X;
        const CellSpaceHost& Y = gsys->host.Y;
        const int nlx = n_local(X.dim, X.p);
        std::vector<int64_t> covered(static_cast<size_t>(Y.nc), 0);
        auto dev = [&](auto& v) {
            using T = typename std::decay_t<decltype(v)>::value_type;
            void* q = nullptr;
            HIP_TRY(hipMalloc(&q, std::max<size_t>(1, v.size()) * sizeof(T)));
            ge->allocs.push_back(q);
            if (!v.empty()) HIP_TRY(hipMemcpy(q, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
            return static_cast<T*>(q);
        };
        for (int gi = 0; gi < n_groups; ++gi) {
            EstGroup G;
            G.space = spaces[gi];
            if (!G.space) throw std::invalid_argument("null group space");
            const apf::Space& S = G.space->S;
            if (S.dim != Y.dim || S.p != Y.p) throw std::invalid_argument("group space differs in dim or p");
            if (!G.space->has_pattern) throw std::invalid_argument("group space needs its K pattern");
            if (G.space->H.coarse.empty() && !sliced)
                throw std::invalid_argument("group space needs coarse meshes (the slab solve's multigrid)");
            const int ns = S.mesh.ns(), m = S.mesh.m(), p = S.p, Ns = S.Ns, Nt = S.Nt, nf = S.nf;
            G.ncells = static_cast<int64_t>(ns) * m;
            G.nYg = S.nY;
            G.nXg = S.nX;
            G.sliced = sliced;
            if (sliced) {
                G.first_y_cell = 0;
                G.src.assign(cell_source + cell_ptr[gi], cell_source + cell_ptr[gi] + G.ncells);
                if (cell_ptr[gi + 1] - cell_ptr[gi] != G.ncells)
                    throw std::invalid_argument("group " + std::to_string(gi) + ": cell_ptr range differs from its mesh");
            } else {
                G.first_y_cell = first_y_cell[gi];
                if (G.first_y_cell < 0 || G.first_y_cell + G.ncells > Y.nc)
                    throw std::invalid_argument("group cells outside the Y cells");
                G.src.resize(static_cast<size_t>(G.ncells));
                for (int64_t q = 0; q < G.ncells; ++q) G.src[q] = static_cast<int>(G.first_y_cell + q);
            }
            for (int v : G.src)
                if (v < 0 || v >= Y.nc) throw std::invalid_argument("group cell source outside the Y cells");
            // Y map: the group's Y cells must be its mesh's cells slab by slab with identity rows
            // (sliced: the source cell's polynomial evaluated at the slab's time nodes)
            std::vector<int> ymap(static_cast<size_t>(S.nY), -1);
            std::vector<std::vector<std::pair<int, double>>> lrows(sliced ? static_cast<size_t>(S.nY) : 0);
            std::vector<char> lset(sliced ? static_cast<size_t>(S.nY) : 0, 0);
            std::vector<double> lv(Nt), ld(Nt);
            for (int k = 0; k < m; ++k)
                for (int c = 0; c < ns; ++c) {
                    const int64_t yc = G.src[static_cast<size_t>(k) * ns + c];
                    if (covered[yc]++ && !sliced) throw std::invalid_argument("a Y cell in two groups");
                    const double* gy = &Y.geo[yc * kGeo];
                    if (sliced) {
                        if (S.mesh.times[k] < gy[kGeo - 2] - 1e-13 || S.mesh.times[k + 1] > gy[kGeo - 1] + 1e-13)
                            throw std::invalid_argument("a slab outside its source Y cell's time interval");
                    } else if (std::fabs(gy[kGeo - 2] - S.mesh.times[k]) > 1e-13 ||
                               std::fabs(gy[kGeo - 1] - S.mesh.times[k + 1]) > 1e-13) {
                        throw std::invalid_argument("group slab times differ from its Y cells'");
                    }
                    for (int v = 0; v <= S.dim; ++v)
                        for (int d = 0; d < S.dim; ++d)
                            if (std::fabs(gy[v * S.dim + d] -
                                          S.mesh.vertices[S.mesh.cells[c * (S.dim + 1) + v] * S.dim + d]) > 1e-13)
                                throw std::invalid_argument("group cell geometry differs from its Y cell's");
                    if (sliced) {
                        for (int is = 0; is < Ns; ++is) {
                            const int f = S.cell_free[static_cast<size_t>(c) * Ns + is];
                            for (int it = 0; it < Nt; ++it) {
                                const int loc = is * Nt + it;
                                const int64_t r0 = Y.ptr[yc * (Ns * Nt) + loc], r1 = Y.ptr[yc * (Ns * Nt) + loc + 1];
                                if (f < 0) {
                                    if (r1 != r0) throw std::invalid_argument("Dirichlet node with a non-empty Y row");
                                    continue;
                                }
                                const int64_t td = (static_cast<int64_t>(k) * Nt + it) * nf + f;
                                if (lset[td]) continue;
                                lset[td] = 1;
                                const double t = S.mesh.times[k] + (S.mesh.times[k + 1] - S.mesh.times[k]) * it / p;
                                apf::segment_lagrange(p, (t - gy[kGeo - 2]) / (gy[kGeo - 1] - gy[kGeo - 2]), lv.data(),
                                                      ld.data());
                                auto& row = lrows[td];
                                for (int j = 0; j < Nt; ++j) {
                                    if (lv[j] == 0.0) continue;
                                    const int lj = is * Nt + j;
                                    for (int64_t e = Y.ptr[yc * (Ns * Nt) + lj]; e < Y.ptr[yc * (Ns * Nt) + lj + 1]; ++e)
                                        row.emplace_back(Y.dof[e], Y.w[e] * lv[j]);
                                }
                            }
                        }
                        continue;
                    }
                    for (int is = 0; is < Ns; ++is)
                        for (int it = 0; it < Nt; ++it) {
                            const int loc = is * Nt + it;
                            const int64_t r0 = Y.ptr[yc * (Ns * Nt) + loc], r1 = Y.ptr[yc * (Ns * Nt) + loc + 1];
                            const int f = S.cell_free[static_cast<size_t>(c) * Ns + is];
                            if (f < 0) {
                                if (r1 != r0) throw std::invalid_argument("Dirichlet node with a non-empty Y row");
                                continue;
                            }
                            if (r1 - r0 != 1 || Y.w[r0] != 1.0)
                                throw std::invalid_argument("constrained Y dofs inside a slab are not supported");
                            const int64_t td = (static_cast<int64_t>(k) * Nt + it) * nf + f;
                            if (ymap[td] >= 0 && ymap[td] != Y.dof[r0])
                                throw std::invalid_argument("inconsistent Y numbering in a slab");
                            ymap[td] = Y.dof[r0];
                        }
                }
            std::vector<int64_t> lptr(1, 0);
            std::vector<int> lcol;
            std::vector<double> lw;
            if (sliced) {
                for (int64_t td = 0; td < S.nY; ++td) {
                    if (!lset[td]) throw std::invalid_argument("group Y dof without a source cell");
                    auto& row = lrows[td];
                    std::sort(row.begin(), row.end());
                    for (size_t q = 0; q < row.size(); ++q) {
                        if (static_cast<int64_t>(lcol.size()) > lptr.back() && lcol.back() == row[q].first)
                            lw.back() += row[q].second;
                        else {
                            lcol.push_back(row[q].first);
                            lw.push_back(row[q].second);
                        }
                    }
                    lptr.push_back(static_cast<int64_t>(lcol.size()));
                }
                std::fill(ymap.begin(), ymap.end(), 0);
            }
            for (int v : ymap)
                if (v < 0) throw std::invalid_argument("group Y dof without a general Y dof");
            // X interpolation: each group X node evaluated on the X ancestor of a Y cell that holds it
            std::vector<int> node_cell(nf, -1), node_loc(nf, -1);
            for (int c = 0; c < ns; ++c)
                for (int is = 0; is < Ns; ++is) {
                    const int f = S.cell_free[static_cast<size_t>(c) * Ns + is];
                    if (f >= 0 && node_cell[f] < 0) {
                        node_cell[f] = c;
                        node_loc[f] = is;
                    }
                }
            std::vector<int64_t> iptr(1, 0);
            std::vector<int> icol;
            std::vector<double> iw, bv;
            const int ntn = m * p + 1;
            for (int tn = 0; tn < ntn; ++tn) {
                const int k = std::min(tn / p, m - 1), l = tn - k * p;
                const double t = S.mesh.times[k] + (S.mesh.times[k + 1] - S.mesh.times[k]) * l / p;
                for (int f = 0; f < nf; ++f) {
                    double x[2];
                    node_position(S.mesh, p, node_cell[f], node_loc[f], x);
                    const int64_t yc = G.src[static_cast<size_t>(k) * ns + node_cell[f]];
                    const int64_t a = gsys->host.anc[yc];
                    x_basis_at(X, a, x, t, bv);
                    std::vector<std::pair<int, double>> acc;
                    for (int loc = 0; loc < nlx; ++loc) {
                        const double vv = bv[loc];
                        if (vv == 0.0) continue;
                        for (int64_t e = X.ptr[a * nlx + loc]; e < X.ptr[a * nlx + loc + 1]; ++e)
                            acc.emplace_back(X.dof[e], X.w[e] * vv);
                    }
                    std::sort(acc.begin(), acc.end());
                    for (size_t q = 0; q < acc.size(); ++q) {
                        if (!icol.empty() && static_cast<int64_t>(icol.size()) > iptr.back() && icol.back() == acc[q].first)
                            iw.back() += acc[q].second;
                        else {
                            icol.push_back(acc[q].first);
                            iw.push_back(acc[q].second);
                        }
                    }
                    iptr.push_back(static_cast<int64_t>(icol.size()));
                }
            }
            G.d_ymap = dev(ymap);
            if (sliced) {
                G.d_lptr = dev(lptr);
                G.d_lcol = dev(lcol);
                G.d_lw = dev(lw);
            }
            G.d_iptr = dev(iptr);
            G.d_icol = dev(icol);
            G.d_iw = dev(iw);
            apf_system* sys_g = nullptr;
            if (int c = apf_system_create(ctx, G.space, &sys_g)) throw std::runtime_error(g_last_error + " (code " + std::to_string(c) + ")");
            sys_g->slab_only = true;   // slab solves only: no X-column shifts in its multigrid
            // a group mesh without coarser meshes: its multigrid is the one level, solved directly
            if (sliced && !sys_g->H) sys_g->H = &G.space->H;
            G.sys = sys_g;
            ge->groups.push_back(G);
        }
        for (int64_t c = 0; c < Y.nc; ++c)
            if (!covered[c]) throw std::invalid_argument("Y cell " + std::to_string(c) + " in no group");
        void* q = nullptr;
        HIP_TRY(hipMalloc(&q, gsys->n * sizeof(double)));
        ge->allocs.push_back(q);
        ge->d_x = static_cast<double*>(q);
        *out = ge.release();
        return APF_OK;
    } catch (const HipFailure& e) {
        cleanup();
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_estimator_create: ") + e.what());
    } catch (const std::exception& e) {
        cleanup();
        return fail(ctx, APF_BAD_ARG, std::string("apf_general_estimator_create: ") + e.what());
    }
}

}  // namespace

extern "C" {

int apf_general_estimator_create(apf_general_system* gsys, int n_groups, apf_space* const* spaces,
                                 const int64_t* first_y_cell, apf_general_estimator** out)
{
    return estimator_create(gsys, n_groups, spaces, first_y_cell, nullptr, nullptr, out);
}

int apf_general_estimator_create_sliced(apf_general_system* gsys, int n_groups, apf_space* const* spaces,
                                        const int64_t* cell_ptr, const int* cell_source, apf_general_estimator** out)
{
    if (!cell_ptr || !cell_source)
        return fail(gsys ? gsys->ctx : nullptr, APF_BAD_ARG, "apf_general_estimator_create_sliced: null argument");
    return estimator_create(gsys, n_groups, spaces, nullptr, cell_ptr, cell_source, out);
}

void apf_general_estimator_destroy(apf_general_estimator* ge)
{
    if (!ge) return;
    (void)hipSetDevice(ge->device);
    for (auto& g : ge->groups)
        if (g.sys) apf_system_destroy(g.sys);
    for (void* q : ge->allocs) (void)hipFree(q);
    delete ge;
}

int apf_general_estimator_assemble(apf_general_estimator* ge, int group, const apf_coeff* coeff)
{
    if (!ge || !coeff || group < 0 || group >= static_cast<int>(ge->groups.size()))
        return fail(ge ? ge->gsys->ctx : nullptr, APF_BAD_ARG, "apf_general_estimator_assemble: bad argument");
    if (int c = apf_assemble_main(ge->groups[group].sys, coeff, APF_ASSEMBLE_KRON)) return c;
    ge->groups[group].assembled = true;
    return APF_OK;
}

int apf_general_estimate(apf_general_estimator* ge, const double* x, int prec, int refine_steps, apf_estimate_out* out,
                         double* flux_by_cell, double* resid_by_cell, double* recon_by_cell)
{
    if (!ge || !x || !out) return fail(ge ? ge->gsys->ctx : nullptr, APF_BAD_ARG, "apf_general_estimate: null argument");
    apf_ctx* ctx = ge->gsys->ctx;
    for (auto& g : ge->groups)
        if (!g.assembled) return fail(ctx, APF_BAD_ARG, "apf_general_estimate: assemble every group first");
    try {
        HIP_TRY(hipSetDevice(ctx->device));
        hipStream_t st = ctx->stream;
        const int64_t nY = ge->gsys->nY;
        HIP_TRY(hipMemcpyAsync(ge->d_x, x, ge->gsys->n * sizeof(double), hipMemcpyHostToDevice, st));
        const int64_t nyc = ge->gsys->nyc;
        for (double* a : {flux_by_cell, resid_by_cell, recon_by_cell})
            if (a) std::fill(a, a + nyc, 0.0);
        apf_estimate_out tot{};
        tot.slab_stop = APF_SLAB_STOP_TOL;
        double div2 = 0.0;
        for (auto& g : ge->groups) {
            apf::System* sg = g.sys;
            if (g.sliced)
                hipLaunchKernelGGL(k_interp, dim3(1024), dim3(256), 0, st, g.nYg, g.d_lptr, g.d_lcol, g.d_lw, ge->d_x,
                                   sg->d_x);
            else
                hipLaunchKernelGGL(k_gather, dim3(1024), dim3(256), 0, st, g.nYg, g.d_ymap, ge->d_x, sg->d_x);
            hipLaunchKernelGGL(k_interp, dim3(1024), dim3(256), 0, st, g.nXg, g.d_iptr, g.d_icol, g.d_iw, ge->d_x + nY,
                               sg->d_x + g.nYg);
            HIP_TRY(hipGetLastError());
            HIP_TRY(hipStreamSynchronize(st));
            apf_estimate_out o{};
            int code;
            if (g.sliced) {
                // per-slice values summed into the source cells (by_source_cell)
                std::vector<double> fl(static_cast<size_t>(g.ncells)), rs(fl.size()), rc(fl.size());
                code = apf_estimate_ex(g.sys, prec, refine_steps, APF_SLAB_SOLVE, &o, fl.data(), rs.data(), rc.data());
                if (code == APF_OK)
                    for (int64_t q = 0; q < g.ncells; ++q) {
                        if (flux_by_cell) flux_by_cell[g.src[q]] += fl[q];
                        if (resid_by_cell) resid_by_cell[g.src[q]] += rs[q];
                        if (recon_by_cell) recon_by_cell[g.src[q]] += rc[q];
                    }
            } else {
                const int64_t off = g.first_y_cell;
                code = apf_estimate_ex(g.sys, prec, refine_steps, APF_SLAB_SOLVE, &o,
                                       flux_by_cell ? flux_by_cell + off : nullptr,
                                       resid_by_cell ? resid_by_cell + off : nullptr,
                                       recon_by_cell ? recon_by_cell + off : nullptr);
            }
            if (code != APF_OK) return code;
            tot.lambda_y2 += o.lambda_y2;
            // ||u0 - u(0)||^2 lives on the group that starts at t = 0 (a later group's
            // bottom is an interior time)
            if (g.space->S.mesh.times[0] == 0.0) tot.trace2 += o.trace2;
            tot.flux2 += o.flux2;
            tot.recon2 += o.recon2;
            div2 += o.div_residual_l2 * o.div_residual_l2;
            tot.max_patch_residual = std::max(tot.max_patch_residual, o.max_patch_residual);
            tot.seconds_patches += o.seconds_patches;
            tot.seconds_diagnostics += o.seconds_diagnostics;
            tot.seconds_slab += o.seconds_slab;
            tot.extra_refinements += o.extra_refinements;
            tot.fullpiv_patches += o.fullpiv_patches;
            tot.seconds_total += o.seconds_total;
            tot.slab_iterations = std::max(tot.slab_iterations, o.slab_iterations);
            tot.slab_relative_residual = std::max(tot.slab_relative_residual, o.slab_relative_residual);
            if (o.slab_stop != APF_SLAB_STOP_TOL) tot.slab_stop = o.slab_stop;
        }
        tot.eta2 = tot.lambda_y2 + tot.trace2;
        tot.y_est2 = tot.flux2 + tot.recon2;
        tot.div_residual_l2 = std::sqrt(div2);
        *out = tot;
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(ctx, APF_HIP_ERROR, std::string("apf_general_estimate: ") + e.what());
    }
}

}  // extern "C"
