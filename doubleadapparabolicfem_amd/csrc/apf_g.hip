// G-estimator (algorithm/adaptive_driver.hpp:2680-2905) on tensor meshes.
//
// G is the Y space one degree higher in space and time, discontinuous in time
// (adaptive_driver.hpp:2697-2710), so it is the Y block of a degree-(p+1)
// system on the same mesh.  Given the degree-p solution (lambda, u):
//   1. interpolate lambda and u nodally into the degree-(p+1) spaces
//      (prolong_true_coefficients_nodal, adaptive_driver.hpp:2832-2836);
//   2. solve A_G lambda_G = f_G - B^T (I u) with the slab solve of the
//      degree-(p+1) system (A_G is block diagonal over the slabs).  B(X_p, G)^T u
//      equals B(X_{p+1}, G)^T (I u): the interpolant is the same function and the
//      degree-(p+1) cell rule integrates the polynomial integrand exactly;
//   3. G^2 = d^T A_G d with d = lambda_G - I lambda (adaptive_driver.hpp:2838-2866),
//      with the negative check of :2869-2883 and the clamp at zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

#include "../../include/apf.h"
#include "apf_internal.hpp"
#include "apf_tables.hpp"

using namespace apf;

extern thread_local std::string g_last_error;

namespace {

// dst[(tr) nfq + f] = sum_e val[e] sum_j It[l][j] src[(s0 + j) nfp + col[e]]
// for the time row tr of the degree-q vector: in X (continuous in time) tr = k q + l,
// in Y (per slab) tr = k (q + 1) + l.  One thread per output entry.
__global__ void k_prolong(int64_t total, int nfq, int nfp, int p, int q, int m, int cont,
                          const int* __restrict__ ptr, const int* __restrict__ col, const double* __restrict__ val,
                          const double* __restrict__ It, const double* __restrict__ src, double* __restrict__ dst)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t tr = t / nfq;
    const int f = static_cast<int>(t - tr * nfq);
    int k, l;
    int64_t s0;
    if (cont) {
        k = min(static_cast<int>(tr / q), m - 1);
        l = static_cast<int>(tr - static_cast<int64_t>(k) * q);
        s0 = static_cast<int64_t>(k) * p;
    } else {
        k = static_cast<int>(tr / (q + 1));
        l = static_cast<int>(tr - static_cast<int64_t>(k) * (q + 1));
        s0 = static_cast<int64_t>(k) * (p + 1);
    }
    double acc = 0.0;
    for (int e = ptr[f]; e < ptr[f + 1]; ++e) {
        const int c = col[e];
        double v = 0.0;
        for (int j = 0; j <= p; ++j) v += It[l * (p + 1) + j] * src[(s0 + j) * nfp + c];
        acc += val[e] * v;
    }
    dst[t] = acc;
}

// t = [lt - il; 0]
__global__ void k_gdiff(int64_t nY, int64_t n, const double* __restrict__ lt, const double* __restrict__ il,
                        double* __restrict__ t)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        t[i] = i < nY ? lt[i] - il[i] : 0.0;
}

// per block: part[2 b] = sum d_i y_i, part[2 b + 1] = sum d_i^2 over i < nY
__global__ void __launch_bounds__(256) k_gdot_parts(int64_t nY, const double* __restrict__ d,
                                                    const double* __restrict__ y, double* __restrict__ part)
{
    __shared__ double s0[256], s1[256];
    double a = 0.0, b = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < nY;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        a += d[i] * y[i];
        b += d[i] * d[i];
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}

// Reference coordinates of local node j of the degree-q element: the lattice
// point where basis function j equals one.
void node_ref(int dim, int q, int j, double& x, double& y)
{
    std::vector<double> v(n_space_basis(dim, q)), gx(v.size()), gy(v.size());
    for (int c = 0; c <= (dim == 2 ? q : 0); ++c)
        for (int b = 0; b <= q - c; ++b) {
            const double xb = static_cast<double>(b) / q, yc = static_cast<double>(c) / q;
            if (dim == 1)
                segment_lagrange(q, xb, v.data(), gx.data());
            else
                triangle_lagrange(q, xb, yc, v.data(), gx.data(), gy.data());
            if (v[j] > 0.5) {
                x = xb;
                y = yc;
                return;
            }
        }
    throw std::runtime_error("apf_g_estimate: no lattice node for a local basis function");
}

void space_basis(int dim, int p, double x, double y, double* v)
{
    std::vector<double> gx(n_space_basis(dim, p)), gy(gx.size());
    if (dim == 1)
        segment_lagrange(p, x, v, gx.data());
    else
        triangle_lagrange(p, x, y, v, gx.data(), gy.data());
}

// Spatial interpolation Sp -> Sq (free nodes; Dirichlet nodes carry zero) and
// the time matrix, uploaded to the degree-q system.
void build_prolongation(System* gsys, const Space& Sp)
{
    const Space& Sq = *gsys->space;
    const int dim = Sq.dim, p = Sp.p, q = Sq.p;
    const int Nsp = Sp.Ns, Nsq = Sq.Ns, ns = Sq.mesh.ns();
    std::vector<double> phi(static_cast<size_t>(Nsq) * Nsp);
    for (int j = 0; j < Nsq; ++j) {
        double x = 0.0, y = 0.0;
        node_ref(dim, q, j, x, y);
        space_basis(dim, p, x, y, &phi[static_cast<size_t>(j) * Nsp]);
    }
    std::vector<int> ptr(Sq.nf + 1, 0), row_cell(Sq.nf, -1), row_node(Sq.nf, -1);
    for (int c = 0; c < ns; ++c)
        for (int j = 0; j < Nsq; ++j) {
            const int f = Sq.cell_free[static_cast<size_t>(c) * Nsq + j];
            if (f >= 0 && row_cell[f] < 0) {
                row_cell[f] = c;
                row_node[f] = j;
            }
        }
    std::vector<int> col;
    std::vector<double> val;
    for (int f = 0; f < Sq.nf; ++f) {
        const int c = row_cell[f], j = row_node[f];
        if (c < 0) throw std::runtime_error("apf_g_estimate: a free node of the degree-(p+1) space has no cell");
        for (int i = 0; i < Nsp; ++i) {
            const int fp = Sp.cell_free[static_cast<size_t>(c) * Nsp + i];
            const double w = phi[static_cast<size_t>(j) * Nsp + i];
            if (fp >= 0 && w != 0.0) {
                col.push_back(fp);
                val.push_back(w);
            }
        }
        ptr[f + 1] = static_cast<int>(col.size());
    }
    std::vector<double> It(static_cast<size_t>(q + 1) * (p + 1)), dv(p + 1);
    for (int l = 0; l <= q; ++l) segment_lagrange(p, static_cast<double>(l) / q, &It[static_cast<size_t>(l) * (p + 1)], dv.data());
    gsys->d_g_ptr = dalloc<int>(gsys, ptr.size());
    gsys->d_g_col = dalloc<int>(gsys, col.size());
    gsys->d_g_val = dalloc<double>(gsys, val.size());
    gsys->d_g_It = dalloc<double>(gsys, It.size());
    if (!gsys->d_g_il) gsys->d_g_il = dalloc<double>(gsys, Sq.nY);
    if (!gsys->d_g_lt) gsys->d_g_lt = dalloc<double>(gsys, Sq.nY);
    HIP_TRY(hipMemcpy(gsys->d_g_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!col.empty()) {
        HIP_TRY(hipMemcpy(gsys->d_g_col, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(gsys->d_g_val, val.data(), val.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMemcpy(gsys->d_g_It, It.data(), It.size() * sizeof(double), hipMemcpyHostToDevice));
    gsys->g_src = &Sp;
}

bool same_mesh(const TensorMesh& a, const TensorMesh& b)
{
    return a.dim == b.dim && a.vertices == b.vertices && a.cells == b.cells && a.times == b.times;
}

}  // namespace

extern "C" int apf_g_estimate(apf_system* sys, apf_system* gsys, double tol, int max_it, apf_g_estimate_out* out)
{
    auto fail = [&](int code, const std::string& m) {
        g_last_error = m;
        if (sys && sys->ctx) sys->ctx->error = m;
        return code;
    };
    if (!sys || !gsys || !out) return fail(APF_BAD_ARG, "apf_g_estimate: null argument");
    if (sys->ctx != gsys->ctx) return fail(APF_BAD_ARG, "apf_g_estimate: the two systems belong to different contexts");
    if (!sys->assembled || !gsys->assembled)
        return fail(APF_BAD_ARG, "apf_g_estimate: assemble both systems and solve the degree-p one first");
    const Space& Sp = *sys->space;
    const Space& Sq = *gsys->space;
    if (Sq.p != Sp.p + 1) return fail(APF_BAD_ARG, "apf_g_estimate: the G system must have degree p+1");
    if (!same_mesh(Sp.mesh, Sq.mesh)) return fail(APF_BAD_ARG, "apf_g_estimate: the two systems are on different meshes");
    if (!(tol > 0.0) || max_it < 1) return fail(APF_BAD_ARG, "apf_g_estimate: tol must be > 0 and max_it >= 1");
    try {
        HIP_TRY(hipSetDevice(sys->ctx->device));
        const auto t0 = std::chrono::steady_clock::now();
        HIP_TRY(hipStreamSynchronize(sys->stream));
        // The numbering of a space is a function of its mesh and degree, both checked
        // above, so the interpolation built on the first call holds for every later one.
        if (!gsys->g_src) build_prolongation(gsys, Sp);
        hipStream_t st = gsys->stream;
        const int m = Sq.mesh.m();
        const int64_t nXq = static_cast<int64_t>(m * Sq.p + 1) * Sq.nf;
        const int64_t nYq = Sq.nY;
        // I u into the X block of the G system's vector, I lambda into its Y block and d_g_il
        hipLaunchKernelGGL(k_prolong, dim3(static_cast<unsigned>((nXq + 255) / 256)), dim3(256), 0, st, nXq, Sq.nf,
                           Sp.nf, Sp.p, Sq.p, m, 1, gsys->d_g_ptr, gsys->d_g_col, gsys->d_g_val, gsys->d_g_It,
                           sys->d_x + Sp.nY, gsys->d_x + nYq);
        hipLaunchKernelGGL(k_prolong, dim3(static_cast<unsigned>((nYq + 255) / 256)), dim3(256), 0, st, nYq, Sq.nf,
                           Sp.nf, Sp.p, Sq.p, m, 0, gsys->d_g_ptr, gsys->d_g_col, gsys->d_g_val, gsys->d_g_It, sys->d_x,
                           gsys->d_g_il);
        HIP_TRY(hipMemcpyAsync(gsys->d_x, gsys->d_g_il, nYq * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipGetLastError());
        // lambda_G from I lambda: A_G lambda_G = f_G - B^T I u
        int its = 0, stop = APF_SLAB_STOP_NONE;
        double rel = 0.0;
        slab_reconstruct(gsys, gsys->d_g_lt, tol, max_it, &its, &rel, &stop);
        out->solver_stop = stop;
        // d^T A_G d: the Y rows of K [d; 0]
        double* t = gsys->d_vec[2];
        double* y = gsys->d_vec[3];
        const int nparts = gsys->nparts;
        hipLaunchKernelGGL(k_gdiff, dim3(nparts), dim3(256), 0, st, nYq, Sq.n, gsys->d_g_lt, gsys->d_g_il, t);
        gsys->spmv_rows(t, y, nYq);
        hipLaunchKernelGGL(k_gdot_parts, dim3(nparts), dim3(256), 0, st, nYq, t, y, gsys->d_part);
        std::vector<double> parts(2 * static_cast<size_t>(nparts));
        HIP_TRY(hipMemcpyAsync(parts.data(), gsys->d_part, parts.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        double raw = 0.0, dd = 0.0;
        for (int b = 0; b < nparts; ++b) {
            raw += parts[2 * b];
            dd += parts[2 * b + 1];
        }
        out->raw = raw;
        out->coefficient_difference_squared = dd;
        out->solver_iterations = its;
        out->solver_relative_residual = rel;
        out->g_true_dofs = nYq;
        out->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double neg_tol = 1000.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, dd);
        if (raw < -neg_tol) {
            out->lambda_difference_squared = out->lambda_difference = -1.0;
            return fail(APF_BAD_ARG, "apf_g_estimate: negative lambda-difference energy norm squared (" +
                                         std::to_string(raw) + ")");
        }
        out->lambda_difference_squared = std::max(0.0, raw);
        out->lambda_difference = std::sqrt(out->lambda_difference_squared);
        return APF_OK;
    } catch (const HipFailure& e) {
        return fail(APF_HIP_ERROR, std::string("apf_g_estimate: ") + e.what());
    } catch (const std::exception& e) {
        return fail(APF_BAD_ARG, std::string("apf_g_estimate: ") + e.what());
    }
}
