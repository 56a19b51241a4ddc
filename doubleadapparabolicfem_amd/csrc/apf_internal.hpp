// Internal types shared by the HIP translation units (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include <memory>

#include "apf_mg.hpp"
#include "apf_mg_dev.hpp"
#include "apf_space.hpp"
#include "apf_tables.hpp"

struct apf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // default priority (no priority streams)
    // second stream, shared by the context's systems and solves (calls on one context are
    // sequential): the estimate's factor phase beside its slab solve, the caller's-CSR
    // MINRES X-block chain beside the lambda block's.  Created once: a stream costs about
    // a millisecond to create and destroy.
    hipStream_t stream2 = nullptr;
    std::string error;
};

// A symmetric CSR matrix on the device (apf_csr.hip).  `owned`: the arrays belong to the
// matrix; a view of another object's K (apf_general_matrix) borrows them.
struct apf_csr_matrix {
    apf_ctx* ctx = nullptr;
    int64_t n = 0, nnz = 0;
    int64_t* rowptr = nullptr;
    int* col = nullptr;
    double* val = nullptr;
    bool owned = true;
    // work buffers of apf_csr_minres_solve, allocated on its first call
    void* work = nullptr;
    size_t work_bytes = 0;
};

struct apf_space {
    apf::Space S;
    apf::RefTables T;
    bool has_pattern = false;
    apf::MGHierarchyHost H;   // nested coarser spatial meshes for the multigrid preconditioner
    bool mg_ready = false;
};

namespace apf {

struct HipFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// temporal tables, Nt <= 5 (p <= 4), passed by value to kernels
struct TimeTables {
    double Mt[25];
    double Dt[25];     // Dt[i][j] = int l_i' l_j
    double Mtinv[25];
    double mt[5];
};

struct MinresScalars {
    double beta, beta1, oldb, alfa, dbar, epsln, oldeps, delta, gamma, cs, sn, phi, phibar;
    int itn;
};

// Per-cell affine metric (restates CellGeometry<Space,dim>::make,
// src/finite_element/geometry/cell_geometry.hpp:161-207):
//   det = |det J|,  G = det J^-1 M J^-T  (stiffness metric),
//   H = J^T M^-1 J / det  (contravariant-Piola mass metric),  Jm = J.
struct CellMetric {
    double det;
    double G[4];
    double H[4];
    double J[4];
};

__host__ __device__ inline CellMetric cell_metric(int dim, int c, const double* verts, const int* cells,
                                                  const double* Mcell)
{
    CellMetric g{};
    if (dim == 1) {
        const double h = verts[cells[2 * c + 1]] - verts[cells[2 * c]];
        const double ah = h < 0 ? -h : h;
        const double m = Mcell[c];
        g.det = ah;
        g.G[0] = m / ah;
        g.H[0] = ah / m;
        g.J[0] = h;
        return g;
    }
    const int v0 = cells[3 * c], v1 = cells[3 * c + 1], v2 = cells[3 * c + 2];
    const double J00 = verts[2 * v1] - verts[2 * v0], J01 = verts[2 * v2] - verts[2 * v0];
    const double J10 = verts[2 * v1 + 1] - verts[2 * v0 + 1], J11 = verts[2 * v2 + 1] - verts[2 * v0 + 1];
    const double det = J00 * J11 - J01 * J10;
    const double ad = det < 0 ? -det : det;
    const double i00 = J11 / det, i01 = -J01 / det, i10 = -J10 / det, i11 = J00 / det;  // J^-1
    const double* M = Mcell + 4 * c;
    // G = ad * Ji M Ji^T
    const double a00 = i00 * M[0] + i01 * M[2], a01 = i00 * M[1] + i01 * M[3];
    const double a10 = i10 * M[0] + i11 * M[2], a11 = i10 * M[1] + i11 * M[3];
    g.G[0] = ad * (a00 * i00 + a01 * i01);
    g.G[1] = ad * (a00 * i10 + a01 * i11);
    g.G[2] = ad * (a10 * i00 + a11 * i01);
    g.G[3] = ad * (a10 * i10 + a11 * i11);
    // H = J^T M^-1 J / ad
    const double md = M[0] * M[3] - M[1] * M[2];
    const double n00 = M[3] / md, n01 = -M[1] / md, n10 = -M[2] / md, n11 = M[0] / md;
    const double b00 = n00 * J00 + n01 * J10, b01 = n00 * J01 + n01 * J11;
    const double b10 = n10 * J00 + n11 * J10, b11 = n10 * J01 + n11 * J11;
    g.H[0] = (J00 * b00 + J10 * b10) / ad;
    g.H[1] = (J00 * b01 + J10 * b11) / ad;
    g.H[2] = (J01 * b00 + J11 * b10) / ad;
    g.H[3] = (J01 * b01 + J11 * b11) / ad;
    g.det = ad;
    g.J[0] = J00; g.J[1] = J01; g.J[2] = J10; g.J[3] = J11;
    return g;
}

struct EstimatorBuffers {
    // device tables (estimator)
    double* RTab = nullptr;
    double* Bhat = nullptr;
    double* qbar = nullptr;
    double* R = nullptr;
    double* Pv = nullptr;
    double* Q = nullptr;
    double* E = nullptr;
    double* Phi = nullptr;
    double* Khat = nullptr;
    double* Mhat = nullptr;
    double* dbary = nullptr;
    double* Wt = nullptr;     // [Nt][Nt] int l_l L_k
    double* Wdt = nullptr;    // [Nt][Nt] int l_l' L_k
    double* Mdd = nullptr;
    double* Lnorm_ = nullptr;
    double* Mt = nullptr;
    double* Igrad = nullptr;
    // patch topology
    int* p_cell_ptr = nullptr;
    int* p_cells = nullptr;
    int* p_lv = nullptr;
    int* p_fmap = nullptr;
    int8_t* p_fsign = nullptr;
    int* p_nflux = nullptr;
    int* p_nscal = nullptr;
    uint8_t* p_boundary = nullptr;
    int* order = nullptr;       // patch processing order (large first)
    int* cell_patch_pos = nullptr;
    // per (space-time cell, local vertex) flux contributions [ncell][dim+1][Nf][Nt]
    double* sig = nullptr;
    // per space-time cell outputs
    double* flux = nullptr;
    double* resid = nullptr;
    double* recon = nullptr;
    double* lam2 = nullptr;
    double* trace = nullptr;
    double* patch_res = nullptr;  // per patch max residual
    double* lt = nullptr;         // slab-reconstructed lambda
    double* sums = nullptr;
    int* extra_rounds = nullptr;  // refinement rounds beyond refine_steps (all blocks)
    int* tune_extra = nullptr;    // the same count for the column-blocking trial launches (discarded)
    // patch slots in processing order (largest first) and size buckets
    int* s_n = nullptr;
    int* s_np = nullptr;
    int64_t* s_off2 = nullptr;
    int64_t* s_off1 = nullptr;
    double* Kd = nullptr;    // FP64 patch matrices, np x np per patch
    double* Ki = nullptr;    // FP64 inverses (and block-elimination workspace)
    float* Ki32 = nullptr;   // FP32 inverses of the mixed path
    double* RHS = nullptr;   // right-hand sides, np x (slabs * Nt) per patch
    // X / R tiles of the patch solves whose 16-column tiles exceed the LDS
    // (one [np][16 CT] FP64 + FT pair per workgroup of the largest such bucket)
    double* solve_ws = nullptr;
    int64_t solve_ws_doubles = 0;
    struct Bucket {
        int q0, count, np, nc_max;
        int blk_max;  // largest flux block or Schur block order in the bucket
        bool big;     // factored by the blocked multi-launch Gauss-Jordan
    };
    std::vector<Bucket> buckets;
    int nmax = 0;
    // patches of the big buckets: slot, workspace offset (doubles), workspace
    int* big_q = nullptr;
    int64_t* big_off = nullptr;
    double* big_ws = nullptr;
    int big_count = 0, big_npmax = 0;
    // full-pivot retry of patches that miss the acceptance: slot list, count, workspace
    int* fp_list = nullptr;
    int* fp_count = nullptr;
    double* fp_ws = nullptr;
    int* fp_perm = nullptr;
    int64_t fp_ws_doubles = 0;
};

struct System {
    apf_ctx* ctx = nullptr;
    const Space* space = nullptr;
    const RefTables* tables = nullptr;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    TimeTables tt{};
    std::vector<double> Mhost;
    double ell = 0.0;
    // mesh / numbering
    double* d_verts = nullptr;
    int* d_cells = nullptr;
    double* d_times = nullptr;
    int* d_cell_free = nullptr;
    int16_t* d_rank = nullptr;
    int* d_sptr = nullptr;
    int* d_scol = nullptr;
    double* d_Mcell = nullptr;
    double* d_u0b = nullptr;
    // tables
    double* d_Khat = nullptr;
    double* d_Mhat = nullptr;
    double* d_mhat = nullptr;
    double* d_bw = nullptr;
    double* d_bphi = nullptr;
    // spatial matrices
    double* d_Ksp = nullptr;
    double* d_Msp = nullptr;
    double* d_msp = nullptr;
    double* d_gsp = nullptr;
    // K
    int64_t* d_rowptr = nullptr;
    int* d_col = nullptr;
    double* d_val = nullptr;
    double* d_b = nullptr;
    double* d_x = nullptr;
    double* d_vec[8] = {};
    int nparts = 0;
    double* d_part = nullptr;
    int64_t part_cap = 0;  // doubles in d_part (at least 4 nparts; grown to the multigrid's transpose tiles)
    MinresScalars* d_scal = nullptr;
    double* d_res = nullptr;
    double* d_cg = nullptr;  // scalars of the slab reconstruction's conjugate gradients
    hipStream_t stream2 = nullptr;  // the context's second stream (factor beside the slab solve)
    std::set<int> est_done;         // precisions with a completed estimate (tiles chosen)
    // pinned host mailbox: a kernel posts a few doubles and a sequence number, the host
    // polls it (cheaper than a device-to-host copy and a stream synchronisation)
    struct Mailbox {
        double v[8];
        unsigned long long seq;
    };
    Mailbox* h_mail = nullptr;  // host address
    Mailbox* d_mail = nullptr;  // the same memory as the device addresses it
    unsigned long long mail_seq = 0;
    // post n <= 8 doubles from device memory and wait for them on the host
    void post_and_wait(const double* src, int n, double* out);
    void ensure_mailbox();
    int2* d_kron_halo = nullptr;  // neighbour range of each block of the matrix-free product
    int kron_span_cap = 0;
    // the spatial K and M in ELL form ([entry j][free node s], rows padded with zero entries
    // on their own diagonal column) for the matrix-free kernels' coalesced row loads
    int ell_w = 0;
    int* d_ell_col = nullptr;
    double* d_ell_K = nullptr;
    double* d_ell_M = nullptr;
    void ell_refresh();  // values from d_Ksp / d_Msp (after each assembly)
    double* d_kdiag = nullptr;
    double* d_band = nullptr;
    float* d_band32 = nullptr;
    EstimatorBuffers est;
    const MGHierarchyHost* H = nullptr;
    std::unique_ptr<MGDevice<double>> mg64;
    std::unique_ptr<MGDevice<float>> mg32;
    uint64_t assembly_version = 0, mg64_version = 0, mg32_version = 0;
    double t_mg_setup_s = 0.0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t evp[4] = {nullptr, nullptr, nullptr, nullptr};  // estimator phase boundaries
    // patch-solve column blocking chosen by timing, per (padded order, bytes of the inverse's type)
    std::map<std::pair<int, int>, int> solve_ct;
    // patch right-hand sides: slabs per workgroup chosen by timing, per padded order
    std::map<int, int> rhs_sg;
    // mixed path: buckets (by padded order) that run the FP64 factor and solve because that
    // measured faster than the FP32 factor with refinement
    std::map<int, bool> mixed_fp64;
    // G-estimator on this (degree p+1) system: nodal interpolation from the degree-p
    // space on the same mesh (spatial CSR over free nodes, time matrix [p+2][p+1]),
    // built on the first call (g_src: the space it was built from), and two Y-sized
    // vectors (I lambda, lambda_G)
    const Space* g_src = nullptr;
    int* d_g_ptr = nullptr;
    int* d_g_col = nullptr;
    double* d_g_val = nullptr;
    double* d_g_It = nullptr;
    double* d_g_il = nullptr;
    double* d_g_lt = nullptr;
    float t_assembly_ms = 0.f, t_patch_ms = 0.f, t_diag_ms = 0.f, t_spmv_ms = 0.f;
    float t_factor_ms = 0.f, t_rhs_ms = 0.f, t_solve_ms = 0.f;
    bool assembled = false;
    bool use_graph = true;   // capture the MINRES iteration in a hipGraph
    // a system that only runs slab solves (the general estimator's slab groups): its multigrid
    // serves the Y columns alone, so the coarsest inverses are formed for shift 0 only
    bool slab_only = false;

    void spmv(const double* x, double* y);
    void spmv_rows(const double* x, double* y, int64_t rows);
    // y[0:nY) = A_y x[0:nY) without reading K (tensor mesh: K_sp (x) h_k Mt per slab)
    void ay_apply(const double* x, double* y);
    void kron_apply(const double* x, double* y);  // y = K x from the Kronecker factors
    void kron_prepare();
};

void estimator_alloc(System* sys);
void mg_prepare(System* sys, bool mg32);
// stop reason: APF_SLAB_STOP_TOL (relative residual <= tol), APF_SLAB_STOP_FLOOR (a restart
// gained less than 10%: the FP64 round-off floor), APF_SLAB_STOP_MAX_IT
void slab_reconstruct(System* sys, double* lt, double tol, int max_it, int* its, double* rel, int* stop);
// host set-up of the slab reconstruction (multigrid hierarchy, scalars, mailbox); idempotent
void slab_prepare(System* sys);

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            throw ::apf::HipFailure(std::string(#expr) + ": " + hipGetErrorString(_e));  \
    } while (0)

template <typename T>
inline T* dalloc(System* sys, size_t count)
{
    void* p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc(&p, count * sizeof(T)));
    sys->allocs.push_back(p);
    return static_cast<T*>(p);
}

}  // namespace apf

struct apf_system : apf::System {};
