// Device state of the space-time multigrid preconditioner.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "apf_mg.hpp"

namespace apf {

struct System;

template <typename R>
struct MGDevice {
    struct Level {
        int nf = 0;
        double lmax = 1.0;
        int* ptr = nullptr;
        int* col = nullptr;
        R* K = nullptr;
        R* M = nullptr;
        R* dK = nullptr;
        R* dM = nullptr;
        R* X = nullptr;   // solution block
        R* B = nullptr;   // right-hand side block
        R* Rr = nullptr;  // residual
        R* D = nullptr;   // Chebyshev direction
        R* Tm = nullptr;  // second Chebyshev direction (the fused step alternates D and Tm)
        int64_t nnz = 0, Pnnz = 0, Rnnz = 0;  // value counts of K / M, P, R
        int Pn = 0, Rn = 0;
        int* Pptr = nullptr;
        int* Pcol = nullptr;
        R* Pval = nullptr;
        int* Rptr = nullptr;
        int* Rcol = nullptr;
        R* Rval = nullptr;
    };
    std::vector<Level> levels;
    int T = 0, nyc = 0, ncol = 0, ld = 0;
    int nu = 2;              // Chebyshev degree (pre and post)
    double cheb_ratio = 4.0; // smoothing interval [lmax / ratio, lmax]
    R* d_sig = nullptr;
    R* d_zero = nullptr;
    int* d_sidx = nullptr;
    int* d_zidx = nullptr;
    int nc_coarse = 0;
    int64_t blk_total = 0;
    R* d_cinv = nullptr;       // [shift][blk_total] block-diagonal coarsest inverses
    int* d_comp_start = nullptr;
    int* d_node_comp = nullptr;
    int64_t* d_blk_off = nullptr;
    double* d_W = nullptr;
    double* d_MtInvH = nullptr;
    bool ready = false;
    // fused coarse tail: levels tail_from..L-1 run in one launch (k_mg_tail), one
    // workgroup per column with the level vectors in LDS; -1 = off
    int tail_from = -1;
    size_t tail_lds = 0;
    void* d_tail = nullptr;  // [L - tail_from] level descriptors on the device

    int64_t cinv_size = 0, W_size = 0, MtInvH_size = 0;

    void build(System* sys, const MGSetupHost& H, int nyc);
    void refresh(const MGSetupHost& H);
    // the values of another precision's hierarchy of the same space (built alike), converted
    // on the device: no host setup
    template <typename S>
    void refresh_from(const MGDevice<S>& src, hipStream_t st);
    void setup_tail(System* sys);
    // the block pointers of every level are offset by c0 columns (c0 = T: the Y columns only)
    void smooth(hipStream_t st, int l, int c0, int nc, const R* sig, bool zero_start, int deg);
    void vcycle(hipStream_t st, int l, int c0, int nc, const R* sig, const int* sidx, int deg);
    // y_only: the caller reads only z's Y part, so r's X part is not read, only the Y
    // columns are smoothed, z's X part is not written and the dot partials cover z_Y . r_Y
    // degree: Chebyshev steps per smoothing (0: nu)
    void apply(hipStream_t st, const Space& S, const double* r, double* z, double* part, int nparts,
               bool y_only = false, int degree = 0);
    // `reps` launches of one fine-level Chebyshev step over all columns (timing only: it
    // changes the level's work vectors); average ms and the algorithmic bytes of a launch
    void bench_step(hipStream_t st, int reps, double* ms, double* bytes);
    int tile_rows() const;
    int n_partials(int nf) const;  // dot partials written by apply()
};

}  // namespace apf
