// Host setup of the space-time multigrid preconditioner for MINRES.
//
// On a tensor mesh with time-independent M the saddle-point blocks are
// Kronecker sums (K, M spatial stiffness / mass; h_k, Mt, Dt time factors):
//   A_y = (+)_k K (x) h_k Mt
//   S   = B A_y^-1 B^T - C = (M K^-1 M) (x) T1 + K (x) T3 + M (x) e_T e_T^T
//   T1  = sum_k E_k Dt Mt^-1 Dt^T E_k^T / h_k,   T3 = sum_k h_k E_k Mt E_k^T
// (the two mixed terms telescope to the final-time trace).  With T1 W = T3 W Theta,
// W^T T3 W = I the Schur complement decouples over the time modes i into
// theta_i M K^-1 M + K, and
//   (K + theta M K^-1 M)^-1  ~  (K + sqrt(theta) M)^-1 K (K + sqrt(theta) M)^-1
// within a factor 2.  The preconditioner is therefore
//   P_Y = K^-1 (x) (h_k Mt)^-1,
//   P_X = (I (x) W) diag_i[(K + s_i M)^-1 K (K + s_i M)^-1] (I (x) W)^T,  s_i = sqrt(theta_i)
// with every spatial inverse replaced by one symmetric multigrid V-cycle on
// nested spatial meshes (Chebyshev-Jacobi smoothing, Galerkin coarse
// operators, dense inverses on the coarsest mesh).  The reference factors A_y
// and its graph-norm Schur approximation exactly with sparse Cholesky
// (src/linear_algebra/eigen_backend/preconditioners/parabolic_graph_norm_minres.hpp:76-88);
// this build keeps the block-diagonal structure and replaces the factorisations
// by the spectrally equivalent operators above.
#pragma once

#include <vector>

#include "apf_space.hpp"

namespace apf {

struct CSR {
    int nrows = 0, ncols = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
};

struct MGLevelHost {
    int nf = 0;
    // spatial pattern (for level 0 it is the main space's adjacency)
    std::vector<int> sptr, scol;
    std::vector<double> K, M;     // values on (sptr, scol)
    std::vector<double> dK, dM;   // diagonals
    double lmax = 1.0;            // bound of lambda_max(D^-1 (K + s M)) for every s >= 0
    CSR P;                        // prolongation: this level <- next coarser level
    CSR R;                        // restriction = P^T
};

struct MGHierarchyHost {
    std::vector<Space> coarse;    // coarse spatial spaces (index l-1 for level l >= 1)
    std::vector<CSR> P;           // P[l]: level l (fine) <- level l+1 (coarse), l = 0..L-2
};

// Nested-mesh prolongations, built once per space (integer + geometry work).
void build_mg_prolongations(const Space& fine, MGHierarchyHost& H);

struct MGSetupHost {
    std::vector<MGLevelHost> levels;   // 0 = finest
    // time factors
    int T = 0;                          // mp + 1 X time nodes
    std::vector<double> W;             // [T][T] generalised eigenvectors (column i = mode i)
    std::vector<double> shift;          // [T] sqrt(theta_i)
    std::vector<double> MtInvH;        // [m][Nt][Nt] (h_k Mt)^-1
    // coarsest-level dense inverses for each distinct shift (0 and shift[i]),
    // one block per connected component of the coarsest graph (components are
    // contiguous node ranges, e.g. independent problems; else one block)
    std::vector<double> shifts_all;     // [nsft] distinct shifts, shifts_all[0] = 0
    std::vector<int> comp_start;        // [ncomp + 1] node range of each block
    std::vector<int> node_comp;         // [nfc] block of each node
    std::vector<int64_t> blk_off;       // [ncomp] offset of each block inside one shift
    int64_t blk_total = 0;              // sum of squared block sizes
    std::vector<double> coarse_inv;     // [nsft][blk_total]
};

// Galerkin coarse operators, smoother bounds, time eigen-decomposition,
// coarsest inverses.  Ksp/Msp are the finest spatial matrices on (S.sptr, S.scol).
// zero_shift_only: the coarsest inverses of shift 0 alone (the slab solve's Y columns); the
// other shifts' blocks stay zero.
void setup_mg(const Space& S, const MGHierarchyHost& H, const std::vector<double>& Ksp,
              const std::vector<double>& Msp, const std::vector<double>& Mt, const std::vector<double>& Dt,
              MGSetupHost& out, bool zero_shift_only = false);

// small dense symmetric eigen-decomposition (cyclic Jacobi), A = V diag(w) V^T
void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V);

}  // namespace apf
