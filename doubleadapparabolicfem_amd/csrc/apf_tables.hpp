// Reference-cell quadrature, bases and integral tables for the hot path.
//
// Everything here is computed once on the host in FP64 and uploaded; the GPU
// kernels contract these geometry-free tables with per-cell metric
// coefficients.  The rules are the ones the reference's runner uses
// (algorithm/runners/detail/runner_case_execution.hpp:29-37):
//   1+1D: QSpace = 2p+1 Gauss points, QTime = 2p+1 Gauss points;
//   2+1D: QSpace = 4p+1 (Duffy exactness degree), QTime = 2p+1 Gauss points.
#pragma once

#include <vector>

namespace apf {

struct Rule1D {
    std::vector<double> x, w;  // on [0,1]
};
struct Rule2D {
    std::vector<double> x, y, w;  // on conv{(0,0),(1,0),(0,1)}
};

// n-point Gauss-Legendre on [0,1] (Newton iteration on the three-term recurrence).
// Restates gauss_legendre_1d<Order> (src/quadrature/gauss_legendre_1d.hpp:157-172).
Rule1D gauss_legendre01(int n);
// Collapsed Gauss rule exact to total degree `degree`, order (degree+3)/2, u outer.
// Restates duffy_triangle<Degree> (src/quadrature/reference_triangle_duffy.hpp:24-61).
Rule2D duffy(int degree);

// Equidistant Lagrange basis on [0,1] (nodes i/p): values and d/ds.
void segment_lagrange(int p, double s, double* v, double* d);
// P_p Lagrange on the reference triangle, node order c outer / b inner
// (src/finite_element/nodes/triangle_equidistant.hpp:39-53).
void triangle_lagrange(int p, double x, double y, double* v, double* gx, double* gy);
// Shifted Legendre L_k(2s-1), k = 0..p, and d/ds.
void legendre01(int p, double s, double* v, double* d);

int n_space_basis(int dim, int p);   // Ns
int n_flux_local(int dim, int p);    // RT_p (2D) or P_{p+1} hats+bubbles (1D) per cell
int n_scalar_local(int dim, int p);  // P_p per cell
int n_facets(int dim);               // 3 edges / 2 end points
int n_facet_moments(int dim, int p); // p+1 / 1

// Local flux basis on the reference cell (moment-dual RT_p in 2D; in 1D the
// end-point hats then bubbles xi(1-xi)L_k(xi), k < p).  Values [N][dim], div [N].
struct FluxBasis {
    int dim = 0, p = 0, N = 0;
    std::vector<double> coef;  // 2D: monomial -> nodal coefficients [N][N]
    void eval(double x, double y, double* val, double* div) const;
};
FluxBasis make_flux_basis(int dim, int p);

// L2(reference cell)-orthonormal P_p scalar basis (2D: Gram-Schmidt of monomials;
// 1D: normalised shifted Legendre).
struct ScalarBasis {
    int dim = 0, p = 0, N = 0;
    std::vector<double> coef;  // [N mono][N basis]
    void eval(double x, double y, double* val) const;
};
ScalarBasis make_scalar_basis(int dim, int p);

// Geometry-free integral tables on the reference cell.  Layouts are row-major.
struct RefTables {
    int dim = 0, p = 0;
    int Ns = 0, Nt = 0, Nf = 0, Nq = 0, Nfc = 0, Nfm = 0;  // Nf: flux local, Nq: scalar local
    // main system
    std::vector<double> Khat;   // [dim][dim][Ns][Ns]  int d_a phi_i d_b phi_j
    std::vector<double> Mhat;   // [Ns][Ns]
    std::vector<double> mhat;   // [Ns]
    std::vector<double> Mt;     // [Nt][Nt]   int l_i l_j
    std::vector<double> Dt;     // [Nt][Nt]   int l_i' l_j
    std::vector<double> mt;     // [Nt]
    std::vector<double> Mdd;    // [Nt][Nt]   int l_i' l_j'
    // bottom (t = 0) rule: points [nb][dim], weights, spatial basis values [nb][Ns]
    int nb = 0;
    std::vector<double> bx, bw, bphi;
    // estimator
    std::vector<double> RTab;   // [dim][dim][Nf][Nf]   int s_i,a s_j,b
    std::vector<double> Bhat;   // [Nq][Nf]             int q_i div s_j
    std::vector<double> qbar;   // [Nq]                 int q_i
    std::vector<double> R;      // [dim+1][Nf][Ns]      int lam_v s_i . grad phi_m
    std::vector<double> Pv;     // [dim+1][Nq]          int lam_v q_i
    std::vector<double> Q;      // [dim+1][Nq][Ns]      int lam_v phi_m q_i
    std::vector<double> E;      // [dim][Nq][Ns]        int d_b phi_m q_i
    std::vector<double> W;      // [Nt][Nt]             int l_l L_k
    std::vector<double> Wd;     // [Nt][Nt]             int l_l' L_k
    std::vector<double> Lnorm;  // [Nt]                 int L_k^2 = 1/(2k+1)
    std::vector<double> dbary;  // [dim+1][dim]         grad of barycentric lam_v
    std::vector<double> Phi;    // [Nq][Ns]             int q_i phi_m
    std::vector<double> Igrad;  // [Nf][dim][dim][Ns]   flux-basis coefficients of e_a d_b phi_m
    // facet structure of the local flux basis: local dof = facet*Nfm + moment
    std::vector<int> facet_v;   // [Nfc][2] local vertices of facet (2D: counter-clockwise)
};
RefTables make_ref_tables(int dim, int p);

// small dense helpers (row-major)
bool invert_dense(int n, std::vector<double>& a);

}  // namespace apf
