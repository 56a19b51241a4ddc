// Host-side integer work for a tensor space-time mesh: DOF numbering, the
// CSR pattern of the saddle-point matrix K and the vertex-patch topology.
//
// The reference keeps an adaptive prism/rectangle tree and numbers DOFs on
// its active antichain (src/finite_element/fespace/dofs/dof_distribution_2d.hpp,
// src/finite_element/fespace/fespace.hpp:111-758).  This build specialises
// to the tensor meshes  T_h x {t_0 < ... < t_m}  of the hot-path workloads
// and uses its own (documented) numbering:
//   spatial nodes: vertices, then (p-1) nodes per edge (edges sorted by
//     (min, max) vertex id, nodes ordered from the min vertex), then interior
//     nodes per triangle in reference tuple order; free = not on dOmega,
//     numbered by first appearance over (cell, local node);
//   Y dof (k, l, s) = (k (p+1) + l) nf + s          (dG in time)
//   X dof (tn, s)   = nY + tn nf + s,  tn = k p + l (continuous in time)
//   K unknown order [lambda (Y); u (X)] as in
//   src/linear_algebra/system/saddle_point_system.hpp:105-118.
#pragma once

#include <cstdint>
#include <vector>

#if defined(__HIPCC__)
#define APF_HD __host__ __device__
#else
#define APF_HD
#endif

namespace apf {

struct TensorMesh {
    int dim = 2;                      // spatial dimension, 1 or 2
    std::vector<double> vertices;     // [nv * dim]
    std::vector<int> cells;           // [ns * (dim+1)], counter-clockwise in 2D
    std::vector<double> times;        // [m + 1], strictly increasing
    int nv() const { return static_cast<int>(vertices.size()) / dim; }
    int ns() const { return static_cast<int>(cells.size()) / (dim + 1); }
    int m() const { return static_cast<int>(times.size()) - 1; }
};

struct PatchTopology {
    // one patch per spatial vertex (shared by every slab of a tensor mesh)
    int n_patches = 0;
    std::vector<int> vertex;          // [np] spatial vertex id
    std::vector<uint8_t> boundary;    // [np]
    std::vector<int> cell_ptr;        // [np+1] into cell lists
    std::vector<int> cells;           // spatial cell ids, sorted
    std::vector<int> local_vertex;    // local index of the patch vertex in each cell
    std::vector<int> n_flux;          // [np] spatial flux dofs
    std::vector<int> n_scalar;        // [np] spatial scalar dofs
    // per (patch, cell) local flux dof -> patch dof (or -1), sign
    std::vector<int> fmap;            // [cell_ptr[np]][Nf]
    std::vector<int8_t> fsign;        // [cell_ptr[np]][Nf]
    // per spatial cell and local vertex v: position of the cell inside patch
    // of vertex v (index into the patch's cell list)
    std::vector<int> cell_patch_pos;  // [ns][dim+1]
};

struct Space {
    int dim = 2, p = 1;
    TensorMesh mesh;
    int Ns = 0, Nt = 0;
    // spatial nodes
    int n_nodes = 0, nf = 0;
    std::vector<int> cell_nodes;      // [ns][Ns]
    std::vector<int> cell_free;       // [ns][Ns] free id or -1
    std::vector<uint8_t> node_boundary;
    std::vector<int> free_index;      // [n_nodes]
    std::vector<int> edges;           // [ne][2] (2D)
    std::vector<uint8_t> edge_boundary;
    std::vector<uint8_t> vertex_boundary;
    // spatial adjacency of free nodes (CSR, sorted, includes the diagonal)
    std::vector<int> sptr, scol;
    // per spatial cell: position of free node j in adj(free node i), -1 if either is Dirichlet
    std::vector<int16_t> rank;        // [ns][Ns][Ns]
    // global sizes and pattern of K
    int64_t nY = 0, nX = 0, n = 0, nnz = 0;
    std::vector<int64_t> rowptr;      // [n+1]; the columns are generated where needed (pattern_columns, or on the device)
    PatchTopology patches;
};

// Build the spatial numbering, adjacency and patch topology (no K pattern).
void build_space(Space& S, const TensorMesh& mesh, int p);
// Build the CSR pattern of K = [A_y B^T; B C] (full element-coupling pattern).
void build_pattern(Space& S);
// column indices of K's pattern in row order, out[nnz]
void pattern_columns(const Space& S, int* out);
// Row-structure helper: number of intervals touching time node tn and the first one.
APF_HD inline void time_node_intervals(int tn, int p, int m, int& k0, int& nk)
{
    if (tn % p != 0) { k0 = tn / p; nk = 1; return; }
    const int k = tn / p;
    k0 = (k > 0) ? k - 1 : 0;
    nk = (k > 0 && k < m) ? 2 : 1;
}

}  // namespace apf
