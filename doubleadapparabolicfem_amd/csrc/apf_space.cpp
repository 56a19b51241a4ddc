// Host-side integer work: numbering, adjacency, K pattern, patch topology.
#include "apf_space.hpp"

#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>

#include "apf_tables.hpp"

namespace apf {

static const int TRI_E[3][2] = {{1, 2}, {2, 0}, {0, 1}};  // edge e opposite vertex e

namespace {
struct Tuple { int a, b, c; };
std::vector<Tuple> triangle_tuples(int p)
{
    std::vector<Tuple> t;
    for (int c = 0; c <= p; ++c)
        for (int b = 0; b <= p - c; ++b) t.push_back({p - b - c, b, c});
    return t;
}
}  // namespace

void build_space(Space& S, const TensorMesh& mesh, int p)
{
    if (mesh.dim != 1 && mesh.dim != 2) throw std::runtime_error("build_space: dim must be 1 or 2");
    if (p < 1 || p > 4) throw std::runtime_error("build_space: p must be in 1..4");
    if (mesh.m() < 1) throw std::runtime_error("build_space: need at least one time interval");
    for (int k = 0; k < mesh.m(); ++k)
        if (!(mesh.times[k + 1] > mesh.times[k])) throw std::runtime_error("build_space: times must increase");
    S = Space();
    S.dim = mesh.dim;
    S.p = p;
    S.mesh = mesh;
    const int d = mesh.dim, nv = mesh.nv(), ns = mesh.ns();
    S.Ns = n_space_basis(d, p);
    S.Nt = p + 1;
    const int Ns = S.Ns;
    for (int c = 0; c < ns; ++c)
        for (int v = 0; v <= d; ++v) {
            const int id = mesh.cells[c * (d + 1) + v];
            if (id < 0 || id >= nv) throw std::runtime_error("build_space: cell vertex id out of range");
        }
    S.cell_nodes.assign(static_cast<size_t>(ns) * Ns, -1);
    if (d == 1) {
        // boundary vertices are those of exactly one cell: the two ends of an
        // interval, or of every segment when the mesh holds several disjoint ones
        std::vector<int> deg(nv, 0);
        for (int c = 0; c < ns; ++c) {
            ++deg[mesh.cells[c * 2]];
            ++deg[mesh.cells[c * 2 + 1]];
        }
        S.vertex_boundary.assign(nv, 0);
        for (int i = 0; i < nv; ++i) S.vertex_boundary[i] = deg[i] == 1;
        S.n_nodes = nv + ns * (p - 1);
        S.node_boundary.assign(S.n_nodes, 0);
        for (int i = 0; i < nv; ++i) S.node_boundary[i] = S.vertex_boundary[i];
        for (int c = 0; c < ns; ++c) {
            S.cell_nodes[c * Ns + 0] = mesh.cells[c * 2 + 0];
            S.cell_nodes[c * Ns + p] = mesh.cells[c * 2 + 1];
            for (int j = 1; j < p; ++j) S.cell_nodes[c * Ns + j] = nv + c * (p - 1) + (j - 1);
        }
    } else {
        std::map<std::pair<int, int>, int> ecount;
        for (int c = 0; c < ns; ++c)
            for (int e = 0; e < 3; ++e) {
                const int a = mesh.cells[c * 3 + TRI_E[e][0]], b = mesh.cells[c * 3 + TRI_E[e][1]];
                ecount[{std::min(a, b), std::max(a, b)}] += 1;
            }
        std::map<std::pair<int, int>, int> eid;
        S.edges.clear();
        S.edge_boundary.clear();
        S.vertex_boundary.assign(nv, 0);
        for (const auto& kv : ecount) {
            eid[kv.first] = static_cast<int>(S.edge_boundary.size());
            S.edges.push_back(kv.first.first);
            S.edges.push_back(kv.first.second);
            const bool bd = kv.second == 1;
            S.edge_boundary.push_back(bd ? 1 : 0);
            if (bd) { S.vertex_boundary[kv.first.first] = 1; S.vertex_boundary[kv.first.second] = 1; }
            if (kv.second > 2) throw std::runtime_error("build_space: non-manifold edge");
        }
        const int ne = static_cast<int>(S.edge_boundary.size());
        const auto tuples = triangle_tuples(p);
        std::vector<int> interior_pos(tuples.size(), -1);
        int ni = 0;
        for (size_t t = 0; t < tuples.size(); ++t)
            if (tuples[t].a > 0 && tuples[t].b > 0 && tuples[t].c > 0) interior_pos[t] = ni++;
        S.n_nodes = nv + ne * (p - 1) + ns * ni;
        S.node_boundary.assign(S.n_nodes, 0);
        for (int i = 0; i < nv; ++i) S.node_boundary[i] = S.vertex_boundary[i];
        for (int c = 0; c < ns; ++c) {
            const int* V = &mesh.cells[c * 3];
            for (size_t t = 0; t < tuples.size(); ++t) {
                const int w[3] = {tuples[t].a, tuples[t].b, tuples[t].c};
                int nz[3], cnt = 0;
                for (int i = 0; i < 3; ++i)
                    if (w[i] > 0) nz[cnt++] = i;
                int nid;
                if (cnt == 1) {
                    nid = V[nz[0]];
                } else if (cnt == 2) {
                    const int vi = V[nz[0]], vj = V[nz[1]];
                    const int e = eid[{std::min(vi, vj), std::max(vi, vj)}];
                    const int off = (vi < vj) ? w[nz[1]] : w[nz[0]];
                    nid = nv + e * (p - 1) + (off - 1);
                    S.node_boundary[nid] = S.edge_boundary[e];
                } else {
                    nid = nv + ne * (p - 1) + c * ni + interior_pos[t];
                }
                S.cell_nodes[c * Ns + static_cast<int>(t)] = nid;
            }
        }
    }
    // free ids by first appearance over (cell, local node): neighbours get
    // nearby ids, which keeps the x gathers of the SpMV cache-local
    S.free_index.assign(S.n_nodes, -1);
    S.nf = 0;
    for (size_t e = 0; e < S.cell_nodes.size(); ++e) {
        const int nid = S.cell_nodes[e];
        if (!S.node_boundary[nid] && S.free_index[nid] < 0) S.free_index[nid] = S.nf++;
    }
    S.cell_free.resize(S.cell_nodes.size());
    for (size_t i = 0; i < S.cell_nodes.size(); ++i) S.cell_free[i] = S.free_index[S.cell_nodes[i]];

    // spatial adjacency of free nodes
    std::vector<std::vector<int>> adj(S.nf);
    for (int c = 0; c < ns; ++c)
        for (int i = 0; i < Ns; ++i) {
            const int fi = S.cell_free[c * Ns + i];
            if (fi < 0) continue;
            for (int j = 0; j < Ns; ++j) {
                const int fj = S.cell_free[c * Ns + j];
                if (fj >= 0) adj[fi].push_back(fj);
            }
        }
    S.sptr.assign(S.nf + 1, 0);
    for (int i = 0; i < S.nf; ++i) {
        auto& a = adj[i];
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
        S.sptr[i + 1] = S.sptr[i] + static_cast<int>(a.size());
    }
    S.scol.resize(S.sptr[S.nf]);
    for (int i = 0; i < S.nf; ++i) std::copy(adj[i].begin(), adj[i].end(), S.scol.begin() + S.sptr[i]);
    S.rank.assign(static_cast<size_t>(ns) * Ns * Ns, -1);
    for (int c = 0; c < ns; ++c)
        for (int i = 0; i < Ns; ++i) {
            const int fi = S.cell_free[c * Ns + i];
            if (fi < 0) continue;
            for (int j = 0; j < Ns; ++j) {
                const int fj = S.cell_free[c * Ns + j];
                if (fj < 0) continue;
                const int* b = &S.scol[S.sptr[fi]];
                const int* e = &S.scol[S.sptr[fi + 1]];
                const int pos = static_cast<int>(std::lower_bound(b, e, fj) - b);
                if (pos > 32767) throw std::runtime_error("build_space: spatial degree too large");
                S.rank[(static_cast<size_t>(c) * Ns + i) * Ns + j] = static_cast<int16_t>(pos);
            }
        }
    const int64_t m = mesh.m();
    S.nY = m * S.Nt * S.nf;
    S.nX = (m * p + 1) * static_cast<int64_t>(S.nf);
    S.n = S.nY + S.nX;

    // ---- vertex patches (shared by every slab)
    PatchTopology& P = S.patches;
    const int Nf = n_flux_local(d, p), Nfm = n_facet_moments(d, p), Nfc = n_facets(d);
    const int nint = Nf - Nfc * Nfm, Nq = n_scalar_local(d, p);
    std::vector<std::vector<int>> cells_of(nv);
    for (int c = 0; c < ns; ++c)
        for (int v = 0; v <= d; ++v) cells_of[mesh.cells[c * (d + 1) + v]].push_back(c);
    P.n_patches = nv;
    P.vertex.resize(nv);
    P.boundary.resize(nv);
    P.cell_ptr.assign(nv + 1, 0);
    P.n_flux.resize(nv);
    P.n_scalar.resize(nv);
    P.cells.clear();
    P.local_vertex.clear();
    P.cell_patch_pos.assign(static_cast<size_t>(ns) * (d + 1), -1);
    for (int a = 0; a < nv; ++a) {
        auto& cl = cells_of[a];
        std::sort(cl.begin(), cl.end());
        P.vertex[a] = a;
        P.boundary[a] = S.vertex_boundary[a];
        P.cell_ptr[a + 1] = P.cell_ptr[a] + static_cast<int>(cl.size());
    }
    const int total = P.cell_ptr[nv];
    P.cells.resize(total);
    P.local_vertex.resize(total);
    P.fmap.assign(static_cast<size_t>(total) * Nf, -1);
    P.fsign.assign(static_cast<size_t>(total) * Nf, 0);
    for (int a = 0; a < nv; ++a) {
        const auto& cl = cells_of[a];
        const int base = P.cell_ptr[a];
        const bool is_b = S.vertex_boundary[a] != 0;
        // facets of the patch and their multiplicity inside the patch
        std::map<std::pair<int, int>, int> inpatch;
        for (int ci = 0; ci < static_cast<int>(cl.size()); ++ci) {
            const int c = cl[ci];
            const int* V = &mesh.cells[c * (d + 1)];
            P.cells[base + ci] = c;
            int lv = -1;
            for (int v = 0; v <= d; ++v)
                if (V[v] == a) lv = v;
            P.local_vertex[base + ci] = lv;
            P.cell_patch_pos[static_cast<size_t>(c) * (d + 1) + lv] = ci;
            for (int f = 0; f < Nfc; ++f) {
                int A, B;
                if (d == 1) { A = B = V[f]; }
                else { A = V[TRI_E[f][0]]; B = V[TRI_E[f][1]]; }
                inpatch[{std::min(A, B), std::max(A, B)}] += 1;
            }
        }
        std::map<std::pair<int, int>, int> slot;
        for (const auto& kv : inpatch) {
            bool on_boundary;
            if (d == 1) on_boundary = S.vertex_boundary[kv.first.first] != 0;
            else {
                // mesh boundary edge: count 1 in the whole mesh
                on_boundary = false;
                const int lo = kv.first.first, hi = kv.first.second;
                // find edge id by binary search on the sorted edge list
                int l = 0, r = static_cast<int>(S.edge_boundary.size()) - 1;
                while (l <= r) {
                    const int mid = (l + r) / 2;
                    const int el = S.edges[2 * mid], eh = S.edges[2 * mid + 1];
                    if (el == lo && eh == hi) { on_boundary = S.edge_boundary[mid] != 0; break; }
                    if (el < lo || (el == lo && eh < hi)) l = mid + 1; else r = mid - 1;
                }
            }
            if (kv.second == 2 || (is_b && on_boundary)) {
                const int s = static_cast<int>(slot.size());
                slot[kv.first] = s;
            }
        }
        int off = static_cast<int>(slot.size()) * Nfm;
        for (int ci = 0; ci < static_cast<int>(cl.size()); ++ci) {
            const int c = cl[ci];
            const int* V = &mesh.cells[c * (d + 1)];
            int* fm = &P.fmap[static_cast<size_t>(base + ci) * Nf];
            int8_t* fs = &P.fsign[static_cast<size_t>(base + ci) * Nf];
            for (int f = 0; f < Nfc; ++f) {
                int A, B;
                if (d == 1) { A = B = V[f]; }
                else { A = V[TRI_E[f][0]]; B = V[TRI_E[f][1]]; }
                auto it = slot.find({std::min(A, B), std::max(A, B)});
                if (it == slot.end()) continue;
                for (int k = 0; k < Nfm; ++k) {
                    fm[f * Nfm + k] = it->second * Nfm + k;
                    int sg = 1;
                    if (d == 2 && A > B) sg = (k % 2 == 0) ? -1 : 1;  // -(-1)^k
                    fs[f * Nfm + k] = static_cast<int8_t>(sg);
                }
            }
            for (int k = 0; k < nint; ++k) {
                fm[Nfc * Nfm + k] = off++;
                fs[Nfc * Nfm + k] = 1;
            }
        }
        P.n_flux[a] = off;
        P.n_scalar[a] = static_cast<int>(cl.size()) * Nq;
    }
}

void build_pattern(Space& S)
{
    const int p = S.p, Nt = S.Nt, nf = S.nf;
    const int m = S.mesh.m();
    S.rowptr.assign(S.n + 1, 0);
    // row lengths
    for (int64_t r = 0; r < S.n; ++r) {
        int64_t len;
        if (r < S.nY) {
            const int s = static_cast<int>(r % nf);
            const int deg = S.sptr[s + 1] - S.sptr[s];
            len = 2LL * Nt * deg;
        } else {
            const int64_t x = r - S.nY;
            const int tn = static_cast<int>(x / nf);
            const int s = static_cast<int>(x % nf);
            const int deg = S.sptr[s + 1] - S.sptr[s];
            int k0, nk;
            time_node_intervals(tn, p, m, k0, nk);
            // C block: full local X x X coupling of the t = 0 cells, as registered
            // by the two-pass pattern (two_pass_full_saddle_2d.hpp:780-789)
            len = static_cast<int64_t>(nk) * Nt * deg + (tn <= p ? Nt * deg : 0);
        }
        S.rowptr[r + 1] = S.rowptr[r] + len;
    }
    S.nnz = S.rowptr[S.n];
    if (S.n > 2147483647LL) throw std::runtime_error("build_pattern: more than 2^31 unknowns");
}

void pattern_columns(const Space& S, int* col)
{
    const int p = S.p, Nt = S.Nt, nf = S.nf;
    const int m = S.mesh.m();
    for (int64_t r = 0; r < S.n; ++r) {
        int* out = &col[S.rowptr[r]];
        if (r < S.nY) {
            const int s = static_cast<int>(r % nf);
            const int k = static_cast<int>(r / nf) / Nt;
            const int* a = &S.scol[S.sptr[s]];
            const int deg = S.sptr[s + 1] - S.sptr[s];
            for (int l = 0; l < Nt; ++l)
                for (int j = 0; j < deg; ++j) *out++ = static_cast<int>((static_cast<int64_t>(k) * Nt + l) * nf + a[j]);
            for (int it = 0; it < Nt; ++it)
                for (int j = 0; j < deg; ++j)
                    *out++ = static_cast<int>(S.nY + (static_cast<int64_t>(k) * p + it) * nf + a[j]);
        } else {
            const int64_t x = r - S.nY;
            const int tn = static_cast<int>(x / nf);
            const int s = static_cast<int>(x % nf);
            const int* a = &S.scol[S.sptr[s]];
            const int deg = S.sptr[s + 1] - S.sptr[s];
            int k0, nk;
            time_node_intervals(tn, p, m, k0, nk);
            for (int kk = 0; kk < nk; ++kk)
                for (int l = 0; l < Nt; ++l)
                    for (int j = 0; j < deg; ++j)
                        *out++ = static_cast<int>((static_cast<int64_t>(k0 + kk) * Nt + l) * nf + a[j]);
            if (tn <= p)
                for (int tq = 0; tq < Nt; ++tq)
                    for (int j = 0; j < deg; ++j)
                        *out++ = static_cast<int>(S.nY + static_cast<int64_t>(tq) * nf + a[j]);
        }
    }
}

}  // namespace apf
