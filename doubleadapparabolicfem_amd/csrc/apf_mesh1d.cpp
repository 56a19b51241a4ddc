// 1+1D adaptive space-time meshes and the X / Y finite-element spaces on them, with the
// reference's cell ids, vertex ids, refinement closure and DoF numbering.
//
// Restates, for rectangles [x0, x1] x [t0, t1]:
//  * the cell tree: root cell (mesh/initialization/root_cell.hpp:60-107), split policy
//    (mesh/refinement/split_policy.hpp:8-26: every 1+1D split is space-time), the four
//    children (left,bottom), (right,bottom), (left,top), (right,top) appended in that order
//    (mesh/refinement/refine_1d.hpp:221-310), midpoint vertices found or created through the
//    coordinate registry (mesh/detail/vertex_registry.hpp, key llround(1e12 x),
//    core/coord_key.hpp), boundary flags from the root's vertex ids (topology/boundary_1d.hpp);
//  * the active partition of an FE space: initialise from a cell list, remove by swap with
//    the last entry in ascending id order, append added ids in ascending order
//    (fespace/partition_view.hpp:197-311, 1014-1045);
//  * adjacency: faces grouped by their fixed vertex, each group sorted by (start, longer
//    first, cell, face), exact matches first, then the nearest containing open face
//    (hanging), unmatched boundary faces last (fespace/adjacency/adjacency_1d.hpp:126-394).
//    The groups live in a std::unordered_map<int, ...> reserved for 2 x (active cells), as
//    in the reference: the interface order, and with it the order in which the grading
//    closure refines cells, is that container's iteration order;
//  * FESpace::refine: marked active cells (duplicates and inactive ids dropped) split in
//    the given order, existing children reused, then recursive one-level grading closure on
//    hanging interfaces (fespace/refinement/refine_1d.hpp:79-321);
//  * distribute_dofs_1d (fespace/dofs/dof_distribution_1d.hpp:100-671): local occurrences
//    in active-cell order, Dirichlet elimination on the spatial boundary, hanging slaves
//    (spatial interfaces first, temporal ones for the space-time continuous X space),
//    conforming unions, vertex unification by coordinate (X only), free DoFs numbered by
//    first occurrence, constrained DoFs after them in (generation, cell, local) order with
//    masters expanded recursively through barycentric Lagrange interpolation
//    (basis/polynomials/segment_lagrange.hpp:60-136) and deduplicated by coordinate;
//  * build_cell_restriction (assembly/constrained_dofs.hpp:133-226): per local DoF the row
//    of (true DoF, weight), sorted by true DoF, empty for an eliminated DoF;
//  * ActiveAncestorCache::find (assembly/detail/active_cell_locator.hpp:21-80): the first
//    active cell of a coarser space on a fine cell's parent chain.
// Everything here is integer / host work; the spaces feed apf_general_system_create.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/apf.h"

extern thread_local std::string g_last_error;

namespace {

long long coord_key(double x) { return std::llround(1e12 * x); }

struct Cell1 {
    int id = -1, parent = -1, generation = 0, slev = 0, tlev = 0;
    bool leaf = true;
    int sv[2] = {-1, -1}, tv[2] = {-1, -1};
    std::vector<int> children;
    bool sb[2] = {false, false}, tb[2] = {false, false};
};

}  // namespace

struct apf_mesh1d {
    std::vector<Cell1> cells;
    std::vector<double> xv, tv;
    std::unordered_map<long long, int> xreg, treg;
    std::vector<int> xbnd, tbnd;

    static int vertex(std::vector<double>& V, std::unordered_map<long long, int>& R, double x)
    {
        const long long k = coord_key(x);
        const auto it = R.find(k);
        if (it != R.end()) return it->second;
        const int id = static_cast<int>(V.size());
        V.push_back(x);
        R.emplace(k, id);
        return id;
    }

    void finalize(Cell1& c) const
    {
        if (xv[c.sv[1]] < xv[c.sv[0]]) std::swap(c.sv[0], c.sv[1]);
        if (tv[c.tv[1]] < tv[c.tv[0]]) std::swap(c.tv[0], c.tv[1]);
        for (int f = 0; f < 2; ++f) {
            c.sb[f] = std::find(xbnd.begin(), xbnd.end(), c.sv[f]) != xbnd.end();
            c.tb[f] = std::find(tbnd.begin(), tbnd.end(), c.tv[f]) != tbnd.end();
        }
    }

    void create_root(double x0, double x1, double t0, double t1)
    {
        Cell1 c;
        c.id = 0;
        c.sv[0] = vertex(xv, xreg, x0);
        c.sv[1] = vertex(xv, xreg, x1);
        c.tv[0] = vertex(tv, treg, t0);
        c.tv[1] = vertex(tv, treg, t1);
        xbnd = {c.sv[0], c.sv[1]};
        tbnd = {c.tv[0], c.tv[1]};
        finalize(c);
        cells.push_back(c);
    }

    // space-time split of a leaf: the only 1+1D split of the generation policy
    void refine(int id)
    {
        if (id < 0 || id >= static_cast<int>(cells.size())) throw std::invalid_argument("refine: cell id out of range");
        if (!cells[id].leaf) throw std::invalid_argument("refine: can only refine a leaf cell");
        const Cell1 parent = cells[id];
        const double xm = 0.5 * (xv[parent.sv[0]] + xv[parent.sv[1]]);
        const double tm = 0.5 * (tv[parent.tv[0]] + tv[parent.tv[1]]);
        const int xmid = vertex(xv, xreg, xm);
        const int tmid = vertex(tv, treg, tm);
        const int next = static_cast<int>(cells.size());
        const int xs[2][2] = {{parent.sv[0], xmid}, {xmid, parent.sv[1]}};
        const int ts[2][2] = {{parent.tv[0], tmid}, {tmid, parent.tv[1]}};
        std::vector<int> kids;
        for (int k = 0; k < 4; ++k) {
            const int ix = k & 1, it = k >> 1;  // (left,bottom) (right,bottom) (left,top) (right,top)
            Cell1 c;
            c.id = next + k;
            c.parent = parent.id;
            c.generation = parent.generation + 1;
            c.slev = parent.slev + 1;
            c.tlev = parent.tlev + 1;
            c.sv[0] = xs[ix][0];
            c.sv[1] = xs[ix][1];
            c.tv[0] = ts[it][0];
            c.tv[1] = ts[it][1];
            finalize(c);
            cells.push_back(c);
            kids.push_back(c.id);
        }
        cells[id].leaf = false;
        cells[id].children = kids;
    }
};

namespace {

struct Iface {
    int mc = -1, mf = -1, sc = -1, sf = -1;
    bool hanging = false, boundary = false;
};

struct FaceRec {
    int cell, face, fixed, lid, rid;
    double left, right;
};

struct DofRec {
    bool constrained = false;
    std::vector<int> masters;
    std::vector<double> weights;
};

// barycentric (second form) equidistant Lagrange values, with the reference's node hit test
std::vector<double> seg_eval_all(int p, double xi)
{
    std::vector<double> nodes(p + 1), w(p + 1), L(p + 1, 0.0);
    for (int i = 0; i <= p; ++i) nodes[i] = static_cast<double>(i) / static_cast<double>(p);
    for (int i = 0; i <= p; ++i) {
        double prod = 1.0;
        for (int j = 0; j <= p; ++j)
            if (i != j) prod *= (nodes[i] - nodes[j]);
        w[i] = 1.0 / prod;
    }
    const double eps = 32.0 * 2.220446049250313e-16;
    for (int i = 0; i <= p; ++i) {
        const double dx = xi - nodes[i];
        const double scale = std::max(1.0, std::abs(xi) + std::abs(nodes[i]));
        if (std::abs(dx) <= eps * scale) {
            L[i] = 1.0;
            return L;
        }
    }
    std::vector<double> t(p + 1);
    double s1 = 0.0;
    for (int i = 0; i <= p; ++i) {
        const double inv = 1.0 / (xi - nodes[i]);
        t[i] = w[i] * inv;
        s1 += t[i];
    }
    const double inv_s1 = 1.0 / s1;
    for (int i = 0; i <= p; ++i) L[i] = t[i] * inv_s1;
    return L;
}

struct DSU {
    std::vector<int> parent, rank;
    explicit DSU(int n) : parent(n), rank(n, 0)
    {
        for (int i = 0; i < n; ++i) parent[i] = i;
    }
    int find(int x)
    {
        while (parent[x] != x) {
            parent[x] = parent[parent[x]];
            x = parent[x];
        }
        return x;
    }
    void unite(int a, int b)
    {
        a = find(a);
        b = find(b);
        if (a == b) return;
        if (rank[a] < rank[b]) std::swap(a, b);
        parent[b] = a;
        if (rank[a] == rank[b]) ++rank[a];
    }
};

}  // namespace

struct apf_fespace1d {
    apf_mesh1d* mesh = nullptr;
    int p = 1;
    bool ct = true;  // continuous in time (X); the Y space is continuous in space only
    std::vector<int> active;
    std::vector<int> pos;  // position in `active` by cell id, -1 when inactive
    std::vector<Iface> sif, tif;
    std::vector<int> cdofs;  // [position in active][n_local] global DoF ids (-1 eliminated)
    std::vector<DofRec> dofs;
    int ntrue = 0;

    int nl() const { return (p + 1) * (p + 1); }
    // local = i_space * (p+1) + j_time (tables/element_tables.hpp:182-236)
    int sfd(int face, int j) const { return (face == 0 ? 0 : p) * (p + 1) + j; }
    int tfd(int face, int i) const { return i * (p + 1) + (face == 0 ? 0 : p); }
    bool is_active(int c) const { return c >= 0 && c < static_cast<int>(pos.size()) && pos[c] >= 0; }

    void grow() { pos.resize(mesh->cells.size(), -1); }

    void set_active(const std::vector<int>& cells)
    {
        grow();
        for (int c : cells) {
            if (c < 0 || c >= static_cast<int>(mesh->cells.size())) throw std::invalid_argument("active cell id out of range");
            if (pos[c] >= 0) throw std::invalid_argument("duplicate active cell id");
            pos[c] = static_cast<int>(active.size());
            active.push_back(c);
        }
    }

    void update_active(const std::unordered_set<int>& add, const std::unordered_set<int>& rem)
    {
        grow();
        std::vector<int> r(rem.begin(), rem.end()), a(add.begin(), add.end());
        std::sort(r.begin(), r.end());
        std::sort(a.begin(), a.end());
        for (int c : r) {
            if (!is_active(c)) continue;
            const int at = pos[c];
            const int moved = active.back();
            active[at] = moved;
            pos[moved] = at;
            active.pop_back();
            pos[c] = -1;
        }
        for (int c : a) {
            if (is_active(c)) continue;
            pos[c] = static_cast<int>(active.size());
            active.push_back(c);
        }
    }

    void match_group(const std::vector<FaceRec>& g, bool spatial)
    {
        auto& out = spatial ? sif : tif;
        std::vector<char> matched(g.size(), 0);
        std::vector<int> open;
        for (int i = 0; i < static_cast<int>(g.size()); ++i) {
            const FaceRec& cur = g[i];
            while (!open.empty() && g[open.back()].right <= cur.left) open.pop_back();
            bool local = false;
            for (auto it = open.rbegin(); it != open.rend(); ++it) {
                const FaceRec& cand = g[*it];
                if (!(cand.lid == cur.lid && cand.rid == cur.rid)) continue;
                if (cand.cell < cur.cell) {
                    Iface f;
                    f.mc = cand.cell;
                    f.mf = cand.face;
                    f.sc = cur.cell;
                    f.sf = cur.face;
                    out.push_back(f);
                }
                matched[i] = 1;
                matched[*it] = 1;
                local = true;
                break;
            }
            if (!local) {
                for (auto it = open.rbegin(); it != open.rend(); ++it) {
                    const FaceRec& cand = g[*it];
                    if (!(cand.left <= cur.left && cur.right <= cand.right)) continue;
                    Iface f;
                    f.mc = cand.cell;
                    f.mf = cand.face;
                    f.sc = cur.cell;
                    f.sf = cur.face;
                    f.hanging = true;
                    out.push_back(f);
                    matched[i] = 1;
                    matched[*it] = 1;
                    break;
                }
            }
            open.push_back(i);
        }
        for (int i = 0; i < static_cast<int>(g.size()); ++i) {
            if (matched[i]) continue;
            const Cell1& c = mesh->cells[g[i].cell];
            if (!(spatial ? c.sb[g[i].face] : c.tb[g[i].face])) continue;
            Iface f;
            f.mc = g[i].cell;
            f.mf = g[i].face;
            f.boundary = true;
            out.push_back(f);
        }
    }

    static void sort_group(std::vector<FaceRec>& g)
    {
        std::sort(g.begin(), g.end(), [](const FaceRec& a, const FaceRec& b) {
            if (a.left != b.left) return a.left < b.left;
            if (a.right != b.right) return a.right > b.right;
            if (a.cell != b.cell) return a.cell < b.cell;
            return a.face < b.face;
        });
    }

    void build_adjacency()
    {
        sif.clear();
        tif.clear();
        const auto& X = mesh->xv;
        const auto& T = mesh->tv;
        for (int pass = 0; pass < 2; ++pass) {
            const bool spatial = pass == 0;
            std::unordered_map<int, std::vector<FaceRec>> groups;
            groups.reserve(active.size() * 2);
            for (int cid : active) {
                const Cell1& c = mesh->cells[cid];
                for (int face = 0; face < 2; ++face) {
                    FaceRec f;
                    f.cell = cid;
                    f.face = face;
                    if (spatial) {
                        f.fixed = c.sv[face];
                        f.lid = c.tv[0];
                        f.rid = c.tv[1];
                        f.left = T[c.tv[0]];
                        f.right = T[c.tv[1]];
                    } else {
                        f.fixed = c.tv[face];
                        f.lid = c.sv[0];
                        f.rid = c.sv[1];
                        f.left = X[c.sv[0]];
                        f.right = X[c.sv[1]];
                    }
                    groups[f.fixed].push_back(f);
                }
            }
            for (auto& kv : groups) sort_group(kv.second);
            for (const auto& kv : groups) match_group(kv.second, spatial);
        }
    }

    // refine_closure_recursive_1d: one wave, then the cells whose hanging neighbours are more
    // than one level finer
    bool closure(const std::vector<int>& marked)
    {
        if (marked.empty()) return false;
        std::vector<int> children;
        std::unordered_set<int> add, rem, seen;
        for (int cid : marked) {
            if (!is_active(cid)) continue;
            if (mesh->cells[cid].leaf) mesh->refine(cid);
            grow();
            const Cell1& parent = mesh->cells[cid];
            if (parent.children.empty()) throw std::runtime_error("refined active cell has no children");
            rem.insert(cid);
            for (int ch : parent.children) {
                add.insert(ch);
                if (seen.insert(ch).second) children.push_back(ch);
            }
        }
        if (children.empty()) return false;
        update_active(add, rem);
        build_adjacency();
        std::vector<int> forced;
        std::unordered_set<int> forced_seen;
        for (const Iface& f : tif) {  // spatial grading, on temporal interfaces
            if (f.boundary || !f.hanging || f.sc < 0) continue;
            if (!is_active(f.mc) || !is_active(f.sc)) continue;
            if (mesh->cells[f.sc].slev > mesh->cells[f.mc].slev + 1 && forced_seen.insert(f.mc).second)
                forced.push_back(f.mc);
        }
        for (const Iface& f : sif) {  // temporal grading, on spatial interfaces
            if (f.boundary || !f.hanging || f.sc < 0) continue;
            if (!is_active(f.mc) || !is_active(f.sc)) continue;
            if (mesh->cells[f.sc].tlev > mesh->cells[f.mc].tlev + 1 && forced_seen.insert(f.mc).second)
                forced.push_back(f.mc);
        }
        (void)closure(forced);
        return true;
    }

    void refine(const std::vector<int>& marked)
    {
        std::vector<int> unique;
        std::unordered_set<int> seen;
        for (int c : marked)
            if (is_active(c) && seen.insert(c).second) unique.push_back(c);
        if (unique.empty()) return;
        const bool current = closure(unique);
        if (!current) build_adjacency();
        distribute();
    }

    void rebuild()
    {
        build_adjacency();
        distribute();
    }

    // physical coordinate of a local DoF (fespace/dofs/physical_dof_coords.hpp:6-26)
    void coord(int cid, int l, double& x, double& t) const
    {
        const Cell1& c = mesh->cells[cid];
        const int i = l / (p + 1), j = l % (p + 1);
        const double xi = static_cast<double>(i) / p, tau = static_cast<double>(j) / p;
        const double x0 = mesh->xv[c.sv[0]], x1 = mesh->xv[c.sv[1]];
        const double t0 = mesh->tv[c.tv[0]], t1 = mesh->tv[c.tv[1]];
        x = x0 + (x1 - x0) * xi;
        t = t0 + (t1 - t0) * tau;
    }

    struct Slave {
        bool is = false, spatial = false;
        int iface = -1;
    };

    void distribute()
    {
        const int NL = nl();
        const int nocc = static_cast<int>(active.size()) * NL;
        auto lin = [&](int cid, int l) {
            if (!is_active(cid)) throw std::runtime_error("DoF distribution failed: local DoF reference is not active");
            return pos[cid] * NL + l;
        };
        std::vector<char> elim(nocc, 0);
        std::vector<Slave> slave(nocc);
        for (int a = 0; a < static_cast<int>(active.size()); ++a) {
            const Cell1& c = mesh->cells[active[a]];
            for (int l = 0; l < NL; ++l) {
                const int i = l / (p + 1);
                const int face = i == 0 ? 0 : (i == p ? 1 : -1);
                if (face >= 0 && c.sb[face]) elim[a * NL + l] = 1;
            }
        }
        auto mark = [&](int cid, int l, bool spatial, int k) {
            const int id = lin(cid, l);
            if (elim[id]) return;
            if (!slave[id].is) {
                slave[id].is = true;
                slave[id].spatial = spatial;
                slave[id].iface = k;
            }
        };
        for (int k = 0; k < static_cast<int>(sif.size()); ++k) {
            if (!sif[k].hanging) continue;
            for (int j = 0; j <= p; ++j) mark(sif[k].sc, sfd(sif[k].sf, j), true, k);
        }
        if (ct)
            for (int k = 0; k < static_cast<int>(tif.size()); ++k) {
                if (!tif[k].hanging) continue;
                for (int i = 0; i <= p; ++i) mark(tif[k].sc, tfd(tif[k].sf, i), false, k);
            }
        DSU dsu(nocc);
        auto unite_pair = [&](int a, int b) {
            if (elim[a] || elim[b]) return;
            if (slave[a].is || slave[b].is) return;
            dsu.unite(a, b);
        };
        for (const Iface& f : sif) {
            if (f.boundary || f.hanging) continue;
            for (int j = 0; j <= p; ++j) unite_pair(lin(f.mc, sfd(f.mf, j)), lin(f.sc, sfd(f.sf, j)));
        }
        if (ct) {
            for (const Iface& f : tif) {
                if (f.boundary || f.hanging) continue;
                for (int i = 0; i <= p; ++i) unite_pair(lin(f.mc, tfd(f.mf, i)), lin(f.sc, tfd(f.sf, i)));
            }
            std::map<std::pair<long long, long long>, int> owner;
            for (int id = 0; id < nocc; ++id) {
                if (elim[id] || slave[id].is) continue;
                const int l = id % NL, i = l / (p + 1), j = l % (p + 1);
                if (!((i == 0 || i == p) && (j == 0 || j == p))) continue;
                double x, t;
                coord(active[id / NL], l, x, t);
                const auto key = std::make_pair(coord_key(x), coord_key(t));
                const auto it = owner.find(key);
                if (it == owner.end()) owner.emplace(key, id);
                else dsu.unite(id, it->second);
            }
        }
        std::vector<int> rep(nocc);
        for (int id = 0; id < nocc; ++id) rep[id] = dsu.find(id);
        cdofs.assign(static_cast<size_t>(nocc), -1);
        dofs.clear();
        std::unordered_map<int, int> root_gid;
        for (int id = 0; id < nocc; ++id) {
            if (elim[id] || slave[id].is) continue;
            const auto it = root_gid.find(rep[id]);
            int gid;
            if (it == root_gid.end()) {
                gid = static_cast<int>(dofs.size());
                dofs.emplace_back();
                root_gid.emplace(rep[id], gid);
            } else {
                gid = it->second;
            }
            cdofs[id] = gid;
        }
        ntrue = static_cast<int>(dofs.size());
        // constrained DoFs: masters expanded through the hanging interfaces, recursively
        std::function<void(int, int, double, std::map<int, double>&)> acc;
        acc = [&](int cid, int l, double factor, std::map<int, double>& out) {
            if (std::abs(factor) < 1e-14) return;
            const int id = lin(cid, l);
            if (elim[id]) return;
            const Slave& s = slave[id];
            if (!s.is) {
                out[root_gid.at(rep[id])] += factor;
                return;
            }
            double x, t;
            coord(cid, l, x, t);
            if (s.spatial) {
                const Iface& f = sif[s.iface];
                const Cell1& m = mesh->cells[f.mc];
                const double tm0 = mesh->tv[m.tv[0]], tm1 = mesh->tv[m.tv[1]];
                const double den = tm1 - tm0;
                if (std::abs(den) < 1e-15) throw std::runtime_error("Degenerate master temporal interval in spatial hanging constraint.");
                const std::vector<double> w = seg_eval_all(p, (t - tm0) / den);
                for (int j = 0; j <= p; ++j) acc(f.mc, sfd(f.mf, j), factor * w[j], out);
            } else {
                const Iface& f = tif[s.iface];
                const Cell1& m = mesh->cells[f.mc];
                const double xm0 = mesh->xv[m.sv[0]], xm1 = mesh->xv[m.sv[1]];
                const double den = xm1 - xm0;
                if (std::abs(den) < 1e-15) throw std::runtime_error("Degenerate master spatial interval in temporal hanging constraint.");
                const std::vector<double> w = seg_eval_all(p, (x - xm0) / den);
                for (int i = 0; i <= p; ++i) acc(f.mc, tfd(f.mf, i), factor * w[i], out);
            }
        };
        std::vector<int> slaves;
        for (int id = 0; id < nocc; ++id)
            if (!elim[id] && slave[id].is) slaves.push_back(id);
        std::sort(slaves.begin(), slaves.end(), [&](int a, int b) {
            const int ca = active[a / NL], cb = active[b / NL];
            const int ga = mesh->cells[ca].generation, gb = mesh->cells[cb].generation;
            if (ga != gb) return ga < gb;
            if (ca != cb) return ca < cb;
            return a % NL < b % NL;
        });
        std::map<std::pair<long long, long long>, int> con_gid;
        for (int id : slaves) {
            const int cid = active[id / NL], l = id % NL;
            std::map<int, double> mw;
            acc(cid, l, 1.0, mw);
            std::vector<std::pair<int, double>> pairs;
            for (const auto& kv : mw)
                if (std::abs(kv.second) > 1e-12) pairs.push_back(kv);
            if (pairs.empty()) continue;
            if (pairs.size() == 1 && std::abs(pairs[0].second - 1.0) < 1e-12) {
                cdofs[id] = pairs[0].first;
                continue;
            }
            double x, t;
            coord(cid, l, x, t);
            const auto key = std::make_pair(coord_key(x), coord_key(t));
            const auto it = con_gid.find(key);
            if (it != con_gid.end()) {
                const DofRec& e = dofs[it->second];
                bool same = e.constrained && e.masters.size() == pairs.size();
                for (size_t k = 0; same && k < pairs.size(); ++k)
                    same = e.masters[k] == pairs[k].first && std::abs(e.weights[k] - pairs[k].second) <= 1e-12;
                if (!same)
                    throw std::runtime_error(
                        "DoF distribution failed: same constrained coordinate has inconsistent masters/weights.");
                cdofs[id] = it->second;
                continue;
            }
            DofRec d;
            d.constrained = true;
            for (const auto& pr : pairs) {
                d.masters.push_back(pr.first);
                d.weights.push_back(pr.second);
            }
            const int gid = static_cast<int>(dofs.size());
            dofs.push_back(d);
            con_gid.emplace(key, gid);
            cdofs[id] = gid;
        }
        for (int id = 0; id < nocc; ++id) {
            if (elim[id]) {
                if (cdofs[id] != -1) throw std::runtime_error("Dirichlet-eliminated DoF should remain -1.");
            } else if (cdofs[id] < 0) {
                throw std::runtime_error("DoF distribution failed: some non-eliminated local DoFs were left unassigned.");
            }
        }
    }

    // restriction rows of the active cells, in active order
    void restriction(std::vector<int64_t>& ptr, std::vector<int>& dof, std::vector<double>& w) const
    {
        ptr.assign(1, 0);
        dof.clear();
        w.clear();
        for (size_t id = 0; id < cdofs.size(); ++id) {
            const int gid = cdofs[id];
            std::vector<std::pair<int, double>> row;
            if (gid >= 0) {
                const DofRec& d = dofs[gid];
                if (!d.constrained) {
                    row.emplace_back(gid, 1.0);
                } else {
                    for (size_t k = 0; k < d.masters.size(); ++k) {
                        const int m = d.masters[k];
                        if (m < 0) continue;
                        if (dofs[m].constrained)
                            throw std::runtime_error("build_cell_restriction: constrained DoF references constrained master.");
                        bool merged = false;
                        for (auto& e : row)
                            if (e.first == m) {
                                e.second += d.weights[k];
                                merged = true;
                            }
                        if (!merged) row.emplace_back(m, d.weights[k]);
                    }
                    std::sort(row.begin(), row.end(),
                              [](const std::pair<int, double>& a, const std::pair<int, double>& b) { return a.first < b.first; });
                }
            }
            for (const auto& e : row) {
                dof.push_back(e.first);
                w.push_back(e.second);
            }
            ptr.push_back(static_cast<int64_t>(dof.size()));
        }
    }
};

namespace {

int bad(const std::string& m)
{
    g_last_error = m;
    return APF_BAD_ARG;
}

}  // namespace

extern "C" {

int apf_mesh1d_create(double x0, double x1, double t0, double t1, apf_mesh1d** out)
{
    if (!out) return bad("apf_mesh1d_create: null argument");
    if (!(std::isfinite(x0) && std::isfinite(x1) && std::isfinite(t0) && std::isfinite(t1)) || !(x1 > x0) || !(t1 > t0))
        return bad("apf_mesh1d_create: need finite x0 < x1 and t0 < t1");
    auto* m = new apf_mesh1d();
    m->create_root(x0, x1, t0, t1);
    *out = m;
    return APF_OK;
}

int apf_mesh1d_create_slab(const apf_mesh1d* src, const int* source_cells, int64_t n, double t_begin, double t_end,
                           apf_mesh1d** out)
{
    if (!src || !out || (n > 0 && !source_cells) || n < 0) return bad("apf_mesh1d_create_slab: null argument");
    if (!(std::isfinite(t_begin) && std::isfinite(t_end)) || !(t_begin < t_end))
        return bad("apf_mesh1d_create_slab: expected t_begin < t_end");
    auto* m = new apf_mesh1d();
    try {
        for (int64_t i = 0; i < n; ++i) {
            const int sc = source_cells[i];
            if (sc < 0 || sc >= static_cast<int>(src->cells.size()))
                throw std::invalid_argument("source cell id out of range");
            const Cell1& c = src->cells[sc];
            const double t0 = src->tv[c.tv[0]], t1 = src->tv[c.tv[1]];
            const double tol = 1e-14;
            if (t_begin < t0 - tol || t_end > t1 + tol)
                throw std::invalid_argument("requested slice is not contained in source cell");
            Cell1 s;
            s.id = static_cast<int>(m->cells.size());
            s.parent = -1;
            s.leaf = true;
            s.generation = c.generation;
            s.slev = c.slev;
            s.tlev = c.tlev;
            s.sv[0] = apf_mesh1d::vertex(m->xv, m->xreg, src->xv[c.sv[0]]);
            s.sv[1] = apf_mesh1d::vertex(m->xv, m->xreg, src->xv[c.sv[1]]);
            s.tv[0] = apf_mesh1d::vertex(m->tv, m->treg, t_begin);
            s.tv[1] = apf_mesh1d::vertex(m->tv, m->treg, t_end);
            if (m->xv[s.sv[1]] < m->xv[s.sv[0]]) std::swap(s.sv[0], s.sv[1]);
            s.sb[0] = c.sb[0];
            s.sb[1] = c.sb[1];
            s.tb[0] = s.tb[1] = true;   // both temporal faces bound the slab-local problem
            for (int f = 0; f < 2; ++f)
                if (s.sb[f] && std::find(m->xbnd.begin(), m->xbnd.end(), s.sv[f]) == m->xbnd.end())
                    m->xbnd.push_back(s.sv[f]);
            m->cells.push_back(s);
        }
        m->tbnd.clear();
        for (int t = 0; t < static_cast<int>(m->tv.size()); ++t) m->tbnd.push_back(t);
    } catch (const std::exception& e) {
        delete m;
        return bad(std::string("apf_mesh1d_create_slab: ") + e.what());
    }
    *out = m;
    return APF_OK;
}

void apf_mesh1d_destroy(apf_mesh1d* m) { delete m; }

int apf_mesh1d_info(const apf_mesh1d* m, int64_t* n_cells, int64_t* n_spatial_vertices, int64_t* n_temporal_vertices)
{
    if (!m) return bad("apf_mesh1d_info: null mesh");
    if (n_cells) *n_cells = static_cast<int64_t>(m->cells.size());
    if (n_spatial_vertices) *n_spatial_vertices = static_cast<int64_t>(m->xv.size());
    if (n_temporal_vertices) *n_temporal_vertices = static_cast<int64_t>(m->tv.size());
    return APF_OK;
}

int apf_mesh1d_vertices(const apf_mesh1d* m, double* x, double* t)
{
    if (!m) return bad("apf_mesh1d_vertices: null mesh");
    if (x) std::copy(m->xv.begin(), m->xv.end(), x);
    if (t) std::copy(m->tv.begin(), m->tv.end(), t);
    return APF_OK;
}

int apf_mesh1d_cells(const apf_mesh1d* m, int* parent, int* generation, int* is_leaf, int* vertex_ids, double* box)
{
    if (!m) return bad("apf_mesh1d_cells: null mesh");
    for (size_t c = 0; c < m->cells.size(); ++c) {
        const Cell1& e = m->cells[c];
        if (parent) parent[c] = e.parent;
        if (generation) generation[c] = e.generation;
        if (is_leaf) is_leaf[c] = e.leaf ? 1 : 0;
        if (vertex_ids) {
            vertex_ids[4 * c + 0] = e.sv[0];
            vertex_ids[4 * c + 1] = e.sv[1];
            vertex_ids[4 * c + 2] = e.tv[0];
            vertex_ids[4 * c + 3] = e.tv[1];
        }
        if (box) {
            box[4 * c + 0] = m->xv[e.sv[0]];
            box[4 * c + 1] = m->xv[e.sv[1]];
            box[4 * c + 2] = m->tv[e.tv[0]];
            box[4 * c + 3] = m->tv[e.tv[1]];
        }
    }
    return APF_OK;
}

int apf_fespace1d_create(apf_mesh1d* m, int p, int continuous_in_time, const int* active_cells, int64_t n_active,
                         apf_fespace1d** out)
{
    if (!m || !out || (n_active > 0 && !active_cells) || n_active < 0) return bad("apf_fespace1d_create: null argument");
    if (p < 1 || p > 3) return bad("apf_fespace1d_create: p must be 1, 2 or 3");
    auto* s = new apf_fespace1d();
    try {
        s->mesh = m;
        s->p = p;
        s->ct = continuous_in_time != 0;
        s->set_active(std::vector<int>(active_cells, active_cells + n_active));
        s->rebuild();
    } catch (const std::exception& e) {
        delete s;
        return bad(std::string("apf_fespace1d_create: ") + e.what());
    }
    *out = s;
    return APF_OK;
}

void apf_fespace1d_destroy(apf_fespace1d* s) { delete s; }

int apf_fespace1d_refine(apf_fespace1d* s, const int* marked, int64_t n_marked)
{
    if (!s || (n_marked > 0 && !marked) || n_marked < 0) return bad("apf_fespace1d_refine: null argument");
    try {
        s->refine(std::vector<int>(marked, marked + n_marked));
    } catch (const std::exception& e) {
        return bad(std::string("apf_fespace1d_refine: ") + e.what());
    }
    return APF_OK;
}

int apf_fespace1d_info(const apf_fespace1d* s, int64_t* n_active, int64_t* n_true_dofs, int64_t* n_constrained_dofs,
                       int64_t* n_restriction_entries)
{
    if (!s) return bad("apf_fespace1d_info: null space");
    if (n_active) *n_active = static_cast<int64_t>(s->active.size());
    if (n_true_dofs) *n_true_dofs = s->ntrue;
    if (n_constrained_dofs) *n_constrained_dofs = static_cast<int64_t>(s->dofs.size()) - s->ntrue;
    if (n_restriction_entries) {
        std::vector<int64_t> ptr;
        std::vector<int> dof;
        std::vector<double> w;
        try {
            s->restriction(ptr, dof, w);
        } catch (const std::exception& e) {
            return bad(std::string("apf_fespace1d_info: ") + e.what());
        }
        *n_restriction_entries = static_cast<int64_t>(dof.size());
    }
    return APF_OK;
}

int apf_fespace1d_active_cells(const apf_fespace1d* s, int* cells)
{
    if (!s || !cells) return bad("apf_fespace1d_active_cells: null argument");
    std::copy(s->active.begin(), s->active.end(), cells);
    return APF_OK;
}

int apf_fespace1d_cell_dofs(const apf_fespace1d* s, int* gids)
{
    if (!s || !gids) return bad("apf_fespace1d_cell_dofs: null argument");
    std::copy(s->cdofs.begin(), s->cdofs.end(), gids);
    return APF_OK;
}

int apf_fespace1d_constraints(const apf_fespace1d* s, int64_t* ptr, int* masters, double* weights)
{
    if (!s || !ptr) return bad("apf_fespace1d_constraints: null argument");
    int64_t k = 0;
    ptr[0] = 0;
    for (size_t g = s->ntrue; g < s->dofs.size(); ++g) {
        const DofRec& d = s->dofs[g];
        for (size_t e = 0; e < d.masters.size(); ++e, ++k) {
            if (masters) masters[k] = d.masters[e];
            if (weights) weights[k] = d.weights[e];
        }
        ptr[g - s->ntrue + 1] = k;
    }
    return APF_OK;
}

int apf_fespace1d_restriction(const apf_fespace1d* s, int64_t* ptr, int* dof, double* weight)
{
    if (!s || !ptr || !dof || !weight) return bad("apf_fespace1d_restriction: null argument");
    try {
        std::vector<int64_t> p;
        std::vector<int> d;
        std::vector<double> w;
        s->restriction(p, d, w);
        std::copy(p.begin(), p.end(), ptr);
        std::copy(d.begin(), d.end(), dof);
        std::copy(w.begin(), w.end(), weight);
    } catch (const std::exception& e) {
        return bad(std::string("apf_fespace1d_restriction: ") + e.what());
    }
    return APF_OK;
}

int apf_fespace1d_neighbours(const apf_fespace1d* s, int temporal, int* max_per_face, int* flat)
{
    if (!s || !max_per_face) return bad("apf_fespace1d_neighbours: null argument");
    // build_{spatial,temporal}_neighbour_lists + flatten_neighbours
    // (fespace/io/detail/write_mesh_binary_detail.hpp:60-186): interfaces in adjacency order
    const auto& ifs = temporal ? s->tif : s->sif;
    std::vector<std::vector<int>> nb(s->active.size() * 2);
    for (const Iface& f : ifs) {
        if (f.boundary) continue;
        if (!s->is_active(f.mc) || f.sc < 0 || !s->is_active(f.sc)) continue;
        nb[s->pos[f.mc] * 2 + f.mf].push_back(f.sc);
        nb[s->pos[f.sc] * 2 + f.sf].push_back(f.mc);
    }
    int mx = 0;
    for (const auto& v : nb) mx = std::max(mx, static_cast<int>(v.size()));
    *max_per_face = mx;
    if (flat) {
        size_t k = 0;
        for (const auto& v : nb)
            for (int e = 0; e < mx; ++e) flat[k++] = e < static_cast<int>(v.size()) ? v[e] : -1;
    }
    return APF_OK;
}

int apf_fespace1d_ancestors(const apf_fespace1d* coarse, const apf_fespace1d* fine, int* index)
{
    if (!coarse || !fine || !index) return bad("apf_fespace1d_ancestors: null argument");
    if (coarse->mesh != fine->mesh)
        return bad("find_active_ancestor_cell: coarse_space and fine_space do not share the same mesh.");
    for (size_t a = 0; a < fine->active.size(); ++a) {
        int c = fine->active[a];
        while (c >= 0 && !coarse->is_active(c)) c = coarse->mesh->cells[c].parent;
        if (c < 0) return bad("ActiveAncestorCache::find: no active ancestor found in coarse_space.");
        index[a] = coarse->pos[c];
    }
    return APF_OK;
}

}  // extern "C"
