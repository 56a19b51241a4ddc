// Host setup of the space-time multigrid preconditioner (see apf_mg.hpp).
#include "apf_mg.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <atomic>
#include <future>
#include <thread>
#include <random>
#include <stdexcept>
#include <string>

#include "apf_tables.hpp"

namespace apf {

namespace {

// coordinates of local basis node i of spatial cell c
void node_coords(const Space& S, int c, int i, double* out)
{
    const int d = S.dim;
    const int* V = &S.mesh.cells[c * (d + 1)];
    if (d == 1) {
        const double a = S.mesh.vertices[V[0]], b = S.mesh.vertices[V[1]];
        out[0] = a + (b - a) * static_cast<double>(i) / S.p;
        return;
    }
    int k = 0;
    for (int cc = 0; cc <= S.p; ++cc)
        for (int bb = 0; bb <= S.p - cc; ++bb) {
            if (k == i) {
                const int aa = S.p - bb - cc;
                for (int q = 0; q < 2; ++q)
                    out[q] = (aa * S.mesh.vertices[2 * V[0] + q] + bb * S.mesh.vertices[2 * V[1] + q] +
                              cc * S.mesh.vertices[2 * V[2] + q]) / S.p;
                return;
            }
            ++k;
        }
}

// reference coordinates of physical point x in cell c
void to_reference(const Space& S, int c, const double* x, double* xi)
{
    const int d = S.dim;
    const int* V = &S.mesh.cells[c * (d + 1)];
    if (d == 1) {
        const double a = S.mesh.vertices[V[0]], b = S.mesh.vertices[V[1]];
        xi[0] = (x[0] - a) / (b - a);
        return;
    }
    const double* v0 = &S.mesh.vertices[2 * V[0]];
    const double* v1 = &S.mesh.vertices[2 * V[1]];
    const double* v2 = &S.mesh.vertices[2 * V[2]];
    const double J00 = v1[0] - v0[0], J01 = v2[0] - v0[0], J10 = v1[1] - v0[1], J11 = v2[1] - v0[1];
    const double det = J00 * J11 - J01 * J10;
    const double dx = x[0] - v0[0], dy = x[1] - v0[1];
    xi[0] = (J11 * dx - J01 * dy) / det;
    xi[1] = (-J10 * dx + J00 * dy) / det;
}

bool inside(int d, const double* xi, double tol)
{
    if (d == 1) return xi[0] >= -tol && xi[0] <= 1.0 + tol;
    return xi[0] >= -tol && xi[1] >= -tol && xi[0] + xi[1] <= 1.0 + tol;
}

struct Locator {
    const Space* S;
    double lo[2], hi[2];
    int nb[2];
    std::vector<std::vector<int>> buckets;
    explicit Locator(const Space& sp) : S(&sp)
    {
        const int d = sp.dim, nv = sp.mesh.nv(), nc = sp.mesh.ns();
        for (int q = 0; q < 2; ++q) { lo[q] = 0.0; hi[q] = 1.0; nb[q] = 1; }
        for (int q = 0; q < d; ++q) {
            lo[q] = 1e300;
            hi[q] = -1e300;
            for (int v = 0; v < nv; ++v) {
                lo[q] = std::min(lo[q], sp.mesh.vertices[v * d + q]);
                hi[q] = std::max(hi[q], sp.mesh.vertices[v * d + q]);
            }
            nb[q] = std::max(1, static_cast<int>(std::pow(static_cast<double>(nc), 1.0 / d)));
        }
        buckets.assign(static_cast<size_t>(nb[0]) * nb[1], {});
        for (int c = 0; c < nc; ++c) {
            double bl[2] = {1e300, 1e300}, bh[2] = {-1e300, -1e300};
            for (int v = 0; v <= d; ++v)
                for (int q = 0; q < d; ++q) {
                    const double x = sp.mesh.vertices[sp.mesh.cells[c * (d + 1) + v] * d + q];
                    bl[q] = std::min(bl[q], x);
                    bh[q] = std::max(bh[q], x);
                }
            int i0[2] = {0, 0}, i1[2] = {0, 0};
            for (int q = 0; q < d; ++q) {
                i0[q] = bucket(q, bl[q] - 1e-12);
                i1[q] = bucket(q, bh[q] + 1e-12);
            }
            for (int a = i0[0]; a <= i1[0]; ++a)
                for (int b = i0[1]; b <= i1[1]; ++b) buckets[static_cast<size_t>(b) * nb[0] + a].push_back(c);
        }
    }
    int bucket(int q, double x) const
    {
        const double w = (hi[q] - lo[q]) > 0 ? (hi[q] - lo[q]) : 1.0;
        int i = static_cast<int>(std::floor((x - lo[q]) / w * nb[q]));
        return std::min(std::max(i, 0), nb[q] - 1);
    }
    int find(const double* x) const
    {
        const int d = S->dim;
        const int a = bucket(0, x[0]);
        const int b = (d == 2) ? bucket(1, x[1]) : 0;
        for (int c : buckets[static_cast<size_t>(b) * nb[0] + a]) {
            double xi[2];
            to_reference(*S, c, x, xi);
            if (inside(d, xi, 1e-10)) return c;
        }
        return -1;
    }
};

void eval_basis(int d, int p, const double* xi, double* v)
{
    if (d == 1) segment_lagrange(p, xi[0], v, nullptr);
    else triangle_lagrange(p, xi[0], xi[1], v, nullptr, nullptr);
}

CSR transpose(const CSR& A)
{
    CSR T;
    T.nrows = A.ncols;
    T.ncols = A.nrows;
    T.ptr.assign(T.nrows + 1, 0);
    for (int c : A.col) T.ptr[c + 1]++;
    for (int i = 0; i < T.nrows; ++i) T.ptr[i + 1] += T.ptr[i];
    T.col.resize(A.col.size());
    T.val.resize(A.val.size());
    std::vector<int> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int r = 0; r < A.nrows; ++r)
        for (int e = A.ptr[r]; e < A.ptr[r + 1]; ++e) {
            const int q = pos[A.col[e]]++;
            T.col[q] = r;
            T.val[q] = A.val[e];
        }
    return T;
}

// C = R A P restricted to the coarse adjacency pattern (sptr, scol) of the
// coarse space; the Galerkin product of nested Lagrange spaces lives there.
// Rows [i0, i1) of the Galerkin product R A P on the coarse pattern (cptr, ccol).
void galerkin_rows(const CSR& R, const std::vector<int>& fptr, const std::vector<int>& fcol,
                   const std::vector<double>& A, const CSR& P, const std::vector<int>& cptr,
                   const std::vector<int>& ccol, int i0, int i1, std::vector<double>& out)
{
    const int nc = R.nrows;
    std::vector<double> acc(nc, 0.0);
    std::vector<int> mark(nc, -1), touched;
    for (int i = i0; i < i1; ++i) {
        touched.clear();
        for (int e = R.ptr[i]; e < R.ptr[i + 1]; ++e) {
            const int r = R.col[e];
            const double wr = R.val[e];
            for (int f = fptr[r]; f < fptr[r + 1]; ++f) {
                const int k = fcol[f];
                const double a = wr * A[f];
                for (int g = P.ptr[k]; g < P.ptr[k + 1]; ++g) {
                    const int j = P.col[g];
                    if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; touched.push_back(j); }
                    acc[j] += a * P.val[g];
                }
            }
        }
        for (int e = cptr[i]; e < cptr[i + 1]; ++e) out[e] = (mark[ccol[e]] == i) ? acc[ccol[e]] : 0.0;
        for (int j : touched) {
            bool found = std::binary_search(ccol.begin() + cptr[i], ccol.begin() + cptr[i + 1], j);
            if (!found && std::fabs(acc[j]) > 1e-12)
                throw std::runtime_error("setup_mg: Galerkin product leaves the coarse pattern (meshes not nested?)");
        }
    }
}

// R A P on the coarse pattern; the rows are split over threads (each row's sums run in
// the same order as on one thread, so the result does not depend on the split)
std::vector<double> galerkin(const CSR& R, const std::vector<int>& fptr, const std::vector<int>& fcol,
                             const std::vector<double>& A, const CSR& P, const std::vector<int>& cptr,
                             const std::vector<int>& ccol, int threads)
{
    const int nc = R.nrows;
    std::vector<double> out(ccol.size(), 0.0);
    const int nt = std::max(1, std::min(threads, nc / 256));
    std::vector<std::future<void>> parts;
    for (int t = 1; t < nt; ++t)
        parts.push_back(std::async(std::launch::async, [&, t] {
            galerkin_rows(R, fptr, fcol, A, P, cptr, ccol, static_cast<int>(static_cast<int64_t>(nc) * t / nt),
                          static_cast<int>(static_cast<int64_t>(nc) * (t + 1) / nt), out);
        }));
    galerkin_rows(R, fptr, fcol, A, P, cptr, ccol, 0, static_cast<int>(static_cast<int64_t>(nc) / nt), out);
    for (auto& f : parts) f.get();
    return out;
}

// The same bound with the rows of each power-iteration step split over `nt` threads (a
// spin barrier between the phases; the norm partials are summed in thread order, so the
// result does not depend on the timing).
double lmax_bound_threads(int n, const std::vector<int>& ptr, const std::vector<int>& col,
                          const std::vector<double>& A, const std::vector<double>& d, int nt)
{
    double gers = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) s += std::fabs(A[e]);
        gers = std::max(gers, s / d[i]);
    }
    std::vector<double> x(n), y(n);
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (auto& v : x) v = U(rng);
    std::vector<double> part(nt, 0.0), xpart(nt, 0.0);
    std::atomic<int> arrived{0}, generation{0};
    auto barrier = [&]() {
        const int g = generation.load(std::memory_order_acquire);
        if (arrived.fetch_add(1, std::memory_order_acq_rel) == nt - 1) {
            arrived.store(0, std::memory_order_relaxed);
            generation.fetch_add(1, std::memory_order_acq_rel);
        } else {
            while (generation.load(std::memory_order_acquire) == g) std::this_thread::yield();
        }
    };
    double lam = 0.0;
    auto work = [&](int t) {
        const int i0 = static_cast<int>(static_cast<int64_t>(n) * t / nt);
        const int i1 = static_cast<int>(static_cast<int64_t>(n) * (t + 1) / nt);
        for (int it = 0; it < 60; ++it) {
            double nrm = 0.0, xn = 0.0;
            for (int i = i0; i < i1; ++i) {
                double s = 0.0;
                for (int e = ptr[i]; e < ptr[i + 1]; ++e) s += A[e] * x[col[e]];
                y[i] = s / d[i];
                nrm += y[i] * y[i];
                xn += x[i] * x[i];
            }
            part[t] = nrm;
            xpart[t] = xn;
            barrier();
            double tn = 0.0, tx = 0.0;
            for (int q = 0; q < nt; ++q) {
                tn += part[q];
                tx += xpart[q];
            }
            tn = std::sqrt(tn);
            if (t == 0) lam = tn / std::sqrt(tx);
            for (int i = i0; i < i1; ++i) x[i] = y[i] / tn;
            barrier();
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& q : th) q.join();
    return std::min(gers, 1.2 * lam);
}

double lmax_bound(int n, const std::vector<int>& ptr, const std::vector<int>& col, const std::vector<double>& A,
                  const std::vector<double>& d)
{
    // Gershgorin bound and a power-iteration estimate; the smaller one with a margin
    double gers = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int e = ptr[i]; e < ptr[i + 1]; ++e) s += std::fabs(A[e]);
        gers = std::max(gers, s / d[i]);
    }
    std::vector<double> x(n), y(n);
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    for (auto& v : x) v = U(rng);
    double lam = 0.0;
    for (int it = 0; it < 60; ++it) {
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int e = ptr[i]; e < ptr[i + 1]; ++e) s += A[e] * x[col[e]];
            y[i] = s / d[i];
            nrm += y[i] * y[i];
        }
        nrm = std::sqrt(nrm);
        double xn = 0.0;
        for (int i = 0; i < n; ++i) xn += x[i] * x[i];
        lam = nrm / std::sqrt(xn);
        for (int i = 0; i < n; ++i) x[i] = y[i] / nrm;
    }
    return std::min(gers, 1.2 * lam);
}

}  // namespace

void sym_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V)
{
    V.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[i * n + i] * A[i * n + i];
            for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        }
        if (off <= 1e-30 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = A[p * n + p], aqq = A[q * n + q];
                const double tau = (aqq - app) / (2.0 * apq);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

void build_mg_prolongations(const Space& fine, MGHierarchyHost& H)
{
    H.P.clear();
    const int L = static_cast<int>(H.coarse.size()) + 1;
    for (int l = 0; l + 1 < L; ++l) {
        const Space& F = (l == 0) ? fine : H.coarse[l - 1];
        const Space& Cs = H.coarse[l];
        if (F.dim != Cs.dim || F.p != Cs.p) throw std::runtime_error("mg: coarse space dim/p mismatch");
        if (Cs.nf >= F.nf) throw std::runtime_error("mg: coarse meshes must be ordered fine to coarse");
        const int d = F.dim, Ns = F.Ns;
        Locator loc(Cs);
        std::vector<int> parent(F.mesh.ns(), -1);
        for (int c = 0; c < F.mesh.ns(); ++c) {
            double cen[2] = {0.0, 0.0};
            for (int v = 0; v <= d; ++v)
                for (int q = 0; q < d; ++q) cen[q] += F.mesh.vertices[F.mesh.cells[c * (d + 1) + v] * d + q] / (d + 1);
            parent[c] = loc.find(cen);
            if (parent[c] < 0) throw std::runtime_error("mg: fine cell " + std::to_string(c) + " has no coarse parent");
        }
        CSR P;
        P.nrows = F.nf;
        P.ncols = Cs.nf;
        std::vector<int> owner(F.nf, -1), owner_loc(F.nf, -1);
        for (int c = 0; c < F.mesh.ns(); ++c)
            for (int i = 0; i < Ns; ++i) {
                const int f = F.cell_free[c * Ns + i];
                if (f >= 0 && owner[f] < 0) { owner[f] = c; owner_loc[f] = i; }
            }
        std::vector<std::vector<std::pair<int, double>>> rows(F.nf);
        std::vector<double> phi(Cs.Ns);
        for (int f = 0; f < F.nf; ++f) {
            double x[2], xi[2];
            node_coords(F, owner[f], owner_loc[f], x);
            const int cc = parent[owner[f]];
            to_reference(Cs, cc, x, xi);
            if (!inside(d, xi, 1e-9)) throw std::runtime_error("mg: fine node outside its parent cell (not nested)");
            eval_basis(d, F.p, xi, phi.data());
            for (int j = 0; j < Cs.Ns; ++j) {
                const int cf = Cs.cell_free[cc * Cs.Ns + j];
                if (cf >= 0 && std::fabs(phi[j]) > 1e-13) rows[f].push_back({cf, phi[j]});
            }
            std::sort(rows[f].begin(), rows[f].end());
        }
        P.ptr.assign(F.nf + 1, 0);
        for (int f = 0; f < F.nf; ++f) P.ptr[f + 1] = P.ptr[f] + static_cast<int>(rows[f].size());
        for (int f = 0; f < F.nf; ++f)
            for (auto& e : rows[f]) { P.col.push_back(e.first); P.val.push_back(e.second); }
        H.P.push_back(std::move(P));
    }
}

// Time factors of the preconditioner (they depend on the time grid only): the generalised
// eigenvectors W and shifts sqrt(theta) of T1 w = theta T3 w, and (h_k Mt)^-1.
static void setup_time_factors(const Space& S, const std::vector<double>& Mt, const std::vector<double>& Dt,
                               MGSetupHost& out)
{
    const int p = S.p, Nt = S.Nt, m = S.mesh.m();
    const int T = m * p + 1;
    out.T = T;
    std::vector<double> T1(static_cast<size_t>(T) * T, 0.0), T3(static_cast<size_t>(T) * T, 0.0);
    std::vector<double> Mti(Mt);
    invert_dense(Nt, Mti);
    std::vector<double> DMD(static_cast<size_t>(Nt) * Nt, 0.0);  // Dt Mt^-1 Dt^T
    for (int i = 0; i < Nt; ++i)
        for (int j = 0; j < Nt; ++j) {
            double s = 0.0;
            for (int a = 0; a < Nt; ++a)
                for (int b = 0; b < Nt; ++b) s += Dt[i * Nt + a] * Mti[a * Nt + b] * Dt[j * Nt + b];
            DMD[i * Nt + j] = s;
        }
    out.MtInvH.assign(static_cast<size_t>(m) * Nt * Nt, 0.0);
    for (int k = 0; k < m; ++k) {
        const double h = S.mesh.times[k + 1] - S.mesh.times[k];
        for (int i = 0; i < Nt; ++i)
            for (int j = 0; j < Nt; ++j) {
                T3[static_cast<size_t>(k * p + i) * T + (k * p + j)] += h * Mt[i * Nt + j];
                T1[static_cast<size_t>(k * p + i) * T + (k * p + j)] += DMD[i * Nt + j] / h;
                out.MtInvH[(static_cast<size_t>(k) * Nt + i) * Nt + j] = Mti[i * Nt + j] / h;
            }
    }
    // generalised eigenproblem T1 w = theta T3 w:  T3 = L L^T, C = L^-1 T1 L^-T
    std::vector<double> Lc(static_cast<size_t>(T) * T, 0.0);
    for (int j = 0; j < T; ++j) {
        double s = T3[static_cast<size_t>(j) * T + j];
        for (int k = 0; k < j; ++k) s -= Lc[static_cast<size_t>(j) * T + k] * Lc[static_cast<size_t>(j) * T + k];
        if (!(s > 0.0)) throw std::runtime_error("setup_mg: time mass not SPD");
        Lc[static_cast<size_t>(j) * T + j] = std::sqrt(s);
        for (int i = j + 1; i < T; ++i) {
            double t = T3[static_cast<size_t>(i) * T + j];
            for (int k = 0; k < j; ++k) t -= Lc[static_cast<size_t>(i) * T + k] * Lc[static_cast<size_t>(j) * T + k];
            Lc[static_cast<size_t>(i) * T + j] = t / Lc[static_cast<size_t>(j) * T + j];
        }
    }
    std::vector<double> Li(Lc);
    invert_dense(T, Li);
    std::vector<double> C(static_cast<size_t>(T) * T, 0.0), tmp(static_cast<size_t>(T) * T, 0.0);
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
            double s = 0.0;
            for (int k = 0; k < T; ++k) s += Li[static_cast<size_t>(i) * T + k] * T1[static_cast<size_t>(k) * T + j];
            tmp[static_cast<size_t>(i) * T + j] = s;
        }
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
            double s = 0.0;
            for (int k = 0; k < T; ++k) s += tmp[static_cast<size_t>(i) * T + k] * Li[static_cast<size_t>(j) * T + k];
            C[static_cast<size_t>(i) * T + j] = s;
        }
    for (int i = 0; i < T; ++i)
        for (int j = i + 1; j < T; ++j) {
            const double a = 0.5 * (C[static_cast<size_t>(i) * T + j] + C[static_cast<size_t>(j) * T + i]);
            C[static_cast<size_t>(i) * T + j] = C[static_cast<size_t>(j) * T + i] = a;
        }
    std::vector<double> theta, Q;
    sym_eig(T, C, theta, Q);
    out.W.assign(static_cast<size_t>(T) * T, 0.0);  // W = L^-T Q
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) {
            double s = 0.0;
            for (int k = 0; k < T; ++k) s += Li[static_cast<size_t>(k) * T + i] * Q[static_cast<size_t>(k) * T + j];
            out.W[static_cast<size_t>(i) * T + j] = s;
        }
    out.shift.resize(T);
    for (int i = 0; i < T; ++i) out.shift[i] = std::sqrt(std::max(theta[i], 0.0));
}

void setup_mg(const Space& S, const MGHierarchyHost& H, const std::vector<double>& Ksp,
              const std::vector<double>& Msp, const std::vector<double>& Mt, const std::vector<double>& Dt,
              MGSetupHost& out, bool zero_shift_only)
{
    const auto t_start = std::chrono::steady_clock::now();
    const int L = static_cast<int>(H.coarse.size()) + 1;
    out.levels.assign(L, MGLevelHost());
    auto tf = std::async(std::launch::async, [&] { setup_time_factors(S, Mt, Dt, out); });
    // smoother bounds of every level and matrix run on their own threads, each
    // started as soon as its level exists (the Galerkin products of the coarser
    // levels, the time factors and the coarsest inverses proceed meanwhile)
    std::vector<std::future<double>> lm;
    for (int l = 0; l < L; ++l) {
        MGLevelHost& lv = out.levels[l];
        const Space& sp = (l == 0) ? S : H.coarse[l - 1];
        lv.nf = sp.nf;
        lv.sptr = sp.sptr;
        lv.scol = sp.scol;
        if (l == 0) {
            lv.K = Ksp;
            lv.M = Msp;
        } else {
            const MGLevelHost& f = out.levels[l - 1];
            const CSR& P = H.P[l - 1];
            const CSR R = transpose(P);
            // the K and M products at once
            auto gm = std::async(std::launch::async,
                                 [&] { return galerkin(R, f.sptr, f.scol, f.M, P, lv.sptr, lv.scol, 4); });
            lv.K = galerkin(R, f.sptr, f.scol, f.K, P, lv.sptr, lv.scol, 4);
            lv.M = gm.get();
        }
        lv.dK.assign(lv.nf, 0.0);
        lv.dM.assign(lv.nf, 0.0);
        for (int i = 0; i < lv.nf; ++i)
            for (int e = lv.sptr[i]; e < lv.sptr[i + 1]; ++e)
                if (lv.scol[e] == i) { lv.dK[i] = lv.K[e]; lv.dM[i] = lv.M[e]; }
        for (int i = 0; i < lv.nf; ++i)
            if (!(lv.dK[i] > 0.0) || !(lv.dM[i] > 0.0)) throw std::runtime_error("setup_mg: non-positive diagonal");
        // the fine level's power iterations split their rows over 4 threads each
        const int pt = (l == 0 && lv.nf >= 4096) ? 4 : 1;
        lm.push_back(std::async(std::launch::async, [&lv, pt] {
            return pt > 1 ? lmax_bound_threads(lv.nf, lv.sptr, lv.scol, lv.K, lv.dK, pt)
                          : lmax_bound(lv.nf, lv.sptr, lv.scol, lv.K, lv.dK);
        }));
        lm.push_back(std::async(std::launch::async, [&lv, pt] {
            return pt > 1 ? lmax_bound_threads(lv.nf, lv.sptr, lv.scol, lv.M, lv.dM, pt)
                          : lmax_bound(lv.nf, lv.sptr, lv.scol, lv.M, lv.dM);
        }));
        if (l + 1 < L) {
            lv.P = H.P[l];
            lv.R = transpose(H.P[l]);
        }
    }
    const auto t_lv = std::chrono::steady_clock::now();
    tf.get();  // the time factors (computed beside the levels)
    const auto t_tf = std::chrono::steady_clock::now();
    const int T = out.T;
    // coarsest inverses for shift 0 and each time-mode shift, per connected
    // component when the components are contiguous node ranges
    const MGLevelHost& cl = out.levels.back();
    const int nc = cl.nf;
    std::vector<int> comp(nc, -1);
    int ncomp = 0;
    std::vector<int> stack;
    for (int s0 = 0; s0 < nc; ++s0) {
        if (comp[s0] >= 0) continue;
        comp[s0] = ncomp;
        stack.assign(1, s0);
        while (!stack.empty()) {
            const int i = stack.back();
            stack.pop_back();
            for (int e = cl.sptr[i]; e < cl.sptr[i + 1]; ++e)
                if (comp[cl.scol[e]] < 0) {
                    comp[cl.scol[e]] = ncomp;
                    stack.push_back(cl.scol[e]);
                }
        }
        ++ncomp;
    }
    bool contiguous = true;
    for (int i = 1; i < nc; ++i)
        if (comp[i] != comp[i - 1] && comp[i] != comp[i - 1] + 1) contiguous = false;
    if (!contiguous) std::fill(comp.begin(), comp.end(), 0), ncomp = nc > 0 ? 1 : 0;
    out.node_comp = comp;
    out.comp_start.assign(ncomp + 1, nc);
    for (int i = nc - 1; i >= 0; --i) out.comp_start[comp[i]] = i;
    out.blk_off.assign(ncomp, 0);
    out.blk_total = 0;
    int bmax = 0;
    for (int c = 0; c < ncomp; ++c) {
        const int b = out.comp_start[c + 1] - out.comp_start[c];
        out.blk_off[c] = out.blk_total;
        out.blk_total += static_cast<int64_t>(b) * b;
        bmax = std::max(bmax, b);
    }
    if (bmax > 4096) throw std::runtime_error("setup_mg: coarsest level too large (" + std::to_string(bmax) +
                                              " coupled nodes); pass coarser nested meshes");
    out.shifts_all.clear();
    out.shifts_all.push_back(0.0);
    for (int i = 0; i < T; ++i) out.shifts_all.push_back(out.shift[i]);
    const int nsft = static_cast<int>(out.shifts_all.size());
    out.coarse_inv.assign(static_cast<size_t>(nsft) * out.blk_total, 0.0);
    for (int q = 0; q < (zero_shift_only ? 1 : nsft); ++q)
        for (int c = 0; c < ncomp; ++c) {
            const int s = out.comp_start[c], b = out.comp_start[c + 1] - s;
            std::vector<double> A(static_cast<size_t>(b) * b, 0.0);
            for (int i = 0; i < b; ++i)
                for (int e = cl.sptr[s + i]; e < cl.sptr[s + i + 1]; ++e)
                    A[static_cast<size_t>(i) * b + (cl.scol[e] - s)] = cl.K[e] + out.shifts_all[q] * cl.M[e];
            if (!invert_dense(b, A)) throw std::runtime_error("setup_mg: singular coarsest operator");
            std::copy(A.begin(), A.end(), out.coarse_inv.begin() + static_cast<size_t>(q) * out.blk_total + out.blk_off[c]);
        }
    const auto t_ci = std::chrono::steady_clock::now();
    for (int l = 0; l < L; ++l) out.levels[l].lmax = std::max(lm[2 * l].get(), lm[2 * l + 1].get());
    if (std::getenv("APF_MG_TIMING")) {
        const auto t_end = std::chrono::steady_clock::now();
        auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
        fprintf(stderr, "setup_mg: levels %.3f, time factors wait %.3f, coarse inverses %.3f, lmax wait %.3f ms\n",
                ms(t_start, t_lv), ms(t_lv, t_tf), ms(t_tf, t_ci), ms(t_ci, t_end));
    }
}

}  // namespace apf
