// Reference-cell quadrature, bases and integral tables (host, FP64).
#include "apf_tables.hpp"

#include <cmath>
#include <stdexcept>

namespace apf {

static void legendre_pair(int n, double z, double& pn, double& pm)
{
    // P_n(z) and P_{n-1}(z) from the three-term recurrence
    pm = 1.0;
    pn = z;
    for (int k = 1; k < n; ++k) {
        const double pk = ((2.0 * k + 1.0) * z * pn - k * pm) / (k + 1.0);
        pm = pn;
        pn = pk;
    }
}

Rule1D gauss_legendre01(int n)
{
    // Newton's method on P_n started at cos(pi (i + 3/4) / (n + 1/2)).
    Rule1D r;
    r.x.resize(n);
    r.w.resize(n);
    if (n == 1) {
        r.x[0] = 0.5;
        r.w[0] = 1.0;
        return r;
    }
    const double pi = 3.141592653589793238462643383279502884;
    for (int i = 0; i < n; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5));
        double pn, pm;
        for (int it = 0; it < 100; ++it) {
            legendre_pair(n, z, pn, pm);
            const double dp = n * (z * pn - pm) / (z * z - 1.0);
            const double dz = pn / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        legendre_pair(n, z, pn, pm);
        const double dp = n * (z * pn - pm) / (z * z - 1.0);
        // ascending order on [0,1]
        r.x[n - 1 - i] = 0.5 * (z + 1.0);
        r.w[n - 1 - i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
    return r;
}

Rule2D duffy(int degree)
{
    const int order = (degree + 3) / 2;
    const Rule1D g = gauss_legendre01(order);
    Rule2D r;
    for (int i = 0; i < order; ++i)
        for (int j = 0; j < order; ++j) {
            const double u = g.x[i];
            r.x.push_back(u);
            r.y.push_back((1.0 - u) * g.x[j]);
            r.w.push_back(g.w[i] * g.w[j] * (1.0 - u));
        }
    return r;
}

void segment_lagrange(int p, double s, double* v, double* d)
{
    for (int i = 0; i <= p; ++i) {
        double val = 1.0, der = 0.0;
        const double xi = static_cast<double>(i) / p;
        for (int j = 0; j <= p; ++j) {
            if (j == i) continue;
            const double xj = static_cast<double>(j) / p;
            const double f = (s - xj) / (xi - xj);
            der = der * f + val / (xi - xj);
            val *= f;
        }
        v[i] = val;
        if (d) d[i] = der;
    }
}

static void bary_factor(int p, int a, double s, double& val, double& der)
{
    val = 1.0;
    der = 0.0;
    for (int m = 0; m < a; ++m) {
        const double f = (p * s - m) / (m + 1.0);
        der = der * f + val * (p / (m + 1.0));
        val *= f;
    }
}

void triangle_lagrange(int p, double x, double y, double* v, double* gx, double* gy)
{
    const double l0 = 1.0 - x - y;
    int k = 0;
    for (int c = 0; c <= p; ++c)
        for (int b = 0; b <= p - c; ++b) {
            const int a = p - b - c;
            double fa, da, fb, db, fc, dc;
            bary_factor(p, a, l0, fa, da);
            bary_factor(p, b, x, fb, db);
            bary_factor(p, c, y, fc, dc);
            v[k] = fa * fb * fc;
            if (gx) gx[k] = -da * fb * fc + fa * db * fc;
            if (gy) gy[k] = -da * fb * fc + fa * fb * dc;
            ++k;
        }
}

void legendre01(int p, double s, double* v, double* d)
{
    const double x = 2.0 * s - 1.0;
    std::vector<double> P(p + 2, 0.0), D(p + 2, 0.0);
    P[0] = 1.0;
    if (p >= 1) { P[1] = x; D[1] = 1.0; }
    for (int k = 1; k < p; ++k) {
        P[k + 1] = ((2.0 * k + 1.0) * x * P[k] - k * P[k - 1]) / (k + 1.0);
        D[k + 1] = D[k - 1] + (2.0 * k + 1.0) * P[k];
    }
    for (int k = 0; k <= p; ++k) {
        v[k] = P[k];
        if (d) d[k] = 2.0 * D[k];
    }
}

int n_space_basis(int dim, int p) { return dim == 1 ? p + 1 : (p + 1) * (p + 2) / 2; }
int n_flux_local(int dim, int p) { return dim == 1 ? p + 2 : (p + 1) * (p + 3); }
int n_scalar_local(int dim, int p) { return dim == 1 ? p + 1 : (p + 1) * (p + 2) / 2; }
int n_facets(int dim) { return dim == 1 ? 2 : 3; }
int n_facet_moments(int dim, int p) { return dim == 1 ? 1 : p + 1; }

bool invert_dense(int n, std::vector<double>& a)
{
    // Gauss-Jordan with partial pivoting, in place.
    std::vector<int> piv(n);
    std::vector<double> inv(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int r = c;
        for (int i = c + 1; i < n; ++i)
            if (std::fabs(a[i * n + c]) > std::fabs(a[r * n + c])) r = i;
        if (std::fabs(a[r * n + c]) < 1e-300) return false;
        if (r != c)
            for (int j = 0; j < n; ++j) {
                std::swap(a[r * n + j], a[c * n + j]);
                std::swap(inv[r * n + j], inv[c * n + j]);
            }
        const double d = 1.0 / a[c * n + c];
        for (int j = 0; j < n; ++j) { a[c * n + j] *= d; inv[c * n + j] *= d; }
        for (int i = 0; i < n; ++i) {
            if (i == c) continue;
            const double f = a[i * n + c];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) {
                a[i * n + j] -= f * a[c * n + j];
                inv[i * n + j] -= f * inv[c * n + j];
            }
        }
    }
    a = inv;
    return true;
}

// ---------------------------------------------------------------- flux basis
// 2D spanning set of RT_p (centred at the barycentre, RT_p is translation
// invariant): (m,0), (0,m) for monomials m of degree <= p, then m~ (x,y) for
// homogeneous m~ of degree p.
static void rt_spanning(int p, double x, double y, double* vx, double* vy, double* dv)
{
    const double X = x - 1.0 / 3.0, Y = y - 1.0 / 3.0;
    auto pw = [](double t, int e) { double r = 1.0; for (int i = 0; i < e; ++i) r *= t; return r; };
    int j = 0;
    for (int tot = 0; tot <= p; ++tot)
        for (int a = tot; a >= 0; --a) {
            const int b = tot - a;
            const double m = pw(X, a) * pw(Y, b);
            const double mx = a > 0 ? a * pw(X, a - 1) * pw(Y, b) : 0.0;
            const double my = b > 0 ? b * pw(X, a) * pw(Y, b - 1) : 0.0;
            vx[j] = m; vy[j] = 0.0; dv[j] = mx; ++j;
            vx[j] = 0.0; vy[j] = m; dv[j] = my; ++j;
        }
    for (int a = p; a >= 0; --a) {
        const int b = p - a;
        const double m = pw(X, a) * pw(Y, b);
        const double mx = a > 0 ? a * pw(X, a - 1) * pw(Y, b) : 0.0;
        const double my = b > 0 ? b * pw(X, a) * pw(Y, b - 1) : 0.0;
        vx[j] = X * m; vy[j] = Y * m; dv[j] = 2.0 * m + X * mx + Y * my; ++j;
    }
}

static const double REF_V[3][2] = {{0.0, 0.0}, {1.0, 0.0}, {0.0, 1.0}};
static const int REF_E[3][2] = {{1, 2}, {2, 0}, {0, 1}};

FluxBasis make_flux_basis(int dim, int p)
{
    FluxBasis fb;
    fb.dim = dim;
    fb.p = p;
    fb.N = n_flux_local(dim, p);
    if (dim == 1) return fb;
    const int N = fb.N;
    std::vector<double> V(static_cast<size_t>(N) * N, 0.0);
    std::vector<double> vx(N), vy(N), dv(N), L(p + 1);
    const Rule1D g = gauss_legendre01(p + 2);
    int row = 0;
    for (int e = 0; e < 3; ++e) {
        const double* A = REF_V[REF_E[e][0]];
        const double* B = REF_V[REF_E[e][1]];
        const double tx = B[0] - A[0], ty = B[1] - A[1];
        const double nx = ty, ny = -tx;
        for (int k = 0; k <= p; ++k) {
            for (size_t q = 0; q < g.x.size(); ++q) {
                rt_spanning(p, A[0] + g.x[q] * tx, A[1] + g.x[q] * ty, vx.data(), vy.data(), dv.data());
                legendre01(p, g.x[q], L.data(), nullptr);
                for (int j = 0; j < N; ++j) V[row * N + j] += g.w[q] * L[k] * (vx[j] * nx + vy[j] * ny);
            }
            ++row;
        }
    }
    const Rule2D t = duffy(2 * p + 2);
    for (int tot = 0; tot < p; ++tot)
        for (int a = tot; a >= 0; --a) {
            const int b = tot - a;
            for (size_t q = 0; q < t.w.size(); ++q) {
                rt_spanning(p, t.x[q], t.y[q], vx.data(), vy.data(), dv.data());
                const double m = std::pow(t.x[q], a) * std::pow(t.y[q], b);
                for (int j = 0; j < N; ++j) {
                    V[row * N + j] += t.w[q] * m * vx[j];
                    V[(row + 1) * N + j] += t.w[q] * m * vy[j];
                }
            }
            row += 2;
        }
    if (row != N) throw std::runtime_error("make_flux_basis: dof count mismatch");
    if (!invert_dense(N, V)) throw std::runtime_error("make_flux_basis: singular moment matrix");
    fb.coef = V;
    return fb;
}

void FluxBasis::eval(double x, double y, double* val, double* div) const
{
    if (dim == 1) {
        std::vector<double> L(p + 1), dL(p + 1);
        legendre01(p, x, L.data(), dL.data());
        val[0] = 1.0 - x; div[0] = -1.0;
        val[1] = x; div[1] = 1.0;
        const double b = x * (1.0 - x), db = 1.0 - 2.0 * x;
        for (int k = 0; k < p; ++k) {
            val[2 + k] = b * L[k];
            div[2 + k] = db * L[k] + b * dL[k];
        }
        return;
    }
    std::vector<double> vx(N), vy(N), dv(N);
    rt_spanning(p, x, y, vx.data(), vy.data(), dv.data());
    for (int j = 0; j < N; ++j) {
        double sx = 0.0, sy = 0.0, sd = 0.0;
        for (int i = 0; i < N; ++i) {
            const double c = coef[i * N + j];
            sx += vx[i] * c; sy += vy[i] * c; sd += dv[i] * c;
        }
        val[2 * j] = sx;
        val[2 * j + 1] = sy;
        div[j] = sd;
    }
}

// ---------------------------------------------------------------- scalar basis
static void scalar_monomials(int dim, int p, double x, double y, double* m)
{
    if (dim == 1) {
        // shifted Legendre, orthogonal already
        legendre01(p, x, m, nullptr);
        return;
    }
    int j = 0;
    for (int tot = 0; tot <= p; ++tot)
        for (int a = tot; a >= 0; --a) m[j++] = std::pow(x, a) * std::pow(y, tot - a);
}

ScalarBasis make_scalar_basis(int dim, int p)
{
    ScalarBasis sb;
    sb.dim = dim;
    sb.p = p;
    sb.N = n_scalar_local(dim, p);
    const int N = sb.N;
    std::vector<double> G(static_cast<size_t>(N) * N, 0.0), m(N);
    if (dim == 1) {
        const Rule1D r = gauss_legendre01(p + 2);
        for (size_t q = 0; q < r.w.size(); ++q) {
            scalar_monomials(dim, p, r.x[q], 0.0, m.data());
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) G[i * N + j] += r.w[q] * m[i] * m[j];
        }
    } else {
        const Rule2D r = duffy(2 * p + 2);
        for (size_t q = 0; q < r.w.size(); ++q) {
            scalar_monomials(dim, p, r.x[q], r.y[q], m.data());
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) G[i * N + j] += r.w[q] * m[i] * m[j];
        }
    }
    // Cholesky G = L L^T, coef = L^{-T}
    std::vector<double> L(static_cast<size_t>(N) * N, 0.0);
    for (int j = 0; j < N; ++j) {
        double s = G[j * N + j];
        for (int k = 0; k < j; ++k) s -= L[j * N + k] * L[j * N + k];
        L[j * N + j] = std::sqrt(s);
        for (int i = j + 1; i < N; ++i) {
            double t = G[i * N + j];
            for (int k = 0; k < j; ++k) t -= L[i * N + k] * L[j * N + k];
            L[i * N + j] = t / L[j * N + j];
        }
    }
    std::vector<double> Li = L;
    if (!invert_dense(N, Li)) throw std::runtime_error("make_scalar_basis: singular Gram");
    sb.coef.assign(static_cast<size_t>(N) * N, 0.0);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) sb.coef[i * N + j] = Li[j * N + i];
    return sb;
}

void ScalarBasis::eval(double x, double y, double* val) const
{
    std::vector<double> m(N);
    scalar_monomials(dim, p, x, y, m.data());
    for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (int i = 0; i < N; ++i) s += m[i] * coef[i * N + j];
        val[j] = s;
    }
}

// ---------------------------------------------------------------- tables
RefTables make_ref_tables(int dim, int p)
{
    if (dim != 1 && dim != 2) throw std::runtime_error("make_ref_tables: dim must be 1 or 2");
    if (p < 1 || p > 4) throw std::runtime_error("make_ref_tables: p must be 1..4");
    RefTables T;
    T.dim = dim;
    T.p = p;
    const int Ns = T.Ns = n_space_basis(dim, p);
    const int Nt = T.Nt = p + 1;
    const int Nf = T.Nf = n_flux_local(dim, p);
    const int Nq = T.Nq = n_scalar_local(dim, p);
    T.Nfc = n_facets(dim);
    T.Nfm = n_facet_moments(dim, p);
    const int D = dim, NV = dim + 1;
    const FluxBasis fb = make_flux_basis(dim, p);
    const ScalarBasis sb = make_scalar_basis(dim, p);

    T.Khat.assign(static_cast<size_t>(D) * D * Ns * Ns, 0.0);
    T.Mhat.assign(static_cast<size_t>(Ns) * Ns, 0.0);
    T.mhat.assign(Ns, 0.0);
    T.RTab.assign(static_cast<size_t>(D) * D * Nf * Nf, 0.0);
    T.Bhat.assign(static_cast<size_t>(Nq) * Nf, 0.0);
    T.qbar.assign(Nq, 0.0);
    T.R.assign(static_cast<size_t>(NV) * Nf * Ns, 0.0);
    T.Pv.assign(static_cast<size_t>(NV) * Nq, 0.0);
    T.Q.assign(static_cast<size_t>(NV) * Nq * Ns, 0.0);
    T.E.assign(static_cast<size_t>(D) * Nq * Ns, 0.0);
    T.Phi.assign(static_cast<size_t>(Nq) * Ns, 0.0);
    std::vector<double> Gphi(static_cast<size_t>(Nf) * D * D * Ns, 0.0);  // int s_j,a d_b phi_m

    std::vector<double> phi(Ns), gphi(static_cast<size_t>(Ns) * 2), fv(static_cast<size_t>(Nf) * 2),
        fd(Nf), q(Nq), lam(NV);
    auto accumulate = [&](double x, double y, double w) {
        if (dim == 1) {
            std::vector<double> d(Ns);
            segment_lagrange(p, x, phi.data(), d.data());
            for (int i = 0; i < Ns; ++i) gphi[i * 2] = d[i];
            lam[0] = 1.0 - x; lam[1] = x;
            fb.eval(x, 0.0, fv.data(), fd.data());
            // 1D flux values are scalars; spread into [Nf][2] layout with y = 0
            for (int i = Nf - 1; i >= 0; --i) { fv[2 * i] = fv[i]; fv[2 * i + 1] = 0.0; }
        } else {
            std::vector<double> gx(Ns), gy(Ns);
            triangle_lagrange(p, x, y, phi.data(), gx.data(), gy.data());
            for (int i = 0; i < Ns; ++i) { gphi[i * 2] = gx[i]; gphi[i * 2 + 1] = gy[i]; }
            lam[0] = 1.0 - x - y; lam[1] = x; lam[2] = y;
            fb.eval(x, y, fv.data(), fd.data());
        }
        sb.eval(x, y, q.data());
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                for (int i = 0; i < Ns; ++i)
                    for (int j = 0; j < Ns; ++j)
                        T.Khat[((a * D + b) * Ns + i) * Ns + j] += w * gphi[i * 2 + a] * gphi[j * 2 + b];
        for (int i = 0; i < Ns; ++i) {
            T.mhat[i] += w * phi[i];
            for (int j = 0; j < Ns; ++j) T.Mhat[i * Ns + j] += w * phi[i] * phi[j];
        }
        for (int a = 0; a < D; ++a)
            for (int b = 0; b < D; ++b)
                for (int i = 0; i < Nf; ++i)
                    for (int j = 0; j < Nf; ++j)
                        T.RTab[((a * D + b) * Nf + i) * Nf + j] += w * fv[i * 2 + a] * fv[j * 2 + b];
        for (int i = 0; i < Nq; ++i) {
            T.qbar[i] += w * q[i];
            for (int j = 0; j < Nf; ++j) T.Bhat[i * Nf + j] += w * q[i] * fd[j];
            for (int m = 0; m < Ns; ++m) {
                T.Phi[i * Ns + m] += w * q[i] * phi[m];
                for (int b = 0; b < D; ++b) T.E[(b * Nq + i) * Ns + m] += w * gphi[m * 2 + b] * q[i];
            }
        }
        for (int j = 0; j < Nf; ++j)
            for (int a = 0; a < D; ++a)
                for (int b = 0; b < D; ++b)
                    for (int mm = 0; mm < Ns; ++mm)
                        Gphi[((j * D + a) * D + b) * Ns + mm] += w * fv[j * 2 + a] * gphi[mm * 2 + b];
        for (int v = 0; v < NV; ++v) {
            for (int i = 0; i < Nf; ++i)
                for (int m = 0; m < Ns; ++m) {
                    double dot = 0.0;
                    for (int a = 0; a < D; ++a) dot += fv[i * 2 + a] * gphi[m * 2 + a];
                    T.R[(v * Nf + i) * Ns + m] += w * lam[v] * dot;
                }
            for (int i = 0; i < Nq; ++i) {
                T.Pv[v * Nq + i] += w * lam[v] * q[i];
                for (int m = 0; m < Ns; ++m) T.Q[(v * Nq + i) * Ns + m] += w * lam[v] * phi[m] * q[i];
            }
        }
    };
    if (dim == 1) {
        const Rule1D r = gauss_legendre01(p + 4);
        for (size_t k = 0; k < r.w.size(); ++k) accumulate(r.x[k], 0.0, r.w[k]);
    } else {
        const Rule2D r = duffy(2 * p + 4);
        for (size_t k = 0; k < r.w.size(); ++k) accumulate(r.x[k], r.y[k], r.w[k]);
    }

    // coefficients of e_a d_b phi_m in the flux basis: L2(flux) projection,
    // exact because e_a d_b phi_m lies in the local flux space
    {
        std::vector<double> Mrt(static_cast<size_t>(Nf) * Nf, 0.0);
        for (int a = 0; a < D; ++a)
            for (int i = 0; i < Nf; ++i)
                for (int j = 0; j < Nf; ++j) Mrt[i * Nf + j] += T.RTab[((a * D + a) * Nf + i) * Nf + j];
        if (!invert_dense(Nf, Mrt)) throw std::runtime_error("make_ref_tables: singular flux mass");
        T.Igrad.assign(static_cast<size_t>(Nf) * D * D * Ns, 0.0);
        for (int j = 0; j < Nf; ++j)
            for (int ab = 0; ab < D * D; ++ab)
                for (int mm = 0; mm < Ns; ++mm) {
                    double s = 0.0;
                    for (int i = 0; i < Nf; ++i) s += Mrt[j * Nf + i] * Gphi[(i * D * D + ab) * Ns + mm];
                    T.Igrad[(j * D * D + ab) * Ns + mm] = s;
                }
    }

    // temporal tables
    T.Mt.assign(static_cast<size_t>(Nt) * Nt, 0.0);
    T.Dt.assign(static_cast<size_t>(Nt) * Nt, 0.0);
    T.Mdd.assign(static_cast<size_t>(Nt) * Nt, 0.0);
    T.mt.assign(Nt, 0.0);
    T.W.assign(static_cast<size_t>(Nt) * Nt, 0.0);
    T.Wd.assign(static_cast<size_t>(Nt) * Nt, 0.0);
    T.Lnorm.assign(Nt, 0.0);
    {
        const Rule1D r = gauss_legendre01(p + 3);
        std::vector<double> l(Nt), dl(Nt), L(Nt);
        for (size_t k = 0; k < r.w.size(); ++k) {
            segment_lagrange(p, r.x[k], l.data(), dl.data());
            legendre01(p, r.x[k], L.data(), nullptr);
            const double w = r.w[k];
            for (int i = 0; i < Nt; ++i) {
                T.mt[i] += w * l[i];
                T.Lnorm[i] += w * L[i] * L[i];
                for (int j = 0; j < Nt; ++j) {
                    T.Mt[i * Nt + j] += w * l[i] * l[j];
                    T.Dt[i * Nt + j] += w * dl[i] * l[j];
                    T.Mdd[i * Nt + j] += w * dl[i] * dl[j];
                    T.W[i * Nt + j] += w * l[i] * L[j];
                    T.Wd[i * Nt + j] += w * dl[i] * L[j];
                }
            }
        }
    }

    // bottom rule (the reference's spatial rule at QSpace)
    if (dim == 1) {
        const Rule1D r = gauss_legendre01(2 * p + 1);
        T.nb = static_cast<int>(r.w.size());
        T.bx = r.x;
        T.bw = r.w;
        T.bphi.assign(static_cast<size_t>(T.nb) * Ns, 0.0);
        for (int k = 0; k < T.nb; ++k) segment_lagrange(p, r.x[k], &T.bphi[k * Ns], nullptr);
    } else {
        const Rule2D r = duffy(4 * p + 1);
        T.nb = static_cast<int>(r.w.size());
        T.bx.resize(2 * T.nb);
        for (int k = 0; k < T.nb; ++k) { T.bx[2 * k] = r.x[k]; T.bx[2 * k + 1] = r.y[k]; }
        T.bw = r.w;
        T.bphi.assign(static_cast<size_t>(T.nb) * Ns, 0.0);
        for (int k = 0; k < T.nb; ++k) triangle_lagrange(p, r.x[k], r.y[k], &T.bphi[k * Ns], nullptr, nullptr);
    }

    T.dbary.assign(static_cast<size_t>(NV) * D, 0.0);
    if (dim == 1) {
        T.dbary = {-1.0, 1.0};
        T.facet_v = {0, 0, 1, 1};
    } else {
        T.dbary = {-1.0, -1.0, 1.0, 0.0, 0.0, 1.0};
        T.facet_v = {1, 2, 2, 0, 0, 1};
    }
    return T;
}

}  // namespace apf
