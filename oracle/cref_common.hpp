// oracle/cref_common.hpp — shared pieces of the CPU restatement of the reference's fit
// chains (qpAdm, qpGraph): small dense linear algebra, the chi-square tail, the
// survivor rule, LOO/total and the block-jackknife covariance.
//
// TEST INFRASTRUCTURE ONLY (see cpu_ref.cpp).  Restated from the reference's
// semantics (paths relative to the reference repository root):
//   LU / solve / inverse  src/core/internal/small_linalg.hpp:55-135
//   one-sided Jacobi SVD  src/core/internal/small_linalg.hpp:137-218
//   chi-square tail       src/core/internal/pchisq.hpp:16-70
//   survivor blocks       src/device/cpu/cpu_backend.cpp:125-150
//   LOO / total           src/device/cpu/cpu_backend.cpp:2674-2723
//   jackknife covariance  src/device/cpu/cpu_backend.cpp:2219-2271, ridge 2758-2764
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace cref {

using Vec = std::vector<double>;

// ---- small dense linear algebra (column-major, element (i, j) at i + n*j) -----------

inline bool lu_decompose(Vec& a, int n, std::vector<int>& perm) {
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    auto A = [&](int i, int j) -> double& { return a[i + static_cast<std::size_t>(n) * j]; };
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double big = std::fabs(A(k, k));
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(A(i, k)) > big) { big = std::fabs(A(i, k)); piv = i; }
        if (big == 0.0) return false;
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A(k, j), A(piv, j));
            std::swap(perm[k], perm[piv]);
        }
        const double rp = 1.0 / A(k, k);
        for (int i = k + 1; i < n; ++i) {
            const double f = A(i, k) * rp;
            A(i, k) = f;
            for (int j = k + 1; j < n; ++j) A(i, j) -= f * A(k, j);
        }
    }
    return true;
}

// Forward (unit lower) then backward substitution; y holds the permuted rhs.
inline void lu_substitute(const Vec& lu, int n, Vec& y, double* out) {
    auto A = [&](int i, int j) { return lu[i + static_cast<std::size_t>(n) * j]; };
    for (int i = 0; i < n; ++i) {
        double s = y[i];
        for (int j = 0; j < i; ++j) s -= A(i, j) * y[j];
        y[i] = s;
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < n; ++j) s -= A(i, j) * out[j];
        out[i] = s / A(i, i);
    }
}

inline bool lin_solve(const Vec& a, int n, const Vec& b, Vec& x) {
    Vec lu = a;
    std::vector<int> perm;
    if (!lu_decompose(lu, n, perm)) return false;
    Vec y(n);
    for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
    x.assign(n, 0.0);
    lu_substitute(lu, n, y, x.data());
    return true;
}

inline bool mat_inverse(const Vec& a, int n, Vec& inv) {
    Vec lu = a;
    std::vector<int> perm;
    if (!lu_decompose(lu, n, perm)) return false;
    inv.assign(static_cast<std::size_t>(n) * n, 0.0);
    Vec y(n);
    for (int c = 0; c < n; ++c) {
        for (int i = 0; i < n; ++i) y[i] = (perm[i] == c) ? 1.0 : 0.0;
        lu_substitute(lu, n, y, inv.data() + static_cast<std::size_t>(n) * c);
    }
    return true;
}

struct Svd {
    Vec S, V;  // S descending (k = min(m, n) values); V n x k (column j = right vector j)
};

// One-sided Jacobi on the columns of the m x n matrix A: cyclic (p < q) sweeps
// until the squared off-diagonal mass is below 1e-30 (at most 60 sweeps).
inline Svd jacobi(const Vec& A, int m, int n) {
    Vec W = A;
    Vec Vf(static_cast<std::size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) Vf[i + static_cast<std::size_t>(n) * i] = 1.0;
    auto col = [&](Vec& M, int ld, int j) { return M.data() + static_cast<std::size_t>(ld) * j; };
    auto dot = [](const double* x, const double* y, int len) {
        double s = 0.0;
        for (int i = 0; i < len; ++i) s += x[i] * y[i];
        return s;
    };
    auto rot = [](double* x, double* y, int len, double c, double s) {
        for (int i = 0; i < len; ++i) {
            const double a = x[i], b = y[i];
            x[i] = c * a - s * b;
            y[i] = s * a + c * b;
        }
    };
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p + 1 < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                double* wp = col(W, m, p);
                double* wq = col(W, m, q);
                const double al = dot(wp, wp, m), be = dot(wq, wq, m), ga = dot(wp, wq, m);
                off += ga * ga;
                if (std::fabs(ga) <= 1e-15 * std::sqrt(al * be) || ga == 0.0) continue;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = c * t;
                rot(wp, wq, m, c, s);
                rot(col(Vf, n, p), col(Vf, n, q), n, c, s);
            }
        if (off < 1e-30) break;
    }
    Vec sig(n);
    for (int j = 0; j < n; ++j) sig[j] = std::sqrt(dot(col(W, m, j), col(W, m, j), m));
    std::vector<int> ord(n);
    for (int j = 0; j < n; ++j) ord[j] = j;
    // Selection-style exchange sort on descending sigma (the reference's order).
    for (int i = 0; i + 1 < n; ++i)
        for (int j = i + 1; j < n; ++j)
            if (sig[ord[j]] > sig[ord[i]]) std::swap(ord[i], ord[j]);
    const int k = std::min(m, n);
    Svd out;
    out.S.resize(k);
    out.V.assign(static_cast<std::size_t>(n) * k, 0.0);
    for (int jj = 0; jj < k; ++jj) {
        out.S[jj] = sig[ord[jj]];
        for (int i = 0; i < n; ++i) out.V[i + static_cast<std::size_t>(n) * jj] = Vf[i + static_cast<std::size_t>(n) * ord[jj]];
    }
    return out;
}

// ---- chi-square upper tail (regularised incomplete gamma; textbook series / CF) ------

inline double gamma_pref(double a, double x) { return std::exp(-x + a * std::log(x) - std::lgamma(a)); }

inline double chisq_upper(double x, int dof) {
    if (dof <= 0) return std::nan("");
    if (x <= 0.0) return 1.0;
    const double a = 0.5 * dof, xx = 0.5 * x;
    if (xx < a + 1.0) {
        double ap = a, sum = 1.0 / a, del = sum;
        for (int n = 0; n < 1000; ++n) {
            ap += 1.0;
            del *= xx / ap;
            sum += del;
            if (std::fabs(del) < std::fabs(sum) * 1e-15) break;
        }
        return 1.0 - sum * gamma_pref(a, xx);
    }
    const double tiny = 1e-300;
    double b = xx + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i <= 1000; ++i) {
        const double an = -static_cast<double>(i) * (static_cast<double>(i) - a);
        b += 2.0;
        d = an * d + b;
        if (std::fabs(d) < tiny) d = tiny;
        c = b + an / c;
        if (std::fabs(c) < tiny) c = tiny;
        d = 1.0 / d;
        const double del = d * c;
        h *= del;
        if (std::fabs(del - 1.0) < 1e-15) break;
    }
    return gamma_pref(a, xx) * h;
}

// ---- f4 basis and jackknife ---------------------------------------------------------

struct F4 {
    int nl = 0, nr = 0, nb = 0;
    Vec x_blocks, x_loo, x_total, tot_line;  // k + m*b
    std::vector<int> bsize;
};

inline void loo_and_total(F4& x) {
    const int m = x.nl * x.nr, nb = x.nb;
    x.x_loo.assign(static_cast<std::size_t>(m) * nb, 0.0);
    x.x_total.assign(m, 0.0);
    x.tot_line.assign(m, 0.0);
    if (m <= 0 || nb <= 0) return;
    long double n_ld = 0.0L;
    for (int b = 0; b < nb; ++b) n_ld += static_cast<long double>(x.bsize[b]);
    const double n = static_cast<double>(n_ld);
    for (int k = 0; k < m; ++k) {
        auto X = [&](int b) { return x.x_blocks[k + static_cast<std::size_t>(m) * b]; };
        auto& LOO = x.x_loo;
        long double s = 0.0L;
        for (int b = 0; b < nb; ++b) s += static_cast<long double>(X(b)) * static_cast<long double>(x.bsize[b]);
        const double tot = static_cast<double>(s / n_ld);
        for (int b = 0; b < nb; ++b) {
            const double rel = static_cast<double>(x.bsize[b]) / n;
            LOO[k + static_cast<std::size_t>(m) * b] = (tot - X(b) * rel) / (1.0 - rel);
        }
        long double wn = 0.0L, wd = 0.0L;
        for (int b = 0; b < nb; ++b) {
            const double w = 1.0 - static_cast<double>(x.bsize[b]) / n;
            wn += static_cast<long double>(LOO[k + static_cast<std::size_t>(m) * b]) * static_cast<long double>(w);
            wd += static_cast<long double>(w);
        }
        x.tot_line[k] = static_cast<double>(wn / wd);
        long double ds = 0.0L;
        for (int b = 0; b < nb; ++b)
            ds += static_cast<long double>(x.tot_line[k]) - static_cast<long double>(LOO[k + static_cast<std::size_t>(m) * b]);
        const double term1 = static_cast<double>(ds / static_cast<long double>(nb)) * static_cast<double>(nb);
        long double wb = 0.0L;
        for (int b = 0; b < nb; ++b)
            wb += static_cast<long double>(LOO[k + static_cast<std::size_t>(m) * b]) * static_cast<long double>(x.bsize[b]);
        x.x_total[k] = term1 + static_cast<double>(wb / n_ld);
    }
}

// Drop blocks where some pairs have vpair <= 0 and others do not.
inline std::vector<int> survivors(const double* vpair, int P, int nb) {
    std::vector<int> s;
    const std::size_t slab = static_cast<std::size_t>(P) * P;
    for (int b = 0; b < nb; ++b) {
        bool keep = true;
        if (vpair) {
            bool miss = false, pres = false;
            for (std::size_t e = 0; e < slab; ++e) {
                if (vpair[slab * b + e] > 0.0) pres = true; else miss = true;
                if (miss && pres) break;
            }
            keep = !(miss && pres);
        }
        if (keep) s.push_back(b);
    }
    return s;
}

inline void add_ridge(Vec& a, int n, double fudge) {
    double tr = 0.0;
    for (int k = 0; k < n; ++k) tr += a[k + static_cast<std::size_t>(n) * k];
    for (int k = 0; k < n; ++k) a[k + static_cast<std::size_t>(n) * k] += fudge * tr;
}

struct Cov {
    int m = 0;
    Vec Q, Qinv;
    bool ok = true;
};

inline Cov jackknife_cov(const F4& x, double fudge) {
    Cov c;
    const int m = x.nl * x.nr, nb = x.nb;
    c.m = m;
    if (m <= 0 || nb <= 0) return c;
    long double n_ld = 0.0L;
    for (int b = 0; b < nb; ++b) n_ld += static_cast<long double>(x.bsize[b]);
    const double n = static_cast<double>(n_ld);
    Vec tau(static_cast<std::size_t>(m) * nb);
    for (int b = 0; b < nb; ++b) {
        const double h = n / static_cast<double>(x.bsize[b]);
        const double sq = std::sqrt(h - 1.0);
        for (int k = 0; k < m; ++k)
            tau[k + static_cast<std::size_t>(m) * b] =
                (x.x_total[k] * h - x.x_loo[k + static_cast<std::size_t>(m) * b] * (h - 1.0) - x.tot_line[k]) / sq;
    }
    c.Q.assign(static_cast<std::size_t>(m) * m, 0.0);
    for (int a = 0; a < m; ++a)
        for (int b2 = a; b2 < m; ++b2) {
            long double acc = 0.0L;
            for (int b = 0; b < nb; ++b)
                acc += static_cast<long double>(tau[a + static_cast<std::size_t>(m) * b]) *
                       static_cast<long double>(tau[b2 + static_cast<std::size_t>(m) * b]);
            const double v = static_cast<double>(acc / static_cast<long double>(nb));
            c.Q[a + static_cast<std::size_t>(m) * b2] = c.Q[b2 + static_cast<std::size_t>(m) * a] = v;
        }
    Vec Qf = c.Q;
    add_ridge(Qf, m, fudge);
    c.ok = mat_inverse(Qf, m, c.Qinv);
    if (!c.ok) c.Qinv.assign(static_cast<std::size_t>(m) * m, 0.0);
    return c;
}

}  // namespace cref
