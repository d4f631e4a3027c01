// oracle/cpu_ref_qpadm.cpp — CPU restatement of the reference's qpAdm fit chain.
//
// TEST INFRASTRUCTURE ONLY (see cpu_ref.cpp): the parity checker for the HIP
// batched qpAdm kernel.  Restated from the reference's semantics, function by
// function (paths relative to the reference repository root):
//
//   survivor blocks       src/device/cpu/cpu_backend.cpp:125-150
//   f4 basis              src/device/cpu/cpu_backend.cpp:2065-2113 (assemble_f4)
//   LOO / total           src/device/cpu/cpu_backend.cpp:2674-2723 (compute_loo_and_total)
//   jackknife covariance  src/device/cpu/cpu_backend.cpp:2219-2271, ridge 2758-2764
//   LU / solve / inverse  src/core/internal/small_linalg.hpp:55-135
//   one-sided Jacobi SVD  src/core/internal/small_linalg.hpp:137-218
//   ALS weights           src/device/cpu/cpu_backend.cpp:2727-2935
//   rank sweep            src/device/cpu/cpu_backend.cpp:2533-2610
//   chi-square tail       src/core/internal/pchisq.hpp:16-70
//   LOO standard errors   src/device/cpu/cpu_backend.cpp:98-122,2625-2668
//   popdrop               src/core/qpadm/ranktest.cpp:25-140
//   fit driver            src/core/qpadm/qpadm_fit.cpp:49-132
//
// The operation order of every floating-point reduction follows the reference
// (plain double sums, or long double where the reference uses it), so results
// agree with the reference to a few ulps; tests pin them against oracle/_ref and
// the ADMIXTOOLS 2 goldens.

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "stp_qpadm_result.h"

#include "cref_common.hpp"

namespace {

using namespace cref;

F4 build_f4(const double* f2, const double* vpair, const int* bsz, int P, int nb, const std::vector<int>& left,
            const std::vector<int>& right) {
    F4 x;
    x.nl = static_cast<int>(left.size()) - 1;
    x.nr = static_cast<int>(right.size()) - 1;
    const int m = x.nl * x.nr;
    if (x.nl <= 0 || x.nr <= 0 || nb <= 0) return x;
    const std::vector<int> surv = survivors(vpair, P, nb);
    x.nb = static_cast<int>(surv.size());
    x.bsize.resize(x.nb);
    for (int s = 0; s < x.nb; ++s) x.bsize[s] = bsz[surv[s]];
    x.x_blocks.assign(static_cast<std::size_t>(m) * x.nb, 0.0);
    const std::size_t slab = static_cast<std::size_t>(P) * P;
    auto F = [&](int i, int j, int b) { return f2[i + static_cast<std::size_t>(P) * j + slab * b]; };
    const int L0 = left[0], R0 = right[0];
    for (int i = 0; i < x.nl; ++i)
        for (int j = 0; j < x.nr; ++j) {
            const int Li = left[i + 1], Rj = right[j + 1];
            const int k = j + x.nr * i;
            for (int s = 0; s < x.nb; ++s) {
                const int b = surv[s];
                x.x_blocks[k + static_cast<std::size_t>(m) * s] =
                    0.5 * (F(Li, R0, b) + F(L0, Rj, b) - F(L0, R0, b) - F(Li, Rj, b));
            }
        }
    loo_and_total(x);
    return x;
}


// ---- ALS ---------------------------------------------------------------------------

// xmat (nl x nr, column-major) from the f4 total vector (k = j + nr*i).
Vec to_xmat(const Vec& tot, int nl, int nr) {
    Vec xm(static_cast<std::size_t>(nl) * nr);
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nr; ++j) xm[i + static_cast<std::size_t>(nl) * j] = tot[j + static_cast<std::size_t>(nr) * i];
    return xm;
}

// Ridge-regularised GLS normal equations for min (x - L c)^T Qinv (x - L c); L is
// given by a callback over (row k in [0,m), column a in [0,t)).
template <class Lop>
bool gls_ridge(const Lop& L, int m, int t, const Vec& xvec, const Vec& qinv, double fudge, Vec& sol) {
    Vec W(static_cast<std::size_t>(m) * t, 0.0);
    for (int kr = 0; kr < m; ++kr)
        for (int a = 0; a < t; ++a) {
            double s = 0.0;
            for (int kc = 0; kc < m; ++kc) s += qinv[kr + static_cast<std::size_t>(m) * kc] * L(kc, a);
            W[kr + static_cast<std::size_t>(m) * a] = s;
        }
    Vec C(static_cast<std::size_t>(t) * t, 0.0), rhs(t, 0.0);
    for (int a = 0; a < t; ++a) {
        for (int c = 0; c < t; ++c) {
            double s = 0.0;
            for (int k = 0; k < m; ++k) s += L(k, a) * W[k + static_cast<std::size_t>(m) * c];
            C[a + static_cast<std::size_t>(t) * c] = s;
        }
        double r = 0.0;
        for (int k = 0; k < m; ++k) r += xvec[k] * W[k + static_cast<std::size_t>(m) * a];
        rhs[a] = r;
    }
    add_ridge(C, t, fudge);
    return lin_solve(C, t, rhs, sol);
}

struct Fit {
    Vec w, A, B;
    double chisq = 0.0;
    bool ok = true;
};

double resid_chisq(const Vec& xm, const Vec& A, const Vec& B, int nl, int nr, int r, const Vec& qinv) {
    const int m = nl * nr;
    Vec e(m);
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nr; ++j) {
            double ab = 0.0;
            for (int p = 0; p < r; ++p) ab += A[i + static_cast<std::size_t>(nl) * p] * B[p + static_cast<std::size_t>(r) * j];
            e[i * nr + j] = xm[i + static_cast<std::size_t>(nl) * j] - ab;
        }
    long double acc = 0.0L;
    for (int a = 0; a < m; ++a) {
        long double row = 0.0L;
        for (int b = 0; b < m; ++b)
            row += static_cast<long double>(qinv[a + static_cast<std::size_t>(m) * b]) * static_cast<long double>(e[b]);
        acc += static_cast<long double>(e[a]) * row;
    }
    return static_cast<double>(acc);
}

Fit als(const Vec& xm, int nl, int nr, int r, const Vec& qinv, double fudge, int iters) {
    Fit f;
    if (r == 0) {
        f.w.assign(nl, 1.0);
        f.chisq = resid_chisq(xm, f.A, f.B, nl, nr, 0, qinv);
        return f;
    }
    const int m = nl * nr;
    Vec xvec(m);
    for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nr; ++j) xvec[i * nr + j] = xm[i + static_cast<std::size_t>(nl) * j];
    // Seed: B from the top r right singular vectors, A = X B^T.
    const Svd sv = jacobi(xm, nl, nr);
    Vec B(static_cast<std::size_t>(r) * nr, 0.0), A(static_cast<std::size_t>(nl) * r, 0.0);
    for (int p = 0; p < r; ++p)
        for (int j = 0; j < nr; ++j) B[p + static_cast<std::size_t>(r) * j] = sv.V[j + static_cast<std::size_t>(nr) * p];
    for (int i = 0; i < nl; ++i)
        for (int p = 0; p < r; ++p) {
            double s = 0.0;
            for (int j = 0; j < nr; ++j) s += xm[i + static_cast<std::size_t>(nl) * j] * B[p + static_cast<std::size_t>(r) * j];
            A[i + static_cast<std::size_t>(nl) * p] = s;
        }
    for (int it = 0; it < iters; ++it) {
        {   // A-step: unknowns A[i][p] at a = i*r + p
            const int t = nl * r;
            auto L = [&](int k, int a) {
                const int i = a / r, p = a % r, ii = k / nr, j = k % nr;
                return i == ii ? B[p + static_cast<std::size_t>(r) * j] : 0.0;
            };
            Vec sol;
            Vec An(static_cast<std::size_t>(nl) * r, 0.0);
            if (gls_ridge(L, m, t, xvec, qinv, fudge, sol))
                for (int i = 0; i < nl; ++i)
                    for (int p = 0; p < r; ++p) An[i + static_cast<std::size_t>(nl) * p] = sol[i * r + p];
            A = An;
        }
        {   // B-step: unknowns B[p][j] at c = p*nr + j
            const int t = r * nr;
            auto L = [&](int k, int c) {
                const int i = k / nr, j = k % nr, p = c / nr, jc = c % nr;
                return j == jc ? A[i + static_cast<std::size_t>(nl) * p] : 0.0;
            };
            Vec sol;
            Vec Bn(static_cast<std::size_t>(r) * nr, 0.0);
            if (gls_ridge(L, m, t, xvec, qinv, fudge, sol))
                for (int p = 0; p < r; ++p)
                    for (int j = 0; j < nr; ++j) Bn[p + static_cast<std::size_t>(r) * j] = sol[p * nr + j];
            B = Bn;
        }
    }
    f.A = A;
    f.B = B;
    // Weights: the least-squares combination of A's columns and the unit vector.
    const int rp = r + 1;
    auto xa = [&](int i, int p) { return p < r ? A[i + static_cast<std::size_t>(nl) * p] : 1.0; };
    Vec M(static_cast<std::size_t>(nl) * nl), rhs(nl);
    for (int i = 0; i < nl; ++i) {
        for (int ip = 0; ip < nl; ++ip) {
            double s = 0.0;
            for (int p = 0; p < rp; ++p) s += xa(i, p) * xa(ip, p);
            M[i + static_cast<std::size_t>(nl) * ip] = s;
        }
        rhs[i] = xa(i, r);
    }
    Vec w;
    if (!lin_solve(M, nl, rhs, w)) { f.ok = false; return f; }
    double sum = 0.0;
    for (double v : w) sum += v;
    f.w.resize(nl);
    for (int i = 0; i < nl; ++i) f.w[i] = w[i] / sum;
    f.chisq = resid_chisq(xm, A, B, nl, nr, r, qinv);
    return f;
}

int dof_of(int nl, int nr, int r) { return (nl - r) * (nr - r); }

struct Sweep {
    Vec chisq, p;
    std::vector<int> dof;
    std::vector<int> rd_f4rank, rd_dof, rd_dofdiff;
    Vec rd_chisq, rd_p, rd_chisqdiff, rd_p_nested;
    int f4rank = 0;
    bool degenerate = false;
};

Sweep rank_sweep(const Vec& tot, int nl, int nr, const Vec& qinv, double alpha, double fudge, int iters) {
    Sweep s;
    const int rmax = std::min(nl, nr) - 1;
    if (rmax < 0) { s.degenerate = true; return s; }
    const Vec xm = to_xmat(tot, nl, nr);
    for (int r = 0; r <= rmax; ++r) {
        const Fit f = als(xm, nl, nr, r, qinv, fudge, iters);
        if (!f.ok) s.degenerate = true;
        s.chisq.push_back(f.chisq);
        s.dof.push_back(dof_of(nl, nr, r));
        s.p.push_back(chisq_upper(f.chisq, s.dof.back()));
    }
    for (int k = 0; k <= rmax; ++k) {
        const int r = rmax - k;
        s.rd_f4rank.push_back(r);
        s.rd_dof.push_back(s.dof[r]);
        s.rd_chisq.push_back(s.chisq[r]);
        s.rd_p.push_back(s.p[r]);
        if (r >= 1) {
            const int dd = s.dof[r - 1] - s.dof[r];
            const double cd = s.chisq[r - 1] - s.chisq[r];
            s.rd_dofdiff.push_back(dd);
            s.rd_chisqdiff.push_back(cd);
            s.rd_p_nested.push_back(chisq_upper(cd, dd));
        } else {
            s.rd_dofdiff.push_back(INT_MIN);
            s.rd_chisqdiff.push_back(std::numeric_limits<double>::quiet_NaN());
            s.rd_p_nested.push_back(std::numeric_limits<double>::quiet_NaN());
        }
    }
    s.f4rank = rmax;
    for (int r = 0; r <= rmax; ++r)
        if (s.p[r] > alpha) { s.f4rank = r; break; }
    return s;
}

bool feasible(const Vec& w) {
    bool any = false;
    for (double v : w) {
        if (std::isnan(v)) continue;
        any = true;
        if (v < 0.0 || v > 1.0) return false;
    }
    return any;
}

template <class T>
void put(T* dst, const std::vector<T>& v, std::size_t cap) {
    for (std::size_t k = 0; k < v.size() && k < cap; ++k) dst[k] = v[k];
}

}  // namespace

extern "C" {

// One qpAdm model (left excludes the target) on a host f2 tensor; jackknife:
// 0 none, 1 feasible-only (SE when the weights are in [0,1]), 2 all.
int cref_qpadm(const double* f2, const double* vpair, const std::int32_t* block_sizes, int P, int nb, int target,
               const std::int32_t* left, int n_left, const std::int32_t* right, int n_right, double fudge,
               int als_iterations, int rank, double rank_alpha, int jackknife, stp_qpadm_result* out) {
    std::memset(out, 0, sizeof(*out));
    std::vector<int> L{target}, R(right, right + n_right);
    for (int i = 0; i < n_left; ++i) L.push_back(left[i]);
    std::vector<int> bsz(block_sizes, block_sizes + nb);
    const F4 x = build_f4(f2, vpair, bsz.data(), P, nb, L, R);
    const int nl = x.nl, nr = x.nr;
    if (nl > STP_QPADM_MAX_NL) return -1;
    const int r = rank < 0 ? nl - 1 : rank;
    out->est_rank = r;
    out->dof = dof_of(nl, nr, r);
    out->nl = 0;
    const Cov cov = jackknife_cov(x, fudge);
    if (!cov.ok) { out->status = 3; return 0; }  // NonSpdCovariance
    const Vec xm = to_xmat(x.x_total, nl, nr);
    const Fit fit = als(xm, nl, nr, r, cov.Qinv, fudge, als_iterations);
    if (!fit.ok) { out->status = 2; return 0; }  // RankDeficient
    out->nl = nl;
    put(out->weight, fit.w, STP_QPADM_MAX_NL);
    out->chisq = fit.chisq;
    out->p = chisq_upper(fit.chisq, out->dof);
    bool do_se = jackknife == 2 || (jackknife == 1 && feasible(fit.w));
    if (do_se && x.nb >= 2) {
        // nb leave-one-block-out refits on x_loo[:, b], scaled by (nb-1)/sqrt(nb),
        // then the sample standard deviation per weight.
        const int m = nl * nr;
        Vec wm(static_cast<std::size_t>(x.nb) * nl, 0.0);
        for (int b = 0; b < x.nb; ++b) {
            Vec tot(x.x_loo.begin() + static_cast<std::size_t>(m) * b, x.x_loo.begin() + static_cast<std::size_t>(m) * (b + 1));
            const Fit fb = als(to_xmat(tot, nl, nr), nl, nr, r, cov.Qinv, fudge, als_iterations);
            for (int i = 0; i < nl; ++i) wm[static_cast<std::size_t>(b) * nl + i] = i < static_cast<int>(fb.w.size()) ? fb.w[i] : 0.0;
        }
        const double scale = static_cast<double>(x.nb - 1) / std::sqrt(static_cast<double>(x.nb));
        for (double& v : wm) v *= scale;
        for (int i = 0; i < nl; ++i) {
            double mean = 0.0;
            for (int b = 0; b < x.nb; ++b) mean += wm[static_cast<std::size_t>(b) * nl + i];
            mean /= x.nb;
            long double acc = 0.0L;
            for (int b = 0; b < x.nb; ++b) {
                const double d = wm[static_cast<std::size_t>(b) * nl + i] - mean;
                acc += static_cast<long double>(d) * static_cast<long double>(d);
            }
            const double se = std::sqrt(static_cast<double>(acc / static_cast<long double>(x.nb - 1)));
            out->se[i] = se;
            out->z[i] = se > 0.0 ? fit.w[i] / se : 0.0;
        }
    }
    // Rank sweep on the full model.
    const Sweep sw = rank_sweep(x.x_total, nl, nr, cov.Qinv, rank_alpha, fudge, als_iterations);
    out->n_rank = static_cast<int>(sw.chisq.size());
    put(out->rank_chisq, sw.chisq, STP_QPADM_MAX_RANK);
    put(out->rank_dof, sw.dof, STP_QPADM_MAX_RANK);
    out->f4rank = sw.f4rank;
    put(out->rankdrop_f4rank, sw.rd_f4rank, STP_QPADM_MAX_RANK);
    put(out->rankdrop_dof, sw.rd_dof, STP_QPADM_MAX_RANK);
    put(out->rankdrop_dofdiff, sw.rd_dofdiff, STP_QPADM_MAX_RANK);
    put(out->rankdrop_chisq, sw.rd_chisq, STP_QPADM_MAX_RANK);
    put(out->rankdrop_p, sw.rd_p, STP_QPADM_MAX_RANK);
    put(out->rankdrop_chisqdiff, sw.rd_chisqdiff, STP_QPADM_MAX_RANK);
    put(out->rankdrop_p_nested, sw.rd_p_nested, STP_QPADM_MAX_RANK);
    // Popdrop: the full model, then each source dropped (last to first).
    std::vector<std::vector<int>> rows;
    std::vector<int> drops;
    {
        std::vector<int> all;
        for (int i = 0; i < nl; ++i) all.push_back(i);
        rows.push_back(all);
        drops.push_back(-1);
        if (nl >= 2)
            for (int d = nl - 1; d >= 0; --d) {
                std::vector<int> s;
                for (int i = 0; i < nl; ++i)
                    if (i != d) s.push_back(i);
                rows.push_back(s);
                drops.push_back(d);
            }
    }
    out->n_popdrop = static_cast<int>(rows.size());
    const int m_full = nl * nr;
    for (std::size_t k = 0; k < rows.size(); ++k) {
        const std::vector<int>& sv = rows[k];
        const int nlr = static_cast<int>(sv.size());
        const int mr = nlr * nr;
        Vec tot(mr);
        std::vector<int> ind(mr);
        for (int ii = 0; ii < nlr; ++ii)
            for (int j = 0; j < nr; ++j) {
                tot[j + nr * ii] = x.x_total[j + nr * sv[ii]];
                ind[j + nr * ii] = j + nr * sv[ii];
            }
        Vec qi(static_cast<std::size_t>(mr) * mr);
        for (int a = 0; a < mr; ++a)
            for (int b = 0; b < mr; ++b) qi[a + static_cast<std::size_t>(mr) * b] = cov.Qinv[ind[a] + static_cast<std::size_t>(m_full) * ind[b]];
        for (int c = 0; c < nl; ++c) out->popdrop_pat[k][c] = '0';
        out->popdrop_pat[k][nl] = '\0';
        if (drops[k] >= 0) out->popdrop_pat[k][drops[k]] = '1';
        out->popdrop_wt[k] = drops[k] >= 0 ? 1 : 0;
        const int rf = nlr - 1;
        const Sweep s2 = rank_sweep(tot, nlr, nr, qi, rank_alpha, fudge, als_iterations);
        out->popdrop_f4rank[k] = rf;
        out->popdrop_dof[k] = rf < static_cast<int>(s2.dof.size()) ? s2.dof[rf] : dof_of(nlr, nr, rf);
        out->popdrop_chisq[k] = rf < static_cast<int>(s2.chisq.size()) ? s2.chisq[rf] : 0.0;
        out->popdrop_p[k] = rf < static_cast<int>(s2.p.size()) ? s2.p[rf] : 0.0;
        Vec wrow(nl, std::numeric_limits<double>::quiet_NaN());
        const Fit f2r = als(to_xmat(tot, nlr, nr), nlr, nr, rf, qi, fudge, als_iterations);
        if (f2r.ok && static_cast<int>(f2r.w.size()) == nlr)
            for (int ii = 0; ii < nlr; ++ii) wrow[sv[ii]] = f2r.w[ii];
        out->popdrop_feasible[k] = feasible(wrow) ? 1 : 0;
    }
    out->status = out->dof <= 0 ? 4 : 0;  // ChisqUndefined
    return 0;
}

}  // extern "C"
