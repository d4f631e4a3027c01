// steppe_amd/csrc/qpadm_record.hpp — the result record of one qpAdm fit, filled from the
// fit tables of a model; shared by the batched kernel (models within one wavefront's
// register budget, qpadm.hip) and the large-model kernels (qpadm_large.hip).
//
// The fields follow the reference's result assembly: the main fit's weights, chi^2 and
// tail probability (src/core/qpadm/qpadm_fit.cpp:49-132), the rank sweep and rankdrop
// table (src/device/cpu/cpu_backend.cpp:2533-2610), the popdrop rows
// (src/core/qpadm/ranktest.cpp:25-140) and the leave-one-block-out standard errors
// (cpu_backend.cpp:98-122, 2625-2668).  Translation units that include this are built
// without FP contraction.
#pragma once

#include <climits>
#include <cstdint>

#include "stp_qpadm_result.h"
#include "wave_linalg.hpp"

namespace stp {

constexpr int kQpMaxNl = STP_QPADM_MAX_NL;
constexpr int kQpMaxRows = kQpMaxNl + 1;   // popdrop rows
constexpr int kQpMaxRanks = kQpMaxNl + 1;  // fits per row: ranks 0..8

// Weights all in [0, 1] (NaN entries ignored; none at all is infeasible).
__device__ inline bool feasible_w(const double* w, int n) {
    bool any = false;
    for (int i = 0; i < n; ++i) {
        const double v = w[i];
        if (v != v) continue;
        any = true;
        if (v < 0.0 || v > 1.0) return false;
    }
    return any;
}

// Whether a model gets leave-one-block-out standard errors (the reference's jackknife
// policy: 2 all fitted models, 1 only feasible ones, optionally only those with
// p >= p_threshold; 0 none), given its main fit.
__device__ inline bool qpadm_do_se(bool main_ok, int nbs, int jackknife, const double* wmain, int nl,
                                   bool se_require_p, double chisq_main, int dof, double p_threshold) {
    if (!main_ok || nbs < 2) return false;
    if (jackknife == 2) return true;
    if (jackknife == 1)
        return feasible_w(wmain, nl) && (!se_require_p || wl::chisq_upper(chisq_main, dof) >= p_threshold);
    return false;
}

// Standard error of weight `i` from its leave-one-block-out refits wm[b * nl + i]:
// the refits are scaled by (nbs - 1) / sqrt(nbs), then the sample standard deviation
// (mean in plain double, squares summed in double-double).
__device__ inline double loo_se(const double* wm, int nbs, int nl, int i) {
    using namespace wl;
    const double scale = static_cast<double>(nbs - 1) / sqrt(static_cast<double>(nbs));
    double mean = 0.0;
    for (int b = 0; b < nbs; ++b) mean += wm[static_cast<std::int64_t>(b) * nl + i] * scale;
    mean /= nbs;
    dd acc{0.0, 0.0};
    for (int b = 0; b < nbs; ++b) {
        const double d = wm[static_cast<std::int64_t>(b) * nl + i] * scale - mean;
        acc = dd_fma(acc, d, d);
    }
    return sqrt(dd_div_d(acc, static_cast<double>(nbs - 1)));
}

// Fills the record of model `mi` from its fit tables; called by one whole wavefront.
//   chisq[k][r], ok[k][r]: fit of popdrop row k (0 = the full model) at rank r
//   wrf[k][i]:             weights of row k's fit at rank (sources of row k) - 1
//   wmain[i]:              weights of the full model's fit at `rank`
//   se:                    lane i < nl: its weight's standard error (when do_se)
//   prank:                 kQpMaxRanks doubles of scratch (LDS or global)
// Lane 0 writes the record fields; lanes < nl write their SE and z.
__device__ inline void write_qpadm_record(stp_qpadm_result* out, int mi, int nl, int nr, int rank, double alpha,
                                          const double (*chisq)[kQpMaxRanks], const int (*ok)[kQpMaxRanks],
                                          const double (*wrf)[kQpMaxNl], const double* wmain, bool main_ok,
                                          bool do_se, double se, double* prank) {
    using namespace wl;
    const int lane = lane_id();
    const int dof = dof_of(nl, nr, rank);
    const int nrows = nl >= 2 ? nl + 1 : 1;
    const int rmax = min(nl, nr) - 1;
    // p of each swept rank, one lane per rank
    if (main_ok && lane <= rmax) prank[lane] = chisq_upper(chisq[0][lane], dof_of(nl, nr, lane));
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        stp_qpadm_result* o = out;
        o->status = 0;
        o->model_index = mi;
        o->est_rank = rank;
        o->dof = dof;
        o->nl = 0;
        o->n_rank = 0;
        o->n_popdrop = 0;
        o->f4rank = 0;
        o->p = 0.0;
        o->chisq = 0.0;
        const bool lanes_write_se = main_ok && do_se;
        for (int i = 0; i < kQpMaxNl; ++i) {
            o->weight[i] = 0.0;
            if (!lanes_write_se || i >= nl) {
                o->se[i] = 0.0;
                o->z[i] = 0.0;
            }
        }
        if (!main_ok) {
            o->status = 2;  // RankDeficient
        } else {
            o->nl = nl;
            o->chisq = chisq[0][rank];
            o->p = chisq_upper(chisq[0][rank], dof);
            for (int i = 0; i < nl; ++i) o->weight[i] = wmain[i];
            // rank sweep of the full model
            o->n_rank = rmax + 1;
            int f4 = rmax;
            for (int rr = rmax; rr >= 0; --rr)
                if (prank[rr] > alpha) f4 = rr;
            o->f4rank = f4;
            for (int rr = 0; rr <= rmax; ++rr) {
                o->rank_chisq[rr] = chisq[0][rr];
                o->rank_dof[rr] = dof_of(nl, nr, rr);
            }
            for (int kk = 0; kk <= rmax; ++kk) {
                const int rr = rmax - kk;
                o->rankdrop_f4rank[kk] = rr;
                o->rankdrop_dof[kk] = dof_of(nl, nr, rr);
                o->rankdrop_chisq[kk] = chisq[0][rr];
                o->rankdrop_p[kk] = prank[rr];
                if (rr >= 1) {
                    const int ddf = dof_of(nl, nr, rr - 1) - dof_of(nl, nr, rr);
                    const double cd = chisq[0][rr - 1] - chisq[0][rr];
                    o->rankdrop_dofdiff[kk] = ddf;
                    o->rankdrop_chisqdiff[kk] = cd;
                    o->rankdrop_p_nested[kk] = chisq_upper(cd, ddf);
                } else {
                    o->rankdrop_dofdiff[kk] = INT_MIN;
                    o->rankdrop_chisqdiff[kk] = __builtin_nan("");
                    o->rankdrop_p_nested[kk] = __builtin_nan("");
                }
            }
            // popdrop rows: the full model, then each source dropped (last first)
            o->n_popdrop = nrows;
            for (int k = 0; k < nrows; ++k) {
                const int drop = k == 0 ? -1 : nl - k;
                const int nlr = k == 0 ? nl : nl - 1;
                const int rmx = min(nlr, nr) - 1, rf = nlr - 1;
                for (int c = 0; c < nl; ++c) o->popdrop_pat[k][c] = (c == drop) ? '1' : '0';
                o->popdrop_pat[k][nl] = '\0';
                o->popdrop_wt[k] = drop >= 0 ? 1 : 0;
                o->popdrop_f4rank[k] = rf;
                const int pdof = dof_of(nlr, nr, rf);
                o->popdrop_dof[k] = pdof;
                o->popdrop_chisq[k] = rf <= rmx ? chisq[k][rf] : 0.0;
                o->popdrop_p[k] = rf <= rmx ? chisq_upper(chisq[k][rf], pdof) : 0.0;
                // feasible: every weight of the row's fit in [0, 1] (the dropped
                // source is NaN and ignored); a failed fit is infeasible
                bool feas = ok[k][rf] != 0 && nlr > 0;
                if (feas)
                    for (int ii = 0; ii < nlr; ++ii) {
                        const double v = wrf[k][ii];
                        if (!(v >= 0.0 && v <= 1.0)) feas = false;
                    }
                o->popdrop_feasible[k] = feas ? 1 : 0;
            }
            o->status = dof <= 0 ? 4 : 0;  // ChisqUndefined
        }
    }
    if (main_ok && do_se && lane < nl) {
        // each lane owns its weight's SE and z (no other lane writes these fields)
        out->se[lane] = se;
        out->z[lane] = se > 0.0 ? wmain[lane] / se : 0.0;
    }
}

}  // namespace stp
