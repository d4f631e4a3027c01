/* stp_qpadm_result.h — the flat, fixed-size result record of one qpAdm fit.
 *
 * One record carries every field of the reference's steppe::QpAdmResult
 * (include/steppe/qpadm.hpp:50-73) that the Python result dict exposes
 * (bindings/internal/bind_common.hpp:179-212), laid out as plain C arrays so a
 * batch of fits is one contiguous array that a kernel, ctypes and numpy can all
 * address directly.  Capacities bound the model shapes the library accepts:
 * at most STP_QPADM_MAX_NL source populations (left, excluding the target),
 * so at most STP_QPADM_MAX_RANK ranks and STP_QPADM_MAX_POPDROP popdrop rows.
 */
#ifndef STP_QPADM_RESULT_H
#define STP_QPADM_RESULT_H

#include <stdint.h>

#define STP_QPADM_MAX_NL 8
#define STP_QPADM_MAX_RANK STP_QPADM_MAX_NL
#define STP_QPADM_MAX_POPDROP (STP_QPADM_MAX_NL + 1)

#ifdef __cplusplus
extern "C" {
#endif

typedef struct stp_qpadm_result {
    int32_t status;     /* steppe::Status ordinal: 0 ok, 2 rank_deficient, 3 non_spd, 4 chisq_undefined */
    int32_t model_index;
    int32_t nl;         /* number of weights (sources) */
    int32_t dof;
    int32_t est_rank;
    int32_t f4rank;
    int32_t n_rank;     /* ranks swept: r = 0 .. n_rank-1 */
    int32_t n_popdrop;  /* popdrop rows */
    double p;
    double chisq;
    double weight[STP_QPADM_MAX_NL];
    double se[STP_QPADM_MAX_NL];
    double z[STP_QPADM_MAX_NL];
    double rank_chisq[STP_QPADM_MAX_RANK];
    int32_t rank_dof[STP_QPADM_MAX_RANK];
    int32_t rankdrop_f4rank[STP_QPADM_MAX_RANK];
    int32_t rankdrop_dof[STP_QPADM_MAX_RANK];
    int32_t rankdrop_dofdiff[STP_QPADM_MAX_RANK]; /* INT32_MIN where undefined */
    double rankdrop_chisq[STP_QPADM_MAX_RANK];
    double rankdrop_p[STP_QPADM_MAX_RANK];
    double rankdrop_chisqdiff[STP_QPADM_MAX_RANK]; /* NaN where undefined */
    double rankdrop_p_nested[STP_QPADM_MAX_RANK];  /* NaN where undefined */
    char popdrop_pat[STP_QPADM_MAX_POPDROP][STP_QPADM_MAX_NL + 1];
    int32_t popdrop_wt[STP_QPADM_MAX_POPDROP];
    int32_t popdrop_dof[STP_QPADM_MAX_POPDROP];
    int32_t popdrop_f4rank[STP_QPADM_MAX_POPDROP];
    int32_t popdrop_feasible[STP_QPADM_MAX_POPDROP];
    double popdrop_chisq[STP_QPADM_MAX_POPDROP];
    double popdrop_p[STP_QPADM_MAX_POPDROP];
} stp_qpadm_result;

#ifdef __cplusplus
}
#endif

#endif /* STP_QPADM_RESULT_H */
