// steppe_amd/csrc/qpadm.hpp — launch interface of the batched qpAdm kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "stp_qpadm_result.h"

namespace stp {

// Model shape limits of the device fit: one f4 entry per lane (m = nl*nr <= 64) and
// one normal-matrix row per lane (nl*r, r*nr <= 64).
constexpr int kMaxQpadmM = 64;
constexpr int kMaxQpadmNr = 32;

struct QpadmArgs {
    const double* f2T;            // survivor-transposed f2: [(i + P*j) * nbs + s]
    const std::int32_t* bsz;      // survivor block sizes [nbs]
    int P = 0, nbs = 0;
    double n_total = 0.0;         // sum of the survivor block sizes
    // model arena: left[left_off[i] .. left_off[i+1]) are model i's sources,
    // right[right_off[i] .. right_off[i+1]) its right set (right[0] is the base)
    const std::int32_t* target;
    const std::int32_t* left_off;
    const std::int32_t* left;
    const std::int32_t* right_off;
    const std::int32_t* right;
    const std::int32_t* order;    // models of this launch, in claim order
    int n_order = 0;
    unsigned* counter;
    double fudge = 1e-4;
    int iters = 20;
    int rank = -1;
    double alpha = 0.05;
    int jackknife = 2;            // 0 none, 1 feasible-only, 2 all
    int se_require_p = 0;
    double p_se_threshold = 0.05;
    double* scratch;              // per workgroup: X/tau [m_max][nbs], LOO [m_max][nbs], wm [nbs][8]
    std::int64_t scratch_stride = 0;
    stp_qpadm_result* out;        // indexed by model
    int m_max = 0, ldq = 0, nr_max = 0, wave_slice = 0;
};

// ---- models beyond the batched kernel (qpadm_large.hip) -----------------------------
// Largest f4 basis (sources x (right - 1)), right set and normal system of the large path.
constexpr int kQpLargeMaxM = 1024;
constexpr int kQpLargeMaxNr = 128;
constexpr int kQpLargeMaxT = 1024;

struct QpLargeModel {
    int mi;                // model index (result record)
    int l0, nl, r0, nr;    // arena offsets and sizes (nr excludes the base)
    int m, rank;
    std::int64_t base;     // offset of the model's setup scratch (doubles)
};

// One fit of a large model: (popdrop row k, rank rr) of its fit table when b < 0, else
// the leave-one-block-out refit without surviving block b.
struct QpLargeTask {
    int slot, k, rr, b;
};

// Per-model fit tables (row k = 0 the full model, then each source dropped).
struct QpLargeTables {
    double chisq[STP_QPADM_MAX_NL + 1][STP_QPADM_MAX_NL + 1];
    int ok[STP_QPADM_MAX_NL + 1][STP_QPADM_MAX_NL + 1];
    double wrf[STP_QPADM_MAX_NL + 1][STP_QPADM_MAX_NL];
    double wmain[STP_QPADM_MAX_NL];
    double prank[STP_QPADM_MAX_NL + 1];
    int cov_ok, pad;
};

struct QpLargeArgs {
    const double* f2T = nullptr;
    const std::int32_t* bsz = nullptr;
    int P = 0, nbs = 0;
    double n_total = 0.0;
    const std::int32_t* target = nullptr;
    const std::int32_t* left = nullptr;
    const std::int32_t* right = nullptr;
    const QpLargeModel* models = nullptr;
    int n_models = 0;
    const QpLargeTask* tasks = nullptr;
    int n_tasks = 0;
    unsigned* counter = nullptr;
    QpLargeTables* tables = nullptr;
    double* setup = nullptr;       // per-model setup scratch
    double* fit = nullptr;         // per-wavefront fit scratch
    std::int64_t fit_stride = 0;
    int m_max = 0, nr_max = 0, t_max = 0;
    double* wm = nullptr;          // leave-one-block-out weights: slot s, block b at s*nbs*8 + b*nl
    stp_qpadm_result* out = nullptr;
    double fudge = 1e-4;
    int iters = 20;
    double alpha = 0.05;
    int jackknife = 2;
    int se_require_p = 0;
    double p_se_threshold = 0.05;
};

std::size_t qpadm_large_setup_doubles(int m, int nbs);
std::size_t qpadm_large_fit_doubles(int m_max, int nr_max, int t_max);
int qpadm_large_fit_waves();
void launch_qpadm_large_setup(const QpLargeArgs& g, hipStream_t st);
void launch_qpadm_large_fits(const QpLargeArgs& g, int n_workgroups, hipStream_t st);
void launch_qpadm_large_finish(const QpLargeArgs& g, hipStream_t st);

void launch_survivors(const double* vpair, std::int64_t slab, int n_block, std::int32_t* keep, hipStream_t st);
void launch_transpose_survivors(const double* f2, std::int64_t slab, const std::int32_t* surv, int nbs, double* f2T,
                                hipStream_t st);
std::size_t qpadm_lds_bytes(int m_max, int ldq, int wave_slice);
int qpadm_wave_slice(int nr_max);
void launch_qpadm(int T, const QpadmArgs& args, int n_workgroups, hipStream_t st);
int qpadm_threads();
// Workgroups of qpadm_kernel<T> that fit on one CU at once with lds bytes of dynamic LDS.
int qpadm_blocks_per_cu(int T, std::size_t lds);

}  // namespace stp
