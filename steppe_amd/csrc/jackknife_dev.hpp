// steppe_amd/csrc/jackknife_dev.hpp — the block-jackknife covariance of a set of
// statistics on the device, shared by the qpAdm (f4 basis) and qpGraph (f3 basis)
// kernels.  Follows the reference's compute_loo_and_total and jackknife_cov
// (src/device/cpu/cpu_backend.cpp:2674-2723, 2219-2271) with double-double sums
// where the reference accumulates in long double, then the ridge
// Qf = Q + fudge tr(Q) I (cpu_backend.cpp:2758-2764) and Qf^-1 by an LDS LU with
// partial pivoting (src/core/internal/small_linalg.hpp:117).
#pragma once

#include <cstdint>

#include "wave_linalg.hpp"

namespace stp {

// Workgroup LU with partial pivoting of the m x m matrix in LDS (row i at lu + i*ld,
// thread i owns row i, m <= kThreads), the reference's pivot rule (first row of largest
// magnitude), rows kept in place with logical positions; then inv = A^-1 by one
// thread per column, each running forward and back substitution in the reference's
// order (small_linalg.hpp:117 lu_substitute).  inv may be LDS or global (row i at
// inv + i*ldi).  Returns false (the same on every thread) on an exactly zero pivot.
template <int kThreads>
__device__ bool wg_lu_inverse(double* lu, int ld, int m, double* inv, int ldi, int* s_rowof, double* s_val,
                              int* s_idx) {
    using namespace wl;
    constexpr int kWaves = kThreads / 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const bool act = tid < m;
    int pos = act ? tid : INT_MAX;
    double* my = lu + (act ? tid : 0) * ld;
    for (int k = 0; k < m; ++k) {
        const double cand = (act && pos >= k) ? fabs(my[k]) : -1.0;
        const int l = argmax_pivot(cand, pos);
        if (lane == l) {
            s_val[wave] = cand;
            s_idx[wave] = pos;
            s_idx[kWaves + wave] = tid;
        }
        __syncthreads();
        if (tid == 0) {
            int best = 0;
            for (int w = 1; w < kWaves; ++w)
                if (s_val[w] > s_val[best] || (s_val[w] == s_val[best] && s_idx[w] < s_idx[best])) best = w;
            s_idx[2 * kWaves] = s_idx[kWaves + best];   // pivot row
            s_idx[2 * kWaves + 1] = s_idx[best];        // its position
            s_val[kWaves] = s_val[best];
        }
        __syncthreads();
        const int pr = s_idx[2 * kWaves];
        const int ppos = s_idx[2 * kWaves + 1];
        if (s_val[kWaves] == 0.0) return false;
        if (tid == pr) pos = k;
        else if (pos == k) pos = ppos;
        const double* prow = lu + pr * ld;
        const double rp = 1.0 / prow[k];
        if (act && pos > k) {
            const double f = my[k] * rp;
            my[k] = f;
            for (int j = k + 1; j < m; ++j) my[j] -= f * prow[j];
        }
        __syncthreads();
    }
    if (act) s_rowof[pos] = tid;
    __syncthreads();
    for (int c = tid; c < m; c += kThreads) {
        for (int i = 0; i < m; ++i) {
            const double* li = lu + s_rowof[i] * ld;
            double sum = (s_rowof[i] == c) ? 1.0 : 0.0;
            for (int j = 0; j < i; ++j) sum -= li[j] * inv[j * ldi + c];
            inv[i * ldi + c] = sum;
        }
        for (int i = m - 1; i >= 0; --i) {
            const double* ui = lu + s_rowof[i] * ld;
            double sum = inv[i * ldi + c];
            for (int j = i + 1; j < m; ++j) sum -= ui[j] * inv[j * ldi + c];
            inv[i * ldi + c] = sum / ui[i];
        }
    }
    __syncthreads();
    return true;
}

// Workgroup-level; every thread of the workgroup calls it.  X [m][nbs] holds the
// statistic per surviving block (global) and is overwritten by tau; LOO [m][nbs]
// (global) receives the leave-one-out values.  Outputs: s_xtot / s_totline (m, LDS),
// s_qdiag (m, the unridged diagonal of Q; may be null), sQi (m rows of ldi, LDS or
// global) = Qf^-1.  sLU (m rows of ldq, LDS) is scratch; s_pos (m), s_lane_of
// (2*waves + 2 ints) and s_val (waves + 1 doubles) are LDS scratch.  m <= kThreads.
// Returns false (the same on every thread) when the LU meets an exactly zero pivot.
template <int kThreads>
__device__ bool jk_covariance(double* X, double* LOO, const std::int32_t* bsz, double n_tot, int nbs, int m,
                              double fudge, double* s_xtot, double* s_totline, double* s_qdiag, double* sLU, int ldq,
                              double* sQi, int ldi, int* s_pos, int* s_lane_of, double* s_val, int* s_flag) {
    using namespace wl;
    constexpr int kWaves = kThreads / 64;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // ---- LOO and total (one thread per statistic) ----
    for (int k = tid; k < m; k += kThreads) {
        const double* Xk = X + static_cast<std::int64_t>(k) * nbs;
        double* Lk = LOO + static_cast<std::int64_t>(k) * nbs;
        dd s{0.0, 0.0};
        for (int b = 0; b < nbs; ++b) s = dd_fma(s, Xk[b], static_cast<double>(bsz[b]));
        const double tot = dd_div_d(s, n_tot);
        for (int b = 0; b < nbs; ++b) {
            const double rel = static_cast<double>(bsz[b]) / n_tot;
            Lk[b] = (tot - Xk[b] * rel) / (1.0 - rel);
        }
        dd wn{0.0, 0.0}, wd{0.0, 0.0};
        for (int b = 0; b < nbs; ++b) {
            const double w = 1.0 - static_cast<double>(bsz[b]) / n_tot;
            wn = dd_fma(wn, Lk[b], w);
            wd = dd_add(wd, w);
        }
        const double tl = dd_div_dd(wn, wd);
        dd ds{0.0, 0.0};
        for (int b = 0; b < nbs; ++b) ds = dd_add(dd_add(ds, tl), -Lk[b]);
        const double term1 = dd_div_d(ds, static_cast<double>(nbs)) * static_cast<double>(nbs);
        dd wb{0.0, 0.0};
        for (int b = 0; b < nbs; ++b) wb = dd_fma(wb, Lk[b], static_cast<double>(bsz[b]));
        s_totline[k] = tl;
        s_xtot[k] = term1 + dd_div_d(wb, n_tot);
    }
    __syncthreads();
    // ---- tau, overwriting X ----
    for (int k = wave; k < m; k += kWaves) {
        const double xt = s_xtot[k], tl = s_totline[k];
        for (int b = lane; b < nbs; b += 64) {
            const double h = n_tot / static_cast<double>(bsz[b]);
            const double sq = sqrt(h - 1.0);
            const std::int64_t o = static_cast<std::int64_t>(k) * nbs + b;
            X[o] = (xt * h - LOO[o] * (h - 1.0) - tl) / sq;
        }
    }
    __syncthreads();
    // ---- Q = sum_b tau tau^T / nb (upper triangle, mirrored) into sLU ----
    for (int e = tid; e < m * m; e += kThreads) {
        const int a = e / m, c = e - (e / m) * m;
        if (c < a) continue;
        const double* ta = X + static_cast<std::int64_t>(a) * nbs;
        const double* tc = X + static_cast<std::int64_t>(c) * nbs;
        dd acc{0.0, 0.0};
        for (int b = 0; b < nbs; ++b) acc = dd_fma(acc, ta[b], tc[b]);
        const double v = dd_div_d(acc, static_cast<double>(nbs));
        sLU[a * ldq + c] = v;
        sLU[c * ldq + a] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double tr = 0.0;
        for (int k = 0; k < m; ++k) tr += sLU[k * ldq + k];
        for (int k = 0; k < m; ++k) {
            if (s_qdiag) s_qdiag[k] = sLU[k * ldq + k];
            sLU[k * ldq + k] += fudge * tr;
        }
        *s_flag = 0;
    }
    __syncthreads();
    if (!wg_lu_inverse<kThreads>(sLU, ldq, m, sQi, ldi, s_pos, s_val, s_lane_of)) return false;
    return true;
}

}  // namespace stp
