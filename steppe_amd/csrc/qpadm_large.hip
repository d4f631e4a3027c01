// steppe_amd/csrc/qpadm_large.hip — qpAdm fits of models too large for the batched
// kernel's one-row-per-lane registers (qpadm.hip handles sources x (right - 1) <= 64
// with normal systems of at most 64 unknowns).
//
// The reference fits such models on its large path (src/core/qpadm/model_search.cpp:
// 85-106, fit_models_batched_default -> run_qpadm_impl, src/core/qpadm/qpadm_fit.cpp:
// 49-132) with the same chain as every other model.  Here the chain is split across
// three kernels so that the work of one large model spreads over the whole device:
//
//   qpl_setup_kernel   one 1024-thread workgroup per model: the f4 basis over the
//                      surviving blocks, LOO / total, the jackknife covariance, its
//                      ridge and inverse (jackknife_dev.hpp, with every matrix in global
//                      memory), and a transposed copy of the inverse
//   qpl_fit_kernel     one wavefront per fit: the rank-sweep and popdrop fits of the
//                      model's fit table, or one leave-one-block-out refit; the ALS
//                      normal matrices (up to 1024 unknowns) live in the wavefront's
//                      global scratch, column-major, rows strided over the lanes
//   qpl_finish_kernel  one wavefront per model: leave-one-block-out standard errors and
//                      the result record (qpadm_record.hpp)
//
// The arithmetic of every fit is the batched kernel's, operation for operation (the
// same sums in the same order, the same pivot rule, column-wise back substitution),
// so a model gives the same numbers on either path.  Built without FP contraction.

#include <climits>
#include <cmath>

#include "common.hpp"
#include "jackknife_dev.hpp"
#include "qpadm.hpp"
#include "qpadm_record.hpp"
#include "stp_qpadm_result.h"
#include "wave_linalg.hpp"

namespace stp {

namespace {

using namespace wl;

constexpr int kSetupThreads = 1024;
constexpr int kFitWaves = 4;
constexpr int kMaxU = kQpLargeMaxT / 64;   // rows of a normal matrix per lane

// Global-memory writes of one lane read by other lanes of the same wavefront.
__device__ inline void wsync() {
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

__device__ inline int src_of(int i, int drop) { return (drop >= 0 && i >= drop) ? i + 1 : i; }

struct Shape {
    int nl, nr, r, drop;  // drop: the full model's source removed (-1: none)
};

// One wavefront's fit scratch (global memory).
struct FitScratch {
    double *x, *y, *wc, *e, *rh, *rlo, *a, *b, *ws, *v, *sig, *wt, *C;
    int* ord;
    int ldc;
};

__device__ FitScratch fit_scratch(double* base, const QpLargeArgs& g) {
    FitScratch W;
    const int m = g.m_max, nr = g.nr_max;
    double* p = base;
    W.x = p; p += m;
    W.y = p; p += m;
    W.wc = p; p += m;
    W.e = p; p += m;
    W.rh = p; p += m;
    W.rlo = p; p += m;
    W.a = p; p += kQpMaxNl * kQpMaxNl;
    W.b = p; p += kQpMaxNl * nr;
    W.ws = p; p += kQpMaxNl * nr;
    W.v = p; p += nr * nr;
    W.sig = p; p += nr;
    W.ord = reinterpret_cast<int*>(p); p += nr;
    W.wt = p; p += kQpMaxNl;
    W.C = p;
    W.ldc = g.t_max;
    return W;
}

// LU solve of the n x n system C (column-major, ldc; row a of the lane-strided rows
// a = lane + 64u) with right-hand sides rhs[u]: the batched kernel's lu_solve_rows with
// the rows in global memory.  Returns false for an exactly zero pivot; x[u] is unknown
// a = lane + 64u.
__device__ bool lu_solve_strided(double* C, int ldc, int n, double (&rhs)[kMaxU], double (&x)[kMaxU]) {
    const int lane = lane_id();
    int pos[kMaxU];
#pragma unroll
    for (int u = 0; u < kMaxU; ++u) {
        const int a = lane + 64 * u;
        pos[u] = a < n ? a : INT_MAX;
        x[u] = 0.0;
    }
    for (int k = 0; k < n; ++k) {
        const double* ck = C + static_cast<std::int64_t>(k) * ldc;
        double best = -1.0;
        int bpos = INT_MAX, brow = 0;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u) {
            const int a = lane + 64 * u;
            if (a < n && pos[u] >= k) {
                const double v = fabs(ck[a]);
                if (v > best || (v == best && pos[u] < bpos)) {
                    best = v;
                    bpos = pos[u];
                    brow = a;
                }
            }
        }
        const int pl = argmax_pivot(best, bpos);
        const int prow = rli(brow, pl);
        const int ppos = rli(bpos, pl);
        const double pv = uni(ck[prow]);
        if (pv == 0.0) return false;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u) {
            const int a = lane + 64 * u;
            if (a == prow) pos[u] = k;
            else if (pos[u] == k) pos[u] = ppos;
        }
        const double rp = 1.0 / pv;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u) {
            const int a = lane + 64 * u;
            if (a < n && pos[u] > k) {
                const double f = ck[a] * rp;
                C[static_cast<std::int64_t>(k) * ldc + a] = f;
                for (int j = k + 1; j < n; ++j) {
                    double* cj = C + static_cast<std::int64_t>(j) * ldc;
                    cj[a] -= f * cj[prow];
                }
            }
        }
        wsync();
    }
    // forward substitution (unit lower), the right-hand side moves with its row
    for (int j = 0; j < n; ++j) {
        double mine = 0.0;
        bool has = false;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u)
            if (pos[u] == j) {
                mine = rhs[u];
                has = true;
            }
        const double yj = rl(mine, first_lane(has));
        const double* cj = C + static_cast<std::int64_t>(j) * ldc;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u) {
            const int a = lane + 64 * u;
            if (a < n && pos[u] > j) rhs[u] -= cj[a] * yj;
        }
    }
    // back substitution, column by column
    for (int ii = n - 1; ii >= 0; --ii) {
        double mine = 0.0;
        int row = 0;
        bool has = false;
#pragma unroll
        for (int u = 0; u < kMaxU; ++u)
            if (pos[u] == ii) {
                mine = rhs[u];
                row = lane + 64 * u;
                has = true;
            }
        const int li = first_lane(has);
        const double* ci = C + static_cast<std::int64_t>(ii) * ldc;
        const double xi = rl(mine, li) / ci[rli(row, li)];
#pragma unroll
        for (int u = 0; u < kMaxU; ++u) {
            const int a = lane + 64 * u;
            if (a < n && pos[u] < ii) rhs[u] -= ci[a] * xi;
            if (a == ii) x[u] = xi;
        }
    }
    return true;
}

// Residual chi^2 = e^T Qsub e, e = x - (A B): row sums over (i, j) then the sum over a,
// both in double-double (the batched kernel's resid_chisq).
__device__ double resid_chisq_large(const double* QT, int ldq, Shape s, const FitScratch& W) {
    const int lane = lane_id();
    const int m = s.nl * s.nr;
    for (int k = lane; k < m; k += 64) {
        const int i = k / s.nr, j = k - (k / s.nr) * s.nr;
        double ab = 0.0;
        for (int p = 0; p < s.r; ++p) ab += W.a[i * s.r + p] * W.b[p * s.nr + j];
        W.e[k] = W.x[k] - ab;
    }
    wsync();
    for (int k = lane; k < m; k += 64) {
        const int ia = k / s.nr, ja = k - ia * s.nr;
        const int R = src_of(ia, s.drop) * s.nr + ja;
        dd row{0.0, 0.0};
        for (int i = 0; i < s.nl; ++i) {
            const int cb = src_of(i, s.drop) * s.nr;
            for (int j = 0; j < s.nr; ++j) row = dd_fma(row, QT[static_cast<std::int64_t>(cb + j) * ldq + R], W.e[i * s.nr + j]);
        }
        W.rh[k] = row.hi;
        W.rlo[k] = row.lo;
    }
    wsync();
    dd acc{0.0, 0.0};
    for (int a = 0; a < m; ++a) acc = dd_fma_dd(acc, W.e[a], dd{W.rh[a], W.rlo[a]});
    wsync();
    return uni(acc.hi + acc.lo);
}

// The ALS fit of one (reduced) model at rank r on W.x: the batched kernel's als_fit with
// the normal matrices in global memory.  Returns ok; the weights (nl) go to w_out and,
// when want_chisq is set, the chi^2 to chisq (the leave-one-block-out refits use only
// the weights).
__device__ bool als_fit_large(const double* Q, const double* QT, int ldq_arg, Shape s_arg, int iters_arg,
                              double fudge_arg, const FitScratch& W, double* w_out, double& chisq,
                              bool want_chisq_arg) {
    const int lane = lane_id();
    // The compiler keeps this function out of line, where arguments arrive in vector
    // registers; moved to scalar registers, the sizes give scalar loop control.
    const int ldq = __builtin_amdgcn_readfirstlane(ldq_arg);
    const Shape s{__builtin_amdgcn_readfirstlane(s_arg.nl), __builtin_amdgcn_readfirstlane(s_arg.nr),
                  __builtin_amdgcn_readfirstlane(s_arg.r), __builtin_amdgcn_readfirstlane(s_arg.drop)};
    const int iters = __builtin_amdgcn_readfirstlane(iters_arg);
    const double fudge = uni(fudge_arg);
    const bool want_chisq = __builtin_amdgcn_readfirstlane(want_chisq_arg ? 1 : 0) != 0;
    const int nl = s.nl, nr = s.nr, r = s.r, m = nl * nr;
    chisq = 0.0;
    if (r == 0) {
        if (lane < nl) w_out[lane] = 1.0;
        wsync();
        if (want_chisq) chisq = resid_chisq_large(QT, ldq, s, W);
        return true;
    }
    // y = Qsub^T x (column kc of Qsub against x, k ascending)
    for (int kc = lane; kc < m; kc += 64) {
        const int ic = kc / nr, jc = kc - (kc / nr) * nr;
        const int col = src_of(ic, s.drop) * nr + jc;
        double acc = 0.0;
        for (int i = 0; i < nl; ++i) {
            const int rb = src_of(i, s.drop) * nr;
            for (int j = 0; j < nr; ++j) acc += Q[static_cast<std::int64_t>(rb + j) * ldq + col] * W.x[i * nr + j];
        }
        W.y[kc] = acc;
    }
    // ---- seed: one-sided Jacobi SVD of xm (nl x nr) on its columns ----
    for (int k = lane; k < m; k += 64) {
        const int i = k / nr, j = k - (k / nr) * nr;
        W.ws[i + nl * j] = W.x[k];
    }
    for (int e = lane; e < nr * nr; e += 64) W.v[e] = (e % nr == e / nr) ? 1.0 : 0.0;
    wsync();
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0;
        for (int p = 0; p + 1 < nr; ++p)
            for (int q = p + 1; q < nr; ++q) {
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int i = 0; i < nl; ++i) {
                    const double wp = W.ws[i + nl * p], wq = W.ws[i + nl * q];
                    al += wp * wp;
                    be += wq * wq;
                    ga += wp * wq;
                }
                al = uni(al);
                be = uni(be);
                ga = uni(ga);
                off += ga * ga;
                if (fabs(ga) <= 1e-15 * sqrt(al * be) || ga == 0.0) continue;
                const double zeta = (be - al) / (2.0 * ga);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t);
                const double sn = c * t;
                if (lane < nl) {
                    const double x0 = W.ws[lane + nl * p], y0 = W.ws[lane + nl * q];
                    W.ws[lane + nl * p] = c * x0 - sn * y0;
                    W.ws[lane + nl * q] = sn * x0 + c * y0;
                }
                for (int jj = lane; jj < nr; jj += 64) {
                    const double x0 = W.v[jj + nr * p], y0 = W.v[jj + nr * q];
                    W.v[jj + nr * p] = c * x0 - sn * y0;
                    W.v[jj + nr * q] = sn * x0 + c * y0;
                }
                wsync();
            }
        if (uni(off) < 1e-30) break;
    }
    for (int jj = lane; jj < nr; jj += 64) {
        double ss = 0.0;
        for (int i = 0; i < nl; ++i) ss += W.ws[i + nl * jj] * W.ws[i + nl * jj];
        W.sig[jj] = sqrt(ss);
        W.ord[jj] = jj;
    }
    wsync();
    // the reference's exchange sort on descending sigma; positions 0..r-1 are final
    // after r outer passes
    if (lane == 0) {
        for (int i = 0; i < r; ++i)
            for (int j = i + 1; j < nr; ++j) {
                const int oi = W.ord[i], oj = W.ord[j];
                if (W.sig[oj] > W.sig[oi]) {
                    W.ord[i] = oj;
                    W.ord[j] = oi;
                }
            }
    }
    wsync();
    for (int k = lane; k < r * nr; k += 64) {
        const int p = k / nr, j = k - (k / nr) * nr;
        W.b[k] = W.v[j + nr * W.ord[p]];
    }
    wsync();
    for (int k = lane; k < nl * r; k += 64) {
        const int i = k / r, p = k - (k / r) * r;
        double acc = 0.0;
        for (int j = 0; j < nr; ++j) acc += W.x[i * nr + j] * W.b[p * nr + j];
        W.a[k] = acc;
    }
    wsync();

    for (int it = 0; it < iters; ++it) {
        // step 0, the A-step: unknowns A[i][p] at a = i*r + p (L(k=(i',j), a) = [i'==i] B[p][j]);
        // step 1, the B-step: unknowns B[p][j] at a = p*nr + j (L(k=(i,j'), a) = [j'==j] A[i][p]).
        for (int step = 0; step < 2; ++step) {
            const bool bstep = step == 1;
            const int t = bstep ? r * nr : nl * r;
            const int div = bstep ? nr : r;
            for (int c = 0; c < t; ++c) {
                const int uc = c / div, vc = c - (c / div) * div;
                // W[:, c] = Qsub L[:, c]
                for (int k = lane; k < m; k += 64) {
                    const int qrow = src_of(k / nr, s.drop) * nr + (k - (k / nr) * nr);
                    double acc = 0.0;
                    if (!bstep) {
                        const int cb = src_of(uc, s.drop) * nr;
                        const double* bb = W.b + vc * nr;
                        for (int j = 0; j < nr; ++j) acc += QT[static_cast<std::int64_t>(cb + j) * ldq + qrow] * bb[j];
                    } else {
                        for (int i = 0; i < nl; ++i)
                            acc += QT[static_cast<std::int64_t>(src_of(i, s.drop) * nr + vc) * ldq + qrow] * W.a[i * r + uc];
                    }
                    W.wc[k] = acc;
                }
                wsync();
                // C[a][c] = L[:, a]^T W[:, c] for this lane's rows
                double* cc = W.C + static_cast<std::int64_t>(c) * W.ldc;
#pragma unroll
                for (int u = 0; u < kMaxU; ++u) {
                    const int a = lane + 64 * u;
                    if (a < t) {
                        const int ua = a / div, va = a - (a / div) * div;
                        double cv = 0.0;
                        if (!bstep) {
                            const double* bb = W.b + va * nr;
                            const double* ww = W.wc + ua * nr;
                            for (int j = 0; j < nr; ++j) cv += bb[j] * ww[j];
                        } else {
                            for (int i = 0; i < nl; ++i) cv += W.a[i * r + ua] * W.wc[i * nr + va];
                        }
                        cc[a] = cv;
                    }
                }
                wsync();
            }
            double tr = 0.0;
            for (int a = 0; a < t; ++a) tr += W.C[static_cast<std::int64_t>(a) * W.ldc + a];
            const double ft = fudge * uni(tr);
            double rhs[kMaxU], sol[kMaxU];
#pragma unroll
            for (int u = 0; u < kMaxU; ++u) {
                const int a = lane + 64 * u;
                rhs[u] = 0.0;
                if (a < t) {
                    W.C[static_cast<std::int64_t>(a) * W.ldc + a] += ft;
                    const int ua = a / div, va = a - (a / div) * div;
                    if (!bstep) {
                        const double* bb = W.b + va * nr;
                        const double* yy = W.y + ua * nr;
                        for (int j = 0; j < nr; ++j) rhs[u] += bb[j] * yy[j];
                    } else {
                        for (int i = 0; i < nl; ++i) rhs[u] += W.a[i * r + ua] * W.y[i * nr + va];
                    }
                }
            }
            wsync();
            const bool ok = lu_solve_strided(W.C, W.ldc, t, rhs, sol);
            double* dst = bstep ? W.b : W.a;
#pragma unroll
            for (int u = 0; u < kMaxU; ++u) {
                const int a = lane + 64 * u;
                if (a < t) dst[a] = ok ? sol[u] : 0.0;
            }
            wsync();
        }
    }
    // ---- weights: (Xa Xa^T) w = Xa[:, r], Xa = [A, 1], normalised to sum 1 ----
    {
        double row[kQpMaxNl];
#pragma unroll
        for (int i2 = 0; i2 < kQpMaxNl; ++i2) {
            double acc = 0.0;
            if (lane < nl && i2 < nl) {
                for (int p = 0; p < r; ++p) acc += W.a[lane * r + p] * W.a[i2 * r + p];
                acc += 1.0 * 1.0;
            }
            row[i2] = acc;
        }
        double wv = 0.0;
        if (!lu_solve_rows<kQpMaxNl>(row, 1.0, nl, wv)) return false;
        double sum = 0.0;
        for (int i = 0; i < nl; ++i) sum += rl(wv, i);
        if (lane < nl) w_out[lane] = wv / sum;
    }
    wsync();
    if (want_chisq) chisq = resid_chisq_large(QT, ldq, s, W);
    return true;
}

// Setup scratch of one model: X/tau and LOO [m][nbs], the LU and the inverse and its
// transpose [m][m], x_total and tot_line [m], the pivot rows [m] (ints).
struct SetupPtrs {
    double *X, *LOO, *LU, *Qi, *QT, *xtot, *totline;
    int* pos;
};

__device__ SetupPtrs setup_ptrs(const QpLargeArgs& g, const QpLargeModel& md) {
    SetupPtrs p;
    const std::int64_t mn = static_cast<std::int64_t>(md.m) * g.nbs, mm = static_cast<std::int64_t>(md.m) * md.m;
    p.X = g.setup + md.base;
    p.LOO = p.X + mn;
    p.LU = p.LOO + mn;
    p.Qi = p.LU + mm;
    p.QT = p.Qi + mm;
    p.xtot = p.QT + mm;
    p.totline = p.xtot + md.m;
    p.pos = reinterpret_cast<int*>(p.totline + md.m);
    return p;
}

__global__ __launch_bounds__(kSetupThreads) void qpl_setup_kernel(QpLargeArgs g) {
    __shared__ int s_L[kQpMaxNl + 1];
    __shared__ int s_R[kQpLargeMaxNr + 1];
    __shared__ int s_lane_of[2 * (kSetupThreads / 64) + 2];
    __shared__ double s_val[kSetupThreads / 64 + 1];
    __shared__ int s_flag;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int kWaves = kSetupThreads / 64;
    const QpLargeModel md = g.models[blockIdx.x];
    const int nl = md.nl, nr = md.nr, m = md.m, nbs = g.nbs;
    if (tid == 0) s_L[0] = g.target[md.mi];
    if (tid < nl) s_L[tid + 1] = g.left[md.l0 + tid];
    if (tid <= nr) s_R[tid] = g.right[md.r0 + tid];
    __syncthreads();
    const SetupPtrs sp = setup_ptrs(g, md);
    {
        const int P = g.P;
        const int L0 = s_L[0], R0 = s_R[0];
        const double* f00 = g.f2T + (L0 + static_cast<std::int64_t>(P) * R0) * nbs;
        for (int k = wave; k < m; k += kWaves) {
            const int i = k / nr, j = k - (k / nr) * nr;
            const int Li = s_L[i + 1], Rj = s_R[j + 1];
            const double* fi0 = g.f2T + (Li + static_cast<std::int64_t>(P) * R0) * nbs;
            const double* f0j = g.f2T + (L0 + static_cast<std::int64_t>(P) * Rj) * nbs;
            const double* fij = g.f2T + (Li + static_cast<std::int64_t>(P) * Rj) * nbs;
            for (int b = lane; b < nbs; b += 64)
                sp.X[static_cast<std::int64_t>(k) * nbs + b] = 0.5 * (fi0[b] + f0j[b] - f00[b] - fij[b]);
        }
    }
    __threadfence_block();
    __syncthreads();
    const bool ok = jk_covariance<kSetupThreads>(sp.X, sp.LOO, g.bsz, g.n_total, nbs, m, g.fudge, sp.xtot,
                                                 sp.totline, nullptr, sp.LU, m, sp.Qi, m, sp.pos, s_lane_of, s_val,
                                                 &s_flag);
    if (ok) {
        __threadfence_block();
        __syncthreads();
        for (std::int64_t e = tid; e < static_cast<std::int64_t>(m) * m; e += kSetupThreads) {
            const int rr = static_cast<int>(e / m), cc = static_cast<int>(e - (e / m) * m);
            sp.QT[static_cast<std::int64_t>(cc) * m + rr] = sp.Qi[e];
        }
    }
    if (tid == 0) g.tables[blockIdx.x].cov_ok = ok ? 1 : 0;
}

// Two wavefronts per SIMD (registers capped, the rest spilled to scratch): 24 models of
// 3-8 sources and 25-40 right populations fit in 6.9 s against 8.5 s uncapped (one per
// SIMD) and 9.4 s with three.
__global__ __launch_bounds__(64 * kFitWaves) __attribute__((amdgpu_waves_per_eu(2))) void qpl_fit_kernel(QpLargeArgs g) {
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const FitScratch W = fit_scratch(g.fit + static_cast<std::int64_t>(blockIdx.x * kFitWaves + wave) * g.fit_stride, g);
    for (;;) {
        int ti = 0;
        if (lane == 0) ti = static_cast<int>(atomicAdd(g.counter, 1u));
        ti = __builtin_amdgcn_readfirstlane(__shfl(ti, 0));
        if (ti >= g.n_tasks) break;
        const QpLargeTask tk = g.tasks[ti];
        const QpLargeModel md = g.models[tk.slot];
        QpLargeTables* tb = g.tables + tk.slot;
        if (!tb->cov_ok) continue;
        const SetupPtrs sp = setup_ptrs(g, md);
        const int nl = md.nl, nr = md.nr, rank = md.rank;
        if (tk.b >= 0) {
            // leave-one-block-out refit of the full model at `rank`, when the jackknife
            // policy gives this model standard errors
            const bool main_ok = tb->ok[0][rank] != 0;
            const double* wmain = rank == nl - 1 ? tb->wrf[0] : tb->wmain;
            if (!qpadm_do_se(main_ok, g.nbs, g.jackknife, wmain, nl, g.se_require_p != 0, tb->chisq[0][rank],
                             dof_of(nl, nr, rank), g.p_se_threshold))
                continue;
            const int m = nl * nr;
            for (int k = lane; k < m; k += 64) W.x[k] = sp.LOO[static_cast<std::int64_t>(k) * g.nbs + tk.b];
            wsync();
            double chisq = 0.0;
            const bool ok = als_fit_large(sp.Qi, sp.QT, md.m, Shape{nl, nr, rank, -1}, g.iters, g.fudge, W,
                                          W.wt, chisq, false);
            wsync();
            double* wm = g.wm + static_cast<std::int64_t>(tk.slot) * g.nbs * kQpMaxNl + static_cast<std::int64_t>(tk.b) * nl;
            if (lane < nl) wm[lane] = ok ? W.wt[lane] : 0.0;
            continue;
        }
        const int k = tk.k, rr = tk.rr;
        const Shape sh{k == 0 ? nl : nl - 1, nr, rr, k == 0 ? -1 : nl - k};
        const int mr = sh.nl * nr;
        for (int kk = lane; kk < mr; kk += 64) {
            const int i = kk / nr, j = kk - (kk / nr) * nr;
            W.x[kk] = sp.xtot[src_of(i, sh.drop) * nr + j];
        }
        wsync();
        const int rf = sh.nl - 1;
        double* wout = rr == rf ? tb->wrf[k] : (k == 0 && rr == rank ? tb->wmain : W.wt);
        double chisq = 0.0;
        const bool ok = als_fit_large(sp.Qi, sp.QT, md.m, sh, g.iters, g.fudge, W, wout, chisq, true);
        if (lane == 0) {
            tb->chisq[k][rr] = ok ? chisq : 0.0;
            tb->ok[k][rr] = ok ? 1 : 0;
        }
        wsync();
    }
}

__global__ __launch_bounds__(64 * kFitWaves) void qpl_finish_kernel(QpLargeArgs g) {
    const int lane = threadIdx.x & 63;
    const int slot = static_cast<int>(blockIdx.x) * kFitWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (slot >= g.n_models) return;
    const QpLargeModel md = g.models[slot];
    QpLargeTables* tb = g.tables + slot;
    stp_qpadm_result* out = g.out + md.mi;
    const int nl = md.nl, nr = md.nr, rank = md.rank;
    if (!tb->cov_ok) {
        if (lane == 0) {
            out->status = 3;  // NonSpdCovariance
            out->model_index = md.mi;
            out->nl = 0;
            out->est_rank = rank;
            out->dof = dof_of(nl, nr, rank);
            out->n_rank = 0;
            out->n_popdrop = 0;
        }
        return;
    }
    const bool main_ok = tb->ok[0][rank] != 0;
    const double* wmain = rank == nl - 1 ? tb->wrf[0] : tb->wmain;
    const bool do_se = qpadm_do_se(main_ok, g.nbs, g.jackknife, wmain, nl, g.se_require_p != 0, tb->chisq[0][rank],
                                   dof_of(nl, nr, rank), g.p_se_threshold);
    double se = 0.0;
    if (do_se && lane < nl) se = loo_se(g.wm + static_cast<std::int64_t>(slot) * g.nbs * kQpMaxNl, g.nbs, nl, lane);
    write_qpadm_record(out, md.mi, nl, nr, rank, g.alpha, tb->chisq, tb->ok, tb->wrf, wmain, main_ok, do_se, se,
                       tb->prank);
}

}  // namespace

std::size_t qpadm_large_setup_doubles(int m, int nbs) {
    const std::size_t mn = static_cast<std::size_t>(m) * nbs, mm = static_cast<std::size_t>(m) * m;
    return 2 * mn + 3 * mm + 3 * static_cast<std::size_t>(m) + 2;
}

std::size_t qpadm_large_fit_doubles(int m_max, int nr_max, int t_max) {
    const std::size_t f = 6 * static_cast<std::size_t>(m_max) + kQpMaxNl * kQpMaxNl +
                          2 * static_cast<std::size_t>(kQpMaxNl) * nr_max + static_cast<std::size_t>(nr_max) * nr_max +
                          2 * static_cast<std::size_t>(nr_max) + kQpMaxNl + static_cast<std::size_t>(t_max) * t_max;
    return (f + 31) / 32 * 32;
}

int qpadm_large_fit_waves() { return kFitWaves; }

void launch_qpadm_large_setup(const QpLargeArgs& g, hipStream_t st) {
    if (g.n_models <= 0) return;
    qpl_setup_kernel<<<static_cast<unsigned>(g.n_models), kSetupThreads, 0, st>>>(g);
    STP_LAUNCH_CHECK();
}

void launch_qpadm_large_fits(const QpLargeArgs& g, int n_workgroups, hipStream_t st) {
    if (g.n_tasks <= 0) return;
    qpl_fit_kernel<<<static_cast<unsigned>(n_workgroups), 64 * kFitWaves, 0, st>>>(g);
    STP_LAUNCH_CHECK();
}

void launch_qpadm_large_finish(const QpLargeArgs& g, hipStream_t st) {
    if (g.n_models <= 0) return;
    qpl_finish_kernel<<<static_cast<unsigned>(ceil_div(g.n_models, kFitWaves)), 64 * kFitWaves, 0, st>>>(g);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
