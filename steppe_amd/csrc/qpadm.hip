// steppe_amd/csrc/qpadm.hip — batched qpAdm fits on the device-resident f2 tensor.
//
// One workgroup (4 wavefronts) fits one model at a time and loops over the models of
// its launch (a shared atomic cursor hands them out in decreasing-cost order).  Per
// model it runs the reference's fit chain (src/core/qpadm/qpadm_fit.cpp:49-132):
//
//   f4 basis        X[k][b] = (f2(Li,R0) + f2(L0,Rj) - f2(L0,R0) - f2(Li,Rj)) / 2,
//                   k = j + nr*i, over the surviving blocks
//                   (src/device/cpu/cpu_backend.cpp:125-150, 2065-2113)
//   LOO / total     src/device/cpu/cpu_backend.cpp:2674-2723
//   covariance      Q = sum_b tau tau^T / nb, Qf = Q + fudge tr(Q) I, Qinv = Qf^-1 by LU
//                   with partial pivoting (cpu_backend.cpp:2219-2271, 2758-2764;
//                   src/core/internal/small_linalg.hpp:55-135)
//   ALS fit         one-sided Jacobi SVD seed, fixed ALS iterations of ridge-GLS
//                   A- and B-steps, weights from [A, 1], residual chi^2
//                   (cpu_backend.cpp:2727-2935; small_linalg.hpp:137-218)
//   rank sweep      ranks 0..min(nl, nr)-1 on the full model (cpu_backend.cpp:2533-2610)
//   popdrop         the full model and each source dropped (src/core/qpadm/ranktest.cpp:25-140)
//   LOO SE          nb leave-one-block-out refits, scaled by (nb-1)/sqrt(nb)
//                   (cpu_backend.cpp:98-122, 2625-2668)
//
// Where the reference accumulates in long double, the device accumulates in
// double-double (TwoSum / FMA TwoProd), which is at least as precise.  Plain double
// sums keep the reference's order; this file is built with FP contraction off so each
// product is rounded before it is added, as in the reference.  The GLS right-hand side
// is evaluated as L^T (Qinv^T x) rather than (Qinv L)^T x (the same quantity, another
// rounding order) and back substitution runs column-wise.
//
// Work per wavefront.  A fit's normal matrices (t x t, t = nl*r or r*nr <= T) are held
// one row per lane in registers; the LU runs the reference's pivot rule (first row of
// largest magnitude) with logical row positions instead of row moves, and pivot rows
// are broadcast by lane reads.  Qinv lives in LDS and is shared by the workgroup; the
// per-wave vectors (x, Qinv^T x, A, B, one W column, the SVD work arrays) live in a
// per-wave LDS slice.

#include <climits>
#include <cmath>
#include <type_traits>
#include <utility>

#include "common.hpp"
#include "qpadm.hpp"
#include "stp_qpadm_result.h"
#include "jackknife_dev.hpp"
#include "qpadm_record.hpp"
#include "wave_linalg.hpp"

namespace stp {

namespace {

using namespace wl;

constexpr int kQWaves = 4;
constexpr int kQThreads = 64 * kQWaves;
constexpr int kMaxNl = kQpMaxNl;
constexpr int kMaxRows = kQpMaxRows;
constexpr int kMaxRanks = kQpMaxRanks;

// ---------------------------------------------------------------------------------
// per-wave LDS slice and one ALS fit
// ---------------------------------------------------------------------------------
struct WaveSlice {
    double* x;    // xvec, k = i*nr + j of the fit's (reduced) model
    double* y;    // Qsub^T x
    double* a;    // A[i][p] at i*r + p
    double* b;    // B[p][j] at p*nr + j
    double* w;    // one W column / the residual
    double* ws;   // Jacobi work matrix, (i, j) at i + nl*j
    double* sig;  // singular values (nr)
    double* v;    // Jacobi right vectors, (i, j) at i + nr*j
    int* ord;     // singular value order (nr)
};

struct FitShape {
    int nl, nr, r, drop;  // drop: full-model source removed (-1: none)
};

// Row of the full Qinv for reduced source index i.
__device__ inline int src_of(int i, int drop) { return (drop >= 0 && i >= drop) ? i + 1 : i; }

// Residual chi^2 = e^T Qsub e, e = x - (A B) (cpu_backend.cpp:2840-2860 ordering:
// row sums over b, then the sum over a, both in extended precision).
__device__ double resid_chisq(const double* Q, int ldq, FitShape s, const WaveSlice& W) {
    const int lane = lane_id();
    const int m = s.nl * s.nr;
    if (lane < m) {
        const int i = lane / s.nr, j = lane - (lane / s.nr) * s.nr;
        double ab = 0.0;
        for (int p = 0; p < s.r; ++p) ab += W.a[i * s.r + p] * W.b[p * s.nr + j];
        W.w[lane] = W.x[lane] - ab;
    }
    __builtin_amdgcn_wave_barrier();
    dd row{0.0, 0.0};
    if (lane < m) {
        const int ia = lane / s.nr, ja = lane - ia * s.nr;
        const double* qrow = Q + (src_of(ia, s.drop) * s.nr + ja) * ldq;
        for (int i = 0; i < s.nl; ++i) {
            const double* qq = qrow + src_of(i, s.drop) * s.nr;
            for (int j = 0; j < s.nr; ++j) row = dd_fma(row, qq[j], W.w[i * s.nr + j]);
        }
    }
    const double e_own = lane < m ? W.w[lane] : 0.0;
    dd acc{0.0, 0.0};
    for (int a = 0; a < m; ++a) {
        const double ea = rl(e_own, a);
        const dd ra{rl(row.hi, a), rl(row.lo, a)};
        acc = dd_fma_dd(acc, ea, ra);
    }
    __builtin_amdgcn_wave_barrier();
    return uni(acc.hi + acc.lo);
}

// The ALS fit of one (reduced) model at rank r on W.x (cpu_backend.cpp:2727-2935).
// Returns ok; writes the weights (nl) to w_out (LDS) and, when want_chisq is set, the
// chi^2 (the leave-one-block-out refits use only the weights).
template <int T>
__device__ __noinline__ bool als_fit(const double* Q, int ldq_arg, FitShape s_arg, int iters_arg, double fudge_arg,
                                     const WaveSlice& W, double* w_out, double& chisq, bool want_chisq_arg) {
    const int lane = lane_id();
    // A called function receives its arguments in vector registers, so the compiler
    // cannot see that they are wavefront-uniform: moved to scalar registers, every
    // loop bound and size test below is a scalar branch rather than a vector compare
    // with exec-mask bookkeeping.
    const int ldq = __builtin_amdgcn_readfirstlane(ldq_arg);
    const FitShape s{__builtin_amdgcn_readfirstlane(s_arg.nl), __builtin_amdgcn_readfirstlane(s_arg.nr),
                     __builtin_amdgcn_readfirstlane(s_arg.r), __builtin_amdgcn_readfirstlane(s_arg.drop)};
    const int iters = __builtin_amdgcn_readfirstlane(iters_arg);
    const double fudge = uni(fudge_arg);
    const bool want_chisq = __builtin_amdgcn_readfirstlane(want_chisq_arg ? 1 : 0) != 0;
    const int nl = s.nl, nr = s.nr, r = s.r, m = nl * nr;
    chisq = 0.0;
    if (r == 0) {
        if (lane < nl) w_out[lane] = 1.0;
        if (want_chisq) chisq = resid_chisq(Q, ldq, s, W);
        return true;
    }
    // x, A and B are also held one entry per lane in registers (xr, ar, br): reads at a
    // wavefront-uniform index are lane reads, in uniform control flow, instead of LDS
    // loads.  Every sum keeps its order, so the values are unchanged.
    const double xr = lane < m ? W.x[lane] : 0.0;
    // y = Qsub^T x (column kc of Qsub against x, k ascending)
    {
        const int ic = lane < m ? lane / nr : 0, jc = lane < m ? lane - ic * nr : 0;
        const int col = src_of(ic, s.drop) * nr + jc;
        double acc = 0.0;
        for (int i = 0; i < nl; ++i) {
            const int rb = src_of(i, s.drop) * nr;
            for (int j = 0; j < nr; ++j) acc += Q[(rb + j) * ldq + col] * rl(xr, i * nr + j);
        }
        if (lane < m) W.y[lane] = acc;
    }
    // ---- seed: one-sided Jacobi SVD of xm (nl x nr) on its columns ----
    if (lane < m) {
        const int i = lane / nr, j = lane - (lane / nr) * nr;
        W.ws[i + nl * j] = xr;
    }
    for (int e = lane; e < nr * nr; e += 64) W.v[e] = (e % nr == e / nr) ? 1.0 : 0.0;
    __builtin_amdgcn_wave_barrier();
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
                if (lane < nr) {
                    const double x0 = W.v[lane + nr * p], y0 = W.v[lane + nr * q];
                    W.v[lane + nr * p] = c * x0 - sn * y0;
                    W.v[lane + nr * q] = sn * x0 + c * y0;
                }
                __builtin_amdgcn_wave_barrier();
            }
        if (uni(off) < 1e-30) break;
    }
    if (lane < nr) {
        double ss = 0.0;
        for (int i = 0; i < nl; ++i) ss += W.ws[i + nl * lane] * W.ws[i + nl * lane];
        W.sig[lane] = sqrt(ss);
        W.ord[lane] = lane;
    }
    __builtin_amdgcn_wave_barrier();
    // The reference's exchange sort on descending sigma; positions 0..r-1 are final
    // after r outer passes.  Every lane runs it on the same LDS values.
    for (int i = 0; i < r; ++i)
        for (int j = i + 1; j < nr; ++j) {
            const int oi = W.ord[i], oj = W.ord[j];
            if (W.sig[oj] > W.sig[oi]) {
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) {
                    W.ord[i] = oj;
                    W.ord[j] = oi;
                }
                __builtin_amdgcn_wave_barrier();
            }
        }
    __builtin_amdgcn_wave_barrier();
    // r * nr < m <= 64: one entry of B per lane
    double br = 0.0, ar = 0.0;
    if (lane < r * nr) {
        const int p = lane / nr, j = lane - (lane / nr) * nr;
        br = W.v[j + nr * W.ord[p]];
        W.b[lane] = br;
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < nl * r) {
        const int i = lane / r, p = lane - (lane / r) * r;
        double acc = 0.0;
        for (int j = 0; j < nr; ++j) acc += W.x[i * nr + j] * W.b[p * nr + j];
        ar = acc;
        W.a[lane] = acc;
    }
    __builtin_amdgcn_wave_barrier();

    // lane -> its row of Qsub (full-matrix row index) for the W columns
    int qrow = 0;
    if (lane < m) {
        const int i = lane / nr, j = lane - (lane / nr) * nr;
        qrow = src_of(i, s.drop) * nr + j;
    }
    const double* Qr = Q + qrow * ldq;

    for (int it = 0; it < iters; ++it) {
        // step 0, the A-step: unknowns A[i][p] at a = i*r + p (L(k=(i',j), a) = [i'==i] B[p][j]);
        // step 1, the B-step: unknowns B[p][j] at a = p*nr + j (L(k=(i,j'), a) = [j'==j] A[i][p]).
        for (int step = 0; step < 2; ++step) {
            const bool bstep = step == 1;
            const int t = bstep ? r * nr : nl * r;
            // lane a's (outer, inner) unknown indices: A-step (i, p), B-step (p, j)
            const int div = bstep ? nr : r;
            const int ua = lane / div, va = lane - (lane / div) * div;
            double row[T];
            double diag = 0.0;
            int uc = 0, vc = 0;
            // Columns in pairs (c0, c0 + 1): both W columns in one pass (the second in
            // the Jacobi work buffer, free after the seed), then both entries of the
            // lane's row of C; each sum keeps its order, so the values are those of a
            // column-by-column build.
            static_for<T / 2>([&](auto cp) {
                constexpr int c0 = 2 * decltype(cp)::value, c1 = c0 + 1;
                row[c0] = 0.0;
                row[c1] = 0.0;
                if (c0 < t) {
                    const bool two = c1 < t;
                    int uc1 = uc, vc1 = vc + 1;
                    if (vc1 == div) {
                        vc1 = 0;
                        ++uc1;
                    }
                    // W[:, c] = Qsub L[:, c] (lanes >= m run on row 0 and keep nothing)
                    {
                        double acc0 = 0.0, acc1 = 0.0;
                        if (!bstep) {
                            const double* qq0 = Qr + src_of(uc, s.drop) * nr;
                            const double* qq1 = Qr + src_of(uc1, s.drop) * nr;
                            for (int j = 0; j < nr; ++j) {
                                acc0 += qq0[j] * rl(br, vc * nr + j);
                                if (two) acc1 += qq1[j] * rl(br, vc1 * nr + j);
                            }
                        } else {
                            for (int i = 0; i < nl; ++i) {
                                const double* qi = Qr + src_of(i, s.drop) * nr;
                                acc0 += qi[vc] * rl(ar, i * r + uc);
                                if (two) acc1 += qi[vc1] * rl(ar, i * r + uc1);
                            }
                        }
                        if (lane < m) {
                            W.w[lane] = acc0;
                            if (two) W.ws[lane] = acc1;
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    // C[a][c] = L[:, a]^T W[:, c]
                    double cv0 = 0.0, cv1 = 0.0;
                    if (lane < t) {
                        if (!bstep) {
                            const double* bb = W.b + va * nr;
                            const double* ww0 = W.w + ua * nr;
                            const double* ww1 = W.ws + ua * nr;
                            for (int j = 0; j < nr; ++j) {
                                const double bj = bb[j];
                                cv0 += bj * ww0[j];
                                if (two) cv1 += bj * ww1[j];
                            }
                        } else {
                            for (int i = 0; i < nl; ++i) {
                                const double ai = W.a[i * r + ua];
                                cv0 += ai * W.w[i * nr + va];
                                if (two) cv1 += ai * W.ws[i * nr + va];
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    row[c0] = cv0;
                    if (lane == c0) diag = cv0;
                    if (two) {
                        row[c1] = cv1;
                        if (lane == c1) diag = cv1;
                    }
                    uc = uc1;
                    vc = vc1;
                    if (++vc == div) {
                        vc = 0;
                        ++uc;
                    }
                }
            });
            double tr = 0.0;
            for (int a = 0; a < t; ++a) tr += rl(diag, a);
            const double ft = fudge * tr;
            static_for<T>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                if (c < t && lane == c) row[c] += ft;
            });
            double rhs = 0.0;
            if (lane < t) {
                if (!bstep) {
                    const double* bb = W.b + va * nr;
                    const double* yy = W.y + ua * nr;
                    for (int j = 0; j < nr; ++j) rhs += bb[j] * yy[j];
                } else {
                    for (int i = 0; i < nl; ++i) rhs += W.a[i * r + ua] * W.y[i * nr + va];
                }
            }
            double sol = 0.0;
            const bool ok = lu_solve_rows<T>(row, rhs, t, sol);
            double* dst = bstep ? W.b : W.a;
            if (lane < t) {
                const double nv = ok ? sol : 0.0;
                dst[lane] = nv;
                if (bstep) br = nv;
                else ar = nv;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    // ---- weights: (Xa Xa^T) w = Xa[:, r], Xa = [A, 1], normalised to sum 1 ----
    {
        double row[kMaxNl];
        const int la = lane < nl ? lane : 0;
#pragma unroll
        for (int i2 = 0; i2 < kMaxNl; ++i2) {
            double acc = 0.0;
            if (i2 < nl) {
                for (int p = 0; p < r; ++p) acc += W.a[la * r + p] * rl(ar, i2 * r + p);
                acc += 1.0 * 1.0;
                if (lane >= nl) acc = 0.0;
            }
            row[i2] = acc;
        }
        double wv = 0.0;
        if (!lu_solve_rows<kMaxNl>(row, 1.0, nl, wv)) return false;
        double sum = 0.0;
        for (int i = 0; i < nl; ++i) sum += rl(wv, i);
        if (lane < nl) w_out[lane] = wv / sum;
    }
    __builtin_amdgcn_wave_barrier();
    if (want_chisq) chisq = resid_chisq(Q, ldq, s, W);
    return true;
}

// ---------------------------------------------------------------------------------
// the batched kernel
// ---------------------------------------------------------------------------------
// Wavefronts per SIMD each system-size bucket is compiled for.  The fits are bound by
// the latency of their lane reads, LDS accesses and short dependent FP64 chains, not by
// register traffic: capping the registers (what does not fit spills to scratch) for
// more resident wavefronts is faster.  On the 10,000-model C4 fleet: no cap 1,923
// fits/s; 4 for T = 16, 3 for T = 32: 2,244; 4 for T <= 32 and 3 for T = 48: 2,654;
// 4 for T = 48: 2,604; 5 for T = 32: 2,494; with the fit's sizes in scalar registers,
// 4 / 4 / 3: 2,885 and 5 / 4 / 3: 2,925 (two runs each; eight wavefronts per workgroup
// instead of four: 2,690).
template <int T>
constexpr int qpadm_waves_per_simd() {
    return T == 16 ? 5 : (T == 32 ? 4 : (T == 48 ? 3 : 1));
}

template <int T>
__global__ __launch_bounds__(kQThreads) __attribute__((amdgpu_waves_per_eu(qpadm_waves_per_simd<T>())))
void qpadm_kernel(QpadmArgs g) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    __shared__ int s_model, s_status, s_next, s_nfit;
    __shared__ int s_L[kMaxNl + 1];
    __shared__ int s_R[kMaxQpadmNr + 1];
    __shared__ int s_pos[kMaxQpadmM], s_lane_of[2 * kQWaves + 2];
    __shared__ double s_val[kQWaves + 1];
    __shared__ signed char s_fit_row[kMaxRows * kMaxRanks + 2], s_fit_rank[kMaxRows * kMaxRanks + 2];
    __shared__ double s_chisq[kMaxRows][kMaxRanks];
    __shared__ double s_wrf[kMaxRows][kMaxNl];  // weights of each popdrop row's fit at rank nlr-1
    __shared__ double s_wmain[kMaxNl];          // weights of the main fit when its rank differs
    __shared__ int s_ok[kMaxRows][kMaxRanks];
    __shared__ double s_xtot[kMaxQpadmM], s_totline[kMaxQpadmM];
    __shared__ double s_prank[kMaxRanks];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int ldq = g.ldq;
    double* const sQi = lds;                                   // Qinv, m x ldq
    double* const sLU = lds + g.m_max * ldq;                   // Qf / its LU (setup only)
    double* const wbase = sLU;                                 // per-wave slices (fit phase)
    const int wstride = g.wave_slice;
    WaveSlice W;
    {
        double* p = wbase + wave * wstride;
        W.x = p;
        W.y = p + 64;
        W.a = p + 128;
        W.b = p + 192;
        W.w = p + 256;
        W.ws = p + 320;
        W.sig = p + 384;
        W.v = p + 416;
        W.ord = reinterpret_cast<int*>(p + 416 + g.nr_max * g.nr_max);
    }
    double* const scr = g.scratch + static_cast<std::int64_t>(blockIdx.x) * g.scratch_stride;
    const int nbs = g.nbs;
    double* const sX = scr;                                    // X, then tau: [m][nbs]
    double* const sLOO = scr + static_cast<std::int64_t>(g.m_max) * nbs;      // [m][nbs]
    double* const sWM = sLOO + static_cast<std::int64_t>(g.m_max) * nbs;      // [nbs][nl]
    const double n_tot = g.n_total;

    for (;;) {
        if (tid == 0) s_model = static_cast<int>(atomicAdd(g.counter, 1u));
        __syncthreads();
        const int slot = __builtin_amdgcn_readfirstlane(s_model);
        __syncthreads();
        if (slot >= g.n_order) break;
        const int mi = g.order[slot];
        const int l0 = g.left_off[mi], l1 = g.left_off[mi + 1];
        const int r0 = g.right_off[mi], r1 = g.right_off[mi + 1];
        const int nl = l1 - l0, nr = r1 - r0 - 1, m = nl * nr;
        const int rank = g.rank < 0 ? nl - 1 : g.rank;
        if (tid == 0) s_L[0] = g.target[mi];
        if (tid < nl) s_L[tid + 1] = g.left[l0 + tid];
        if (tid <= nr) s_R[tid] = g.right[r0 + tid];
        __syncthreads();
        stp_qpadm_result* out = g.out + mi;

        // ---- f4 basis over the surviving blocks ----
        {
            const int P = g.P;
            const std::int64_t ld2 = nbs;
            const int L0 = s_L[0], R0 = s_R[0];
            const double* f00 = g.f2T + (L0 + static_cast<std::int64_t>(P) * R0) * ld2;
            for (int k = wave; k < m; k += kQWaves) {
                const int i = k / nr, j = k - (k / nr) * nr;
                const int Li = s_L[i + 1], Rj = s_R[j + 1];
                const double* fi0 = g.f2T + (Li + static_cast<std::int64_t>(P) * R0) * ld2;
                const double* f0j = g.f2T + (L0 + static_cast<std::int64_t>(P) * Rj) * ld2;
                const double* fij = g.f2T + (Li + static_cast<std::int64_t>(P) * Rj) * ld2;
                for (int b = lane; b < nbs; b += 64)
                    sX[static_cast<std::int64_t>(k) * nbs + b] = 0.5 * (fi0[b] + f0j[b] - f00[b] - fij[b]);
            }
        }
        __syncthreads();
        // ---- LOO / total, jackknife covariance, ridge and its inverse ----
        const bool cov_ok = jk_covariance<kQThreads>(sX, sLOO, g.bsz, n_tot, nbs, m, g.fudge, s_xtot, s_totline,
                                                     nullptr, sLU, ldq, sQi, ldq, s_pos, s_lane_of, s_val,
                                                     &s_status);
        if (!cov_ok) {
            if (tid == 0) {
                out->status = 3;  // NonSpdCovariance
                out->model_index = mi;
                out->nl = 0;
                out->est_rank = rank;
                out->dof = dof_of(nl, nr, rank);
                out->n_rank = 0;
                out->n_popdrop = 0;
            }
            continue;
        }

        // ---- fit table: rank sweeps of the full model and of each popdrop row ----
        const int nrows = nl >= 2 ? nl + 1 : 1;
        if (tid == 0) {
            int n = 0;
            for (int k = 0; k < nrows; ++k) {
                const int nlr = k == 0 ? nl : nl - 1;
                const int rmax = min(nlr, nr) - 1, rf = nlr - 1;
                for (int rr = 0; rr <= rmax; ++rr) {
                    s_fit_row[n] = static_cast<signed char>(k);
                    s_fit_rank[n++] = static_cast<signed char>(rr);
                }
                if (rf > rmax) {
                    s_fit_row[n] = static_cast<signed char>(k);
                    s_fit_rank[n++] = static_cast<signed char>(rf);
                }
                if (k == 0 && rank > rmax && rank != rf) {
                    s_fit_row[n] = 0;
                    s_fit_rank[n++] = static_cast<signed char>(rank);
                }
            }
            s_nfit = n;
            s_next = 0;
        }
        __syncthreads();
        for (;;) {
            int f = 0;
            if (lane == 0) f = atomicAdd(&s_next, 1);
            f = __builtin_amdgcn_readfirstlane(__shfl(f, 0));
            if (f >= s_nfit) break;
            const int k = s_fit_row[f], rr = s_fit_rank[f];
            FitShape sh{k == 0 ? nl : nl - 1, nr, rr, k == 0 ? -1 : nl - k};
            const int mr = sh.nl * nr;
            if (lane < mr) {
                const int i = lane / nr, j = lane - (lane / nr) * nr;
                W.x[lane] = s_xtot[src_of(i, sh.drop) * nr + j];
            }
            __builtin_amdgcn_wave_barrier();
            double chisq = 0.0;
            const int rf = sh.nl - 1;
            double* wout = rr == rf ? s_wrf[k] : (k == 0 && rr == rank ? s_wmain : W.sig);
            const bool ok = als_fit<T>(sQi, ldq, sh, g.iters, g.fudge, W, wout, chisq, true);
            if (lane == 0) {
                s_chisq[k][rr] = ok ? chisq : 0.0;
                s_ok[k][rr] = ok ? 1 : 0;
            }
        }
        __syncthreads();
        const bool main_ok = s_ok[0][rank] != 0;
        const double* wmain = rank == nl - 1 ? s_wrf[0] : s_wmain;
        const int dof = dof_of(nl, nr, rank);
        bool do_se = false;
        if (main_ok && nbs >= 2) {
            if (g.jackknife == 2) do_se = true;
            else if (g.jackknife == 1)
                do_se = feasible_w(wmain, nl) &&
                        (!g.se_require_p || chisq_upper(s_chisq[0][rank], dof) >= g.p_se_threshold);
        }
        if (do_se) {
            if (tid == 0) s_next = 0;
            __syncthreads();
            for (;;) {
                int b = 0;
                if (lane == 0) b = atomicAdd(&s_next, 1);
                b = __builtin_amdgcn_readfirstlane(__shfl(b, 0));
                if (b >= nbs) break;
                if (lane < m) W.x[lane] = sLOO[static_cast<std::int64_t>(lane) * nbs + b];
                __builtin_amdgcn_wave_barrier();
                FitShape sh{nl, nr, rank, -1};
                double chisq = 0.0;
                double* wrow = W.sig;  // nl weights; sig is free after the seed
                const bool ok = als_fit<T>(sQi, ldq, sh, g.iters, g.fudge, W, wrow, chisq, false);
                __builtin_amdgcn_wave_barrier();
                if (lane < nl) sWM[static_cast<std::int64_t>(b) * nl + lane] = ok ? wrow[lane] : 0.0;
            }
            __syncthreads();
        }

        // ---- results (wave 0; lane 0 writes the record fields) ----
        if (tid < 64) {
            double se = 0.0;
            if (do_se && lane < nl) se = loo_se(sWM, nbs, nl, lane);
            write_qpadm_record(out, mi, nl, nr, rank, g.alpha, s_chisq, s_ok, s_wrf, wmain, main_ok, do_se, se,
                               s_prank);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// survivors and the survivor-transposed tensor
// ---------------------------------------------------------------------------------
// keep[b] = !(some pair has vpair <= 0 and another has vpair > 0)
// (the reference's survivor_blocks, cpu_backend.cpp:125-150).
__global__ __launch_bounds__(256) void survivor_kernel(const double* __restrict__ vpair, std::int64_t slab,
                                                       std::int32_t* __restrict__ keep) {
    const std::int64_t b = blockIdx.x;
    const double* v = vpair + b * slab;
    int pres = 0, miss = 0;
    for (std::int64_t e = threadIdx.x; e < slab; e += blockDim.x) {
        if (v[e] > 0.0) pres = 1;
        else miss = 1;
    }
    pres = __syncthreads_or(pres);
    miss = __syncthreads_or(miss);
    if (threadIdx.x == 0) keep[b] = (pres && miss) ? 0 : 1;
}

// f2T[pair * nbs + s] = f2[pair + slab * surv[s]] by 64 x 64 LDS tiles.
__global__ __launch_bounds__(256) void transpose_survivors_kernel(const double* __restrict__ f2, std::int64_t slab,
                                                                  const std::int32_t* __restrict__ surv, int nbs,
                                                                  double* __restrict__ f2T) {
    __shared__ double tile[64][65];
    const std::int64_t p0 = static_cast<std::int64_t>(blockIdx.x) * 64;
    const int s0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int sy = ty; sy < 64; sy += 4) {
        const int s = s0 + sy;
        const std::int64_t p = p0 + tx;
        tile[sy][tx] = (s < nbs && p < slab) ? f2[p + slab * surv[s]] : 0.0;
    }
    __syncthreads();
    for (int py = ty; py < 64; py += 4) {
        const std::int64_t p = p0 + py;
        const int s = s0 + tx;
        if (p < slab && s < nbs) f2T[p * nbs + s] = tile[tx][py];
    }
}

}  // namespace

void launch_survivors(const double* vpair, std::int64_t slab, int n_block, std::int32_t* keep, hipStream_t st) {
    if (n_block <= 0) return;
    survivor_kernel<<<static_cast<unsigned>(n_block), 256, 0, st>>>(vpair, slab, keep);
    STP_LAUNCH_CHECK();
}

void launch_transpose_survivors(const double* f2, std::int64_t slab, const std::int32_t* surv, int nbs, double* f2T,
                                hipStream_t st) {
    if (nbs <= 0 || slab <= 0) return;
    const dim3 grid(static_cast<unsigned>(ceil_div(slab, 64)), static_cast<unsigned>(ceil_div(nbs, 64)));
    transpose_survivors_kernel<<<grid, 256, 0, st>>>(f2, slab, surv, nbs, f2T);
    STP_LAUNCH_CHECK();
}

std::size_t qpadm_lds_bytes(int m_max, int ldq, int wave_slice) {
    const std::size_t a = static_cast<std::size_t>(m_max) * ldq;
    const std::size_t b = std::max<std::size_t>(a, static_cast<std::size_t>(kQWaves) * wave_slice);
    return (a + b) * sizeof(double);
}

int qpadm_wave_slice(int nr_max) { return 416 + nr_max * nr_max + (kMaxQpadmNr + 1) / 2 + 1; }

void launch_qpadm(int T, const QpadmArgs& args, int n_workgroups, hipStream_t st) {
    if (args.n_order <= 0) return;
    const std::size_t lds = qpadm_lds_bytes(args.m_max, args.ldq, args.wave_slice);
    const dim3 grid(static_cast<unsigned>(n_workgroups));
    auto go = [&](auto kernel) {
        if (lds > 64 * 1024)
            STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        kernel<<<grid, kQThreads, lds, st>>>(args);
    };
    switch (T) {
        case 16: go(qpadm_kernel<16>); break;
        case 32: go(qpadm_kernel<32>); break;
        case 48: go(qpadm_kernel<48>); break;
        case 64: go(qpadm_kernel<64>); break;
        default: throw ArgError("qpadm: unsupported system size bucket " + std::to_string(T));
    }
    STP_LAUNCH_CHECK();
}

int qpadm_threads() { return kQThreads; }

int qpadm_blocks_per_cu(int T, std::size_t lds) {
    int n = 0;
    auto q = [&](auto kernel) {
        if (lds > 64 * 1024)
            STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        STP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kQThreads, lds));
    };
    switch (T) {
        case 16: q(qpadm_kernel<16>); break;
        case 32: q(qpadm_kernel<32>); break;
        case 48: q(qpadm_kernel<48>); break;
        case 64: q(qpadm_kernel<64>); break;
        default: throw ArgError("qpadm: unsupported system size bucket " + std::to_string(T));
    }
    return n > 0 ? n : 1;
}

}  // namespace stp
