// steppe_amd/csrc/qpgraph.hip — qpGraph on the device: the f3 basis with its jackknife
// covariance, and a batched score + adjoint kernel.
//
// Score (the reference's qpgraph_score, src/core/qpadm/qpgraph_objective.hpp:24-190):
// for mixture weights theta the leaf weights pwts follow from the path tables, the
// centred weights pc = pwts[:, leaf] - pwts[:, base], the pair products
// P[k][e] = pc[e][a_k] pc[e][b_k]; with G = Qinv, cc = P^T G P, q = P^T G f, the ridge
// C = cc + (fudge tr(cc) / ne) I is scaled by s_e = sqrt(C_ee) and the edge lengths
// solve the scaled system by Lawson-Hanson NNLS (constrained) or LU (free);
// score = res^T G res with res = f - P bl.
//
// Adjoint (hand-derived, reverse mode).  On the NNLS active set A, bl_A = C_AA^-1 q_A
// (the scaling cancels), so with rbar = (G + G^T) res, blbar = -P^T rbar,
// lambda = C_AA^-T blbar_A (zero off A) and kappa = -(fudge / ne) sum_A lambda_e bl_e:
//   Pbar[k][e] = -bl_e (rbar + G^T P lambda)_k + lambda_e (G res)_k
//                + kappa ((G + G^T) P)[k][e],
// then pcbar from the pair products, pwbar = pcbar (leaf columns) and minus their sum
// (base column), pathbar by the path sums, and thetabar from the path products.  In
// the joint mode (bl given) Pbar = -rbar bl^T and the gradient in bl is blbar.
// The derivative is that of the active-set branch; at a change of active set the
// score has a kink and this is one of its one-sided derivatives.
//
// One workgroup of 256 threads evaluates one instance (a topology, its data and a
// theta), persistent over the instances of the launch.  Everything per instance sits in
// LDS; Qinv is read from global memory (row-major).  The NNLS and the small LU solves
// run on wavefront 0 with one matrix row per lane.

#include <climits>
#include <cmath>

#include "common.hpp"
#include "jackknife_dev.hpp"
#include "qpgraph.hpp"
#include "wave_linalg.hpp"

namespace stp {

namespace {

using namespace wl;

constexpr int kThreads = 256;
constexpr double kNnlsEps = 1e-12;
constexpr int kMaxD = 16;

struct EvalLds {
    double *path, *pathbar, *pw, *pc, *pcbar, *P, *W, *C;
    double *q, *q1, *sc, *bl, *blbar, *lam;
    double *qf, *res, *Gr, *GTr, *v, *u;
};

__device__ inline EvalLds eval_layout(double* base, const QgEvalArgs& g) {
    const int NP = g.np_max, NE = g.ne_max, NPOP = g.npop_max, NPATH = g.npath_max;
    EvalLds L;
    double* p = base;
    auto take = [&](int n) { double* r = p; p += n; return r; };
    L.path = take(NPATH);
    L.pathbar = take(NPATH);
    L.pw = take(NE * NPOP);
    L.pc = take(NE * (NPOP - 1));
    L.pcbar = take(NE * (NPOP - 1));
    L.P = take(NP * NE);
    L.W = take(NP * NE);
    L.C = take(NE * NE);
    L.q = take(NE);
    L.q1 = take(NE);
    L.sc = take(NE);
    L.bl = take(NE);
    L.blbar = take(NE);
    L.lam = take(NE);
    L.qf = take(NP);
    L.res = take(NP);
    L.Gr = take(NP);
    L.GTr = take(NP);
    L.v = take(NP);
    L.u = take(NP);
    return L;
}

// Solves the n x n sub-system of the LDS matrix C (leading dimension ld) on the rows /
// columns idx[0..n) (identity when idx is null), transposed when requested, scaled to
// C[i][j] / (sc_i sc_j) when sc is given; lane a holds right-hand side b and receives
// x[a].  One out-of-line instance per T keeps a single copy of the register LU.
template <int T>
__device__ __noinline__ bool solve_sub(const double* C, int ld_arg, const int* idx, int n_arg, bool transpose_arg,
                                       const double* sc, double b, double& x) {
    const int lane = lane_id();
    // Arguments of a called function arrive in vector registers; the sizes are
    // wavefront-uniform, and in scalar registers they give scalar branches in the LU.
    const int ld = __builtin_amdgcn_readfirstlane(ld_arg);
    const int n = __builtin_amdgcn_readfirstlane(n_arg);
    const bool transpose = __builtin_amdgcn_readfirstlane(transpose_arg ? 1 : 0) != 0;
    const int ia = lane < n ? (idx ? idx[lane] : lane) : 0;
    double row[T];
    static_for<T>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        double v = 0.0;
        if (lane < n && c < n) {
            const int jc = idx ? idx[c] : c;
            v = transpose ? C[jc * ld + ia] : C[ia * ld + jc];
            if (sc) v = v / (sc[ia] * sc[jc]);
        }
        row[c] = v;
    });
    return lu_solve_rows<T>(row, b, n, x);
}

// Lawson-Hanson NNLS on the scaled system A(i,j) = C[i][j] / (s_i s_j), q1, on one
// wavefront (lane i <-> unknown i, n <= T <= 64), as the reference's nnls_active_set
// (qpgraph_objective.hpp:41-100): tolerance 1e-12, at most 3n+30 outer and inner steps.
// The solution goes to bls (LDS), the final passive set to passive (LDS).
// With warm set, the passive set held in passive_out (from an earlier solve of a nearby
// problem) is tried first: when the system restricted to it has an all-positive
// solution, that solution is the starting point of the outer loop (a feasible point
// that is optimal on its passive set, as the algorithm requires); otherwise the solve
// starts from zero.  The minimiser is unique, so the result is the same.
template <int T>
__device__ bool nnls_wave(int n, const double* C, const double* sc, const double* q1, double* bls, int* passive_out,
                          int* slot, bool warm) {
    const int lane = lane_id();
    const bool act = lane < n;
    auto A = [&](int i, int j) { return C[i * n + j] / (sc[i] * sc[j]); };
    if (act) bls[lane] = 0.0;
    bool passive = false;
    __builtin_amdgcn_wave_barrier();
    if (warm) {
        passive = act && passive_out[lane] != 0;
        const unsigned long long mask = __ballot(passive);
        const int p = __builtin_popcountll(mask);
        if (p > 0) {
            if (passive) slot[__builtin_popcountll(mask & ((1ull << lane) - 1ull))] = lane;
            __builtin_amdgcn_wave_barrier();
            const int pa = lane < p ? slot[lane] : 0;
            double z = 0.0;
            const bool ok = solve_sub<T>(C, n, slot, p, false, sc, lane < p ? q1[pa] : 0.0, z);
            const bool all_pos = ok && popc(lane < p && z <= kNnlsEps) == 0;
            if (all_pos) {
                if (lane < p) bls[pa] = z;
            } else {
                passive = false;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    const int max_iter = 3 * n + 30;
    for (int outer = 0; outer < max_iter; ++outer) {
        double w = 0.0;
        if (act) {
            double s = q1[lane];
            for (int j = 0; j < n; ++j) s -= A(lane, j) * bls[j];
            w = s;
        }
        const bool cand = act && !passive && w > kNnlsEps;
        if (popc(cand) == 0) break;
        const int t = argmax_pivot(cand ? w : -1.0, lane);
        if (lane == t) passive = true;
        for (int inner = 0; inner < max_iter; ++inner) {
            const unsigned long long mask = __ballot(passive);
            const int p = __builtin_popcountll(mask);
            if (passive) slot[__builtin_popcountll(mask & ((1ull << lane) - 1ull))] = lane;
            __builtin_amdgcn_wave_barrier();
            const int pa = lane < p ? slot[lane] : 0;
            double z = 0.0;
            if (!solve_sub<T>(C, n, slot, p, false, sc, lane < p ? q1[pa] : 0.0, z)) return false;
            const bool all_pos = popc(lane < p && z <= kNnlsEps) == 0;
            if (all_pos) {
                if (act) bls[lane] = 0.0;
                __builtin_amdgcn_wave_barrier();
                if (lane < p) bls[pa] = z;
                __builtin_amdgcn_wave_barrier();
                break;
            }
            double alpha = 1.0;
            for (int a = 0; a < p; ++a) {
                const double za = rl(z, a);
                if (za <= kNnlsEps) {
                    const double blp = bls[slot[a]];
                    const double denom = blp - za;
                    if (denom > kNnlsEps) {
                        const double r = blp / denom;
                        if (r < alpha) alpha = r;
                    }
                }
            }
            alpha = uni(alpha);
            if (lane < p) bls[pa] += alpha * (z - bls[pa]);
            __builtin_amdgcn_wave_barrier();
            if (passive && bls[lane] <= kNnlsEps) {
                passive = false;
                bls[lane] = 0.0;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (act) passive_out[lane] = passive ? 1 : 0;
    __builtin_amdgcn_wave_barrier();
    return true;
}

// Per-workgroup state of one instance evaluation (LDS).
struct EvalShared {
    int* passive;    // [64] NNLS passive set
    int* slot;       // [64] compaction scratch
    double* theta;   // [kMaxD] the instance's mixture weights (input)
    double* thbar;   // [kMaxD] d score / d theta (mode >= 1)
    double* scal;    // [4]: 0 kappa, 1 score
    int* ok;
};

// Evaluates one instance; every thread of the workgroup calls it.  The score of
// sh.theta goes to sh.scal[1], the edge lengths to L.bl (in the joint mode, mode 2,
// L.bl holds the given edge lengths on entry), for mode >= 1 the adjoint to sh.thbar,
// and in the joint mode d score / d bl to L.blbar.  fit (global, may be null) receives
// P bl.  Returns false on every thread when an edge-length system is singular; the
// score is then +inf.
template <int T>
__device__ bool qg_eval_one(const QgEvalArgs& g, const QgTopoDev& tp, const double* __restrict__ G,
                            const double* __restrict__ f, int mode, const EvalLds& L, const EvalShared& sh,
                            double* fit, bool warm = false) {
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const std::int32_t* ia = g.iarena;
    const int np = tp.npair, ne = tp.ne, npop = tp.npop, ncc = npop - 1, D = tp.nadmix, npath = tp.npath;
    // ---- leaf weights from the path tables ----
    for (int p = tid; p < npath; p += kThreads) {
        double w = 1.0;
        for (int t = ia[tp.o_pae_off + p]; t < ia[tp.o_pae_off + p + 1]; ++t) {
            const int id = ia[tp.o_pae_id + t];
            const double th = sh.theta[(id - 1) / 2];
            w *= (id % 2 == 1) ? th : (1.0 - th);
        }
        L.path[p] = w;
    }
    for (int x = tid; x < ne * npop; x += kThreads) L.pw[x] = g.darena[tp.o_pwts0 + x];
    __syncthreads();
    for (int gr = tid; gr < tp.ngrp; gr += kThreads) {
        double s = 0.0;
        for (int t = ia[tp.o_grp_off + gr]; t < ia[tp.o_grp_off + gr + 1]; ++t) s += L.path[ia[tp.o_grp_paths + t]];
        L.pw[ia[tp.o_grp_edge + gr] + ne * ia[tp.o_grp_leaf + gr]] = s;
    }
    __syncthreads();
    for (int x = tid; x < ne * ncc; x += kThreads) {
        const int col = x / ne, e = x - (x / ne) * ne;
        const int leaf = col < tp.base ? col : col + 1;
        L.pc[x] = L.pw[e + ne * leaf] - L.pw[e + ne * tp.base];
    }
    __syncthreads();
    for (int x = tid; x < np * ne; x += kThreads) {
        const int k = x / ne, e = x - (x / ne) * ne;
        L.P[x] = L.pc[e + ne * ia[tp.o_cmb1 + k]] * L.pc[e + ne * ia[tp.o_cmb2 + k]];
    }
    __syncthreads();
    if (mode != 2) {
        // ---- W = G P, qf = G f ----
        for (int x = tid; x < np * ne; x += kThreads) {
            const int r = x / ne, e = x - (x / ne) * ne;
            const double* Gr = G + static_cast<std::int64_t>(r) * np;
            double acc = 0.0;
#pragma unroll 8
            for (int c = 0; c < np; ++c) acc += Gr[c] * L.P[c * ne + e];
            L.W[x] = acc;
        }
        for (int r = tid; r < np; r += kThreads) {
            const double* Gr = G + static_cast<std::int64_t>(r) * np;
            double acc = 0.0;
#pragma unroll 8
            for (int c = 0; c < np; ++c) acc += Gr[c] * f[c];
            L.qf[r] = acc;
        }
        __syncthreads();
        // ---- cc = P^T W, q = P^T qf ----
        for (int x = tid; x < ne * ne; x += kThreads) {
            const int e1 = x / ne, e2 = x - (x / ne) * ne;
            double acc = 0.0;
            for (int r = 0; r < np; ++r) acc += L.P[r * ne + e1] * L.W[r * ne + e2];
            L.C[x] = acc;
        }
        for (int e1 = tid; e1 < ne; e1 += kThreads) {
            double rr = 0.0;
            for (int r = 0; r < np; ++r) rr += L.P[r * ne + e1] * L.qf[r];
            L.q[e1] = rr;
        }
        __syncthreads();
        if (tid == 0) {
            double trm = 0.0;
            for (int e = 0; e < ne; ++e) trm += L.C[e * ne + e];
            const double ridge = g.fudge * trm / static_cast<double>(ne);
            for (int e = 0; e < ne; ++e) L.C[e * ne + e] += ridge;
            *sh.ok = 1;
        }
        __syncthreads();
        for (int e = tid; e < ne; e += kThreads) {
            L.sc[e] = sqrt(L.C[e * ne + e]);
            L.q1[e] = L.q[e] / L.sc[e];
        }
        __syncthreads();
        // ---- edge lengths (wavefront 0) ----
        if (wave == 0) {
            bool ok;
            if (g.constrained) {
                ok = nnls_wave<T>(ne, L.C, L.sc, L.q1, L.lam, sh.passive, sh.slot, warm);
            } else {
                double z = 0.0;
                ok = solve_sub<T>(L.C, ne, nullptr, ne, false, L.sc, lane < ne ? L.q1[lane] : 0.0, z);
                if (lane < ne) {
                    L.lam[lane] = z;
                    sh.passive[lane] = 1;
                }
            }
            if (lane == 0 && !ok) *sh.ok = 0;
        }
        __syncthreads();
        if (!*sh.ok) {
            __syncthreads();
            if (tid == 0) sh.scal[1] = __builtin_inf();
            __syncthreads();
            return false;
        }
        for (int e = tid; e < ne; e += kThreads) L.bl[e] = L.lam[e] / L.sc[e];
    }
    __syncthreads();
    // ---- residual and score ----
    for (int k = tid; k < np; k += kThreads) {
        double fv = 0.0;
        for (int e = 0; e < ne; ++e) fv += L.P[k * ne + e] * L.bl[e];
        L.res[k] = f[k] - fv;
        if (fit) fit[k] = fv;
    }
    __syncthreads();
    for (int a = tid; a < np; a += kThreads) {
        const double* Ga = G + static_cast<std::int64_t>(a) * np;
        double row = 0.0, col = 0.0;
#pragma unroll 8
        for (int b = 0; b < np; ++b) row += Ga[b] * L.res[b];
        L.Gr[a] = row;
        if (mode != 0) {
#pragma unroll 8
            for (int b = 0; b < np; ++b) col += G[static_cast<std::int64_t>(b) * np + a] * L.res[b];
            L.GTr[a] = col;
        }
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int a = 0; a < np; ++a) s += L.res[a] * L.Gr[a];
        sh.scal[1] = s;
    }
    if (mode == 0) {
        __syncthreads();
        return true;
    }
    // ---- adjoint ----
    for (int k = tid; k < np; k += kThreads) L.u[k] = L.Gr[k] + L.GTr[k];  // rbar
    __syncthreads();
    for (int e = tid; e < ne; e += kThreads) {
        double s = 0.0;
        for (int k = 0; k < np; ++k) s += L.P[k * ne + e] * L.u[k];
        L.blbar[e] = -s;
    }
    __syncthreads();
    if (mode == 2) {
        for (int x = tid; x < np * ne; x += kThreads) {
            const int k = x / ne, e = x - (x / ne) * ne;
            L.W[x] = -L.u[k] * L.bl[e];
        }
    } else {
        // lambda = C_AA^-T blbar_A on wavefront 0
        if (wave == 0) {
            const bool pas = lane < ne && sh.passive[lane] != 0;
            const unsigned long long mask = __ballot(pas);
            const int nA = __builtin_popcountll(mask);
            if (pas) sh.slot[__builtin_popcountll(mask & ((1ull << lane) - 1ull))] = lane;
            __builtin_amdgcn_wave_barrier();
            const int ea = lane < nA ? sh.slot[lane] : 0;
            double z = 0.0;
            const bool ok =
                nA == 0 || solve_sub<T>(L.C, ne, sh.slot, nA, true, nullptr, lane < nA ? L.blbar[ea] : 0.0, z);
            if (lane < ne) L.lam[lane] = 0.0;
            __builtin_amdgcn_wave_barrier();
            if (lane < nA) L.lam[ea] = ok ? z : 0.0;
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                double rho = 0.0;
                for (int e = 0; e < ne; ++e) rho -= L.lam[e] * L.bl[e];
                sh.scal[0] = g.fudge / static_cast<double>(ne) * rho;  // kappa
            }
        }
        __syncthreads();
        const double kappa = sh.scal[0];
        for (int c = tid; c < np; c += kThreads) {
            double s = 0.0;
            for (int e = 0; e < ne; ++e) s += L.P[c * ne + e] * L.lam[e];
            L.v[c] = s;  // P lambda
        }
        __syncthreads();
        for (int k = tid; k < np; k += kThreads) {
            double s = 0.0;
#pragma unroll 8
            for (int c = 0; c < np; ++c) s += G[static_cast<std::int64_t>(c) * np + k] * L.v[c];
            L.qf[k] = s;  // G^T P lambda
        }
        __syncthreads();
        for (int x = tid; x < np * ne; x += kThreads) {
            const int k = x / ne, e = x - (x / ne) * ne;
            double w2 = 0.0;
#pragma unroll 8
            for (int c = 0; c < np; ++c) w2 += G[static_cast<std::int64_t>(c) * np + k] * L.P[c * ne + e];
            L.W[x] = -L.bl[e] * (L.u[k] + L.qf[k]) + L.lam[e] * L.Gr[k] + kappa * (L.W[x] + w2);
        }
    }
    __syncthreads();
    // ---- Pbar -> pcbar -> pwbar -> pathbar -> thetabar ----
    for (int e = tid; e < ne; e += kThreads) {
        for (int c = 0; c < ncc; ++c) L.pcbar[e + ne * c] = 0.0;
        for (int k = 0; k < np; ++k) {
            const int a = ia[tp.o_cmb1 + k], b = ia[tp.o_cmb2 + k];
            const double pb = L.W[k * ne + e];
            L.pcbar[e + ne * a] += pb * L.pc[e + ne * b];
            L.pcbar[e + ne * b] += pb * L.pc[e + ne * a];
        }
        double sum = 0.0;
        for (int c = 0; c < ncc; ++c) {
            const int leaf = c < tp.base ? c : c + 1;
            L.pw[e + ne * leaf] = L.pcbar[e + ne * c];
            sum += L.pcbar[e + ne * c];
        }
        L.pw[e + ne * tp.base] = -sum;
    }
    __syncthreads();
    if (tid == 0) {
        for (int p = 0; p < npath; ++p) L.pathbar[p] = 0.0;
        for (int gr = 0; gr < tp.ngrp; ++gr) {
            const double v = L.pw[ia[tp.o_grp_edge + gr] + ne * ia[tp.o_grp_leaf + gr]];
            for (int t = ia[tp.o_grp_off + gr]; t < ia[tp.o_grp_off + gr + 1]; ++t) L.pathbar[ia[tp.o_grp_paths + t]] += v;
        }
        for (int d = 0; d < D; ++d) sh.thbar[d] = 0.0;
        for (int p = 0; p < npath; ++p) {
            const double pb = L.pathbar[p];
            if (pb == 0.0) continue;
            const int t0 = ia[tp.o_pae_off + p], t1 = ia[tp.o_pae_off + p + 1];
            for (int t = t0; t < t1; ++t) {
                double others = 1.0;
                for (int u = t0; u < t1; ++u) {
                    if (u == t) continue;
                    const int id = ia[tp.o_pae_id + u];
                    const double th = sh.theta[(id - 1) / 2];
                    others *= (id % 2 == 1) ? th : (1.0 - th);
                }
                const int id = ia[tp.o_pae_id + t];
                sh.thbar[(id - 1) / 2] += (id % 2 == 1 ? pb : -pb) * others;
            }
        }
    }
    __syncthreads();
    return true;
}

template <int T>
__global__ __launch_bounds__(kThreads) void qg_eval_kernel(QgEvalArgs g) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    __shared__ int s_inst, s_ok;
    __shared__ int s_passive[64], s_slot[64];
    __shared__ double s_theta[kMaxD], s_thbar[kMaxD], s_scal[4];
    const int tid = threadIdx.x;
    const EvalLds L = eval_layout(lds, g);
    const EvalShared sh{s_passive, s_slot, s_theta, s_thbar, s_scal, &s_ok};
    for (;;) {
        if (tid == 0) s_inst = static_cast<int>(atomicAdd(g.counter, 1u));
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (i >= g.n_inst) break;
        const QgTopoDev tp = g.topo[g.inst_topo[i]];
        const int ne = tp.ne, D = tp.nadmix;
        const double* G = g.qinv + g.inst_qinv[i];
        const double* f = g.fobs + g.inst_fobs[i];
        if (tid < D) s_theta[tid] = g.theta[static_cast<std::int64_t>(i) * g.dmax + tid];
        if (g.mode == 2)
            for (int e = tid; e < ne; e += kThreads) L.bl[e] = g.bl_in[static_cast<std::int64_t>(i) * g.emax + e];
        __syncthreads();
        double* fit = g.fit_out ? g.fit_out + static_cast<std::int64_t>(i) * g.np_max : nullptr;
        const bool ok = qg_eval_one<T>(g, tp, G, f, g.mode, L, sh, fit);
        if (tid == 0) g.score[i] = s_scal[1];
        if (ok) {
            if (g.bl_out)
                for (int e = tid; e < ne; e += kThreads) g.bl_out[static_cast<std::int64_t>(i) * g.emax + e] = L.bl[e];
            if (g.mode != 0 && tid < D) g.grad_theta[static_cast<std::int64_t>(i) * g.dmax + tid] = s_thbar[tid];
            if (g.mode == 2)
                for (int e = tid; e < ne; e += kThreads)
                    g.grad_bl[static_cast<std::int64_t>(i) * g.emax + e] = L.blbar[e];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// Projected L-BFGS on [0, 1]^D, one instance per workgroup, to convergence on the
// device (replacing the reference's finite-difference coordinate search,
// cpu_backend.cpp:2936-3045).  Per iteration: variables at a bound whose gradient
// points outward are held; the two-loop recursion over the last kHist pairs gives the
// direction on the free variables (steepest descent when it is not a descent
// direction, with the history reset); the projected step is accepted by Armijo
// backtracking (c1 = 1e-4, halving, at most 30 trials).  Stops when the projected
// gradient is below gtol, the step below xtol, the relative decrease below ftol, the
// line search fails, or after maxit iterations.  The L-BFGS bookkeeping (D <= 16) runs
// on one thread; every evaluation uses the whole workgroup.
// ---------------------------------------------------------------------------------
constexpr int kHist = 8;
constexpr int kMaxBacktrack = 30;
constexpr double kArmijo = 1e-4;

template <int T>
__global__ __launch_bounds__(kThreads) void qg_fit_kernel(QgEvalArgs g, QgFitArgs fa) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    __shared__ int s_inst, s_ok, s_stop, s_acc;
    __shared__ int s_passive[64], s_slot[64];
    __shared__ double s_theta[kMaxD], s_thbar[kMaxD], s_scal[4];
    __shared__ double s_x[kMaxD], s_gx[kMaxD], s_d[kMaxD], s_S[kHist][kMaxD], s_Y[kHist][kMaxD], s_rho[kHist];
    // thread 0's L-BFGS work vectors (LDS rather than registers: the evaluation
    // already needs most of the register file)
    __shared__ double s_pg[kMaxD], s_q[kMaxD], s_alpha[kHist];
    __shared__ int s_held[kMaxD];
    __shared__ double s_f;
    const int tid = threadIdx.x;
    const EvalLds L = eval_layout(lds, g);
    const EvalShared sh{s_passive, s_slot, s_theta, s_thbar, s_scal, &s_ok};
    for (;;) {
        if (tid == 0) s_inst = static_cast<int>(atomicAdd(g.counter, 1u));
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(s_inst);
        __syncthreads();
        if (i >= g.n_inst) break;
        const QgTopoDev tp = g.topo[g.inst_topo[i]];
        const int D = tp.nadmix;
        const double* G = g.qinv + g.inst_qinv[i];
        const double* f = g.fobs + g.inst_fobs[i];
        if (tid < D) {
            const double x0 = fa.theta0[static_cast<std::int64_t>(i) * g.dmax + tid];
            s_x[tid] = s_theta[tid] = x0 < 0.0 ? 0.0 : (x0 > 1.0 ? 1.0 : x0);
        }
        __syncthreads();
        int iters = 0, status = 0;
        if (!qg_eval_one<T>(g, tp, G, f, D > 0 ? 1 : 0, L, sh, nullptr)) {
            status = 1;
        } else if (D > 0) {
            if (tid == 0) {
                s_f = s_scal[1];
                for (int d = 0; d < D; ++d) s_gx[d] = s_thbar[d];
            }
            int hist = 0;  // pairs stored so far (uniform: every thread tracks it)
            for (int it = 0; it < fa.maxit; ++it) {
                __syncthreads();
                if (tid == 0) {
                    int* held = s_held;
                    double *pg = s_pg, *q = s_q, *alpha = s_alpha;
                    double pgmax = 0.0;
                    for (int d = 0; d < D; ++d) {
                        held[d] = (s_x[d] <= 0.0 && s_gx[d] > 0.0) || (s_x[d] >= 1.0 && s_gx[d] < 0.0);
                        pg[d] = held[d] ? 0.0 : s_gx[d];
                        q[d] = pg[d];
                        pgmax = fmax(pgmax, fabs(pg[d]));
                    }
                    s_stop = pgmax <= fa.gtol;
                    const int k = hist < kHist ? hist : kHist;
                    for (int a = 0; a < k; ++a) {  // newest to oldest
                        const int sl = (hist - 1 - a) % kHist;
                        double sq = 0.0;
                        for (int d = 0; d < D; ++d)
                            if (!held[d]) sq += s_S[sl][d] * q[d];
                        alpha[a] = s_rho[sl] * sq;
                        for (int d = 0; d < D; ++d)
                            if (!held[d]) q[d] -= alpha[a] * s_Y[sl][d];
                    }
                    double gamma = 1.0;
                    if (hist > 0) {
                        const int sl = (hist - 1) % kHist;
                        double sy = 0.0, yy = 0.0;
                        for (int d = 0; d < D; ++d)
                            if (!held[d]) {
                                sy += s_S[sl][d] * s_Y[sl][d];
                                yy += s_Y[sl][d] * s_Y[sl][d];
                            }
                        if (yy > 0.0) gamma = sy / yy;
                        if (!(gamma > 0.0)) gamma = 1.0;
                    }
                    for (int d = 0; d < D; ++d) q[d] *= gamma;
                    for (int a = k - 1; a >= 0; --a) {  // oldest to newest
                        const int sl = (hist - 1 - a) % kHist;
                        double yr = 0.0;
                        for (int d = 0; d < D; ++d)
                            if (!held[d]) yr += s_Y[sl][d] * q[d];
                        const double b = s_rho[sl] * yr;
                        for (int d = 0; d < D; ++d)
                            if (!held[d]) q[d] += (alpha[a] - b) * s_S[sl][d];
                    }
                    double slope = 0.0;
                    for (int d = 0; d < D; ++d) {
                        s_d[d] = held[d] ? 0.0 : -q[d];
                        slope += s_d[d] * pg[d];
                    }
                    s_acc = 0;
                    if (slope >= 0.0) {
                        for (int d = 0; d < D; ++d) s_d[d] = -pg[d];
                        s_acc = 1;  // history reset
                    }
                }
                __syncthreads();
                if (s_stop) break;
                if (s_acc) hist = 0;
                double step = 1.0;
                bool accepted = false;
                for (int bt = 0; bt < kMaxBacktrack; ++bt) {
                    __syncthreads();
                    if (tid < D) {
                        const double xn = s_x[tid] + step * s_d[tid];
                        s_theta[tid] = xn < 0.0 ? 0.0 : (xn > 1.0 ? 1.0 : xn);
                    }
                    __syncthreads();
                    const bool ok = qg_eval_one<T>(g, tp, G, f, 1, L, sh, nullptr, true);
                    if (tid == 0) {
                        double dec = 0.0;
                        for (int d = 0; d < D; ++d) dec += s_gx[d] * (s_theta[d] - s_x[d]);
                        s_acc = ok && s_scal[1] <= s_f + kArmijo * dec;
                    }
                    __syncthreads();
                    if (s_acc) {
                        accepted = true;
                        break;
                    }
                    step *= 0.5;
                }
                if (!accepted) break;
                __syncthreads();
                if (tid == 0) {
                    double sy = 0.0, ss = 0.0, yy = 0.0, dx = 0.0;
                    double *sv = s_pg, *yv = s_q;  // the direction's work vectors are free now
                    for (int d = 0; d < D; ++d) {
                        sv[d] = s_theta[d] - s_x[d];
                        yv[d] = s_thbar[d] - s_gx[d];
                        sy += sv[d] * yv[d];
                        ss += sv[d] * sv[d];
                        yy += yv[d] * yv[d];
                        dx = fmax(dx, fabs(sv[d]));
                    }
                    s_acc = sy > 1e-16 * sqrt(ss) * sqrt(yy);
                    if (s_acc) {
                        const int sl = hist % kHist;
                        for (int d = 0; d < D; ++d) {
                            s_S[sl][d] = sv[d];
                            s_Y[sl][d] = yv[d];
                        }
                        s_rho[sl] = 1.0 / sy;
                    }
                    const double fn = s_scal[1];
                    const double df = s_f - fn;
                    for (int d = 0; d < D; ++d) {
                        s_x[d] = s_theta[d];
                        s_gx[d] = s_thbar[d];
                    }
                    s_f = fn;
                    s_stop = dx <= fa.xtol || df <= fa.ftol * (1.0 + fabs(fn));
                }
                __syncthreads();
                if (s_acc) ++hist;
                ++iters;
                if (s_stop) break;
            }
        }
        __syncthreads();
        if (tid == 0) {
            fa.score_out[i] = status ? __builtin_inf() : (D > 0 ? s_f : s_scal[1]);
            fa.iters_out[i] = status ? -1 : iters;
        }
        if (tid < D) fa.theta_out[static_cast<std::int64_t>(i) * g.dmax + tid] = s_x[tid];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------
// f3 basis X[k][s] = (f2(c, a_k) + f2(c, b_k) - f2(a_k, b_k)) / 2 over the surviving
// blocks of each job's tensor (cpu_backend.cpp:2170-2217), then LOO / total and the
// jackknife covariance with the diag_f3 ridge.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void qg_basis_kernel(QgBasisArgs g) {
    extern __shared__ __attribute__((aligned(16))) double lds[];
    __shared__ int s_job, s_nbs, s_flag;
    __shared__ int s_idx[2 * (kThreads / 64) + 2];
    __shared__ double s_val[kThreads / 64 + 1];
    __shared__ double s_ntot;
    const int tid = threadIdx.x;
    const int NP = g.np_max;
    double* sLU = lds;                                  // NP x (NP + 1)
    double* s_xtot = sLU + NP * (NP + 1);
    double* s_totline = s_xtot + NP;
    double* s_qdiag = s_totline + NP;
    int* s_pos = reinterpret_cast<int*>(s_qdiag + NP);  // NP
    int* s_surv = s_pos + NP;                           // nb_max
    int* s_bsz = s_surv + g.nb_max;                     // nb_max
    double* X = g.scratch + static_cast<std::int64_t>(blockIdx.x) * g.scratch_stride;
    for (;;) {
        if (tid == 0) s_job = static_cast<int>(atomicAdd(g.counter, 1u));
        __syncthreads();
        const int j = __builtin_amdgcn_readfirstlane(s_job);
        __syncthreads();
        if (j >= g.n_jobs) break;
        const QgTopoDev tp = g.topo[g.job_topo[j]];
        const int P = g.job_P[j], nb = g.job_nb[j], m = tp.npair;
        const double* f2 = g.f2[j];
        const double* vp = g.vpair[j];
        const std::int32_t* bsz = g.bsz[j];
        const std::int64_t slab = static_cast<std::int64_t>(P) * P;
        // survivors: a block is dropped when some pairs are missing and others present
        for (int b = tid; b < nb; b += kThreads) {
            bool pres = false, miss = false;
            if (vp)
                for (std::int64_t e = 0; e < slab && !(pres && miss); ++e) {
                    if (vp[b * slab + e] > 0.0) pres = true;
                    else miss = true;
                }
            s_surv[b] = (pres && miss) ? 0 : 1;
        }
        __syncthreads();
        if (tid == 0) {
            int n = 0;
            double tot = 0.0;
            for (int b = 0; b < nb; ++b)
                if (s_surv[b]) {
                    s_surv[n] = b;
                    s_bsz[n] = bsz[b];
                    tot += static_cast<double>(bsz[b]);
                    ++n;
                }
            s_nbs = n;
            s_ntot = tot;
        }
        __syncthreads();
        const int nbs = s_nbs;
        double* LOO = X + static_cast<std::int64_t>(NP) * g.nb_max;
        if (nbs < 1 || m < 1) {
            if (tid == 0) g.status[j] = 3;
            __syncthreads();
            continue;
        }
        const int* l2f = g.job_leaf_to_f2 + static_cast<std::int64_t>(j) * g.npop_max;
        const int pc = l2f[tp.base];
        for (int x = tid; x < m * nbs; x += kThreads) {
            const int k = x / nbs, s = x - (x / nbs) * nbs;
            const int ca = g.iarena[tp.o_cmb1 + k], cb = g.iarena[tp.o_cmb2 + k];
            const int pa = l2f[ca < tp.base ? ca : ca + 1], pb = l2f[cb < tp.base ? cb : cb + 1];
            const double* blk = f2 + static_cast<std::int64_t>(s_surv[s]) * slab;
            X[static_cast<std::int64_t>(k) * nbs + s] =
                0.5 * (blk[pc + static_cast<std::int64_t>(P) * pa] + blk[pc + static_cast<std::int64_t>(P) * pb] -
                       blk[pa + static_cast<std::int64_t>(P) * pb]);
        }
        __syncthreads();
        double* qinv = g.qinv + static_cast<std::int64_t>(j) * NP * NP;
        const bool ok = jk_covariance<kThreads>(X, LOO, s_bsz, s_ntot, nbs, m, g.diag_f3, s_xtot, s_totline, s_qdiag,
                                                sLU, NP + 1, qinv, m, s_pos, s_idx, s_val, &s_flag);
        for (int k = tid; k < m; k += kThreads) {
            g.fobs[static_cast<std::int64_t>(j) * NP + k] = s_xtot[k];
            g.qdiag[static_cast<std::int64_t>(j) * NP + k] = s_qdiag[k];
        }
        if (tid == 0) g.status[j] = ok ? 0 : 3;
        __syncthreads();
    }
}

}  // namespace

std::size_t qg_eval_lds_bytes(const QgEvalArgs& a) {
    const std::size_t NP = a.np_max, NE = a.ne_max, NPOP = a.npop_max, NPATH = a.npath_max;
    const std::size_t n = 2 * NPATH + NE * NPOP + 2 * NE * (NPOP - 1) + 2 * NP * NE + NE * NE + 6 * NE + 6 * NP;
    return n * sizeof(double);
}

int qg_blocks_per_cu(int T, bool fit, std::size_t lds) {
    int n = 0;
    auto q = [&](auto kernel) {
        if (lds > 64 * 1024)
            STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        STP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, kThreads, lds));
    };
    if (T == 32) fit ? q(qg_fit_kernel<32>) : q(qg_eval_kernel<32>);
    else fit ? q(qg_fit_kernel<64>) : q(qg_eval_kernel<64>);
    return n > 0 ? n : 1;
}

void launch_qg_eval(int T, const QgEvalArgs& a, int n_workgroups, hipStream_t st) {
    if (a.n_inst <= 0) return;
    const std::size_t lds = qg_eval_lds_bytes(a);
    auto go = [&](auto kernel) {
        if (lds > 64 * 1024)
            STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        kernel<<<static_cast<unsigned>(n_workgroups), kThreads, lds, st>>>(a);
    };
    switch (T) {
        case 32: go(qg_eval_kernel<32>); break;
        case 64: go(qg_eval_kernel<64>); break;
        default: throw ArgError("qpgraph: unsupported edge-count bucket " + std::to_string(T));
    }
    STP_LAUNCH_CHECK();
}

void launch_qg_fit(int T, const QgEvalArgs& a, const QgFitArgs& fa, int n_workgroups, hipStream_t st) {
    if (a.n_inst <= 0) return;
    const std::size_t lds = qg_eval_lds_bytes(a);
    auto go = [&](auto kernel) {
        if (lds > 64 * 1024)
            STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
        kernel<<<static_cast<unsigned>(n_workgroups), kThreads, lds, st>>>(a, fa);
    };
    switch (T) {
        case 32: go(qg_fit_kernel<32>); break;
        case 64: go(qg_fit_kernel<64>); break;
        default: throw ArgError("qpgraph: unsupported edge-count bucket " + std::to_string(T));
    }
    STP_LAUNCH_CHECK();
}

std::size_t qg_basis_lds_bytes(const QgBasisArgs& a) {
    const std::size_t NP = a.np_max;
    return (NP * (NP + 1) + 3 * NP) * sizeof(double) + (NP + 2 * static_cast<std::size_t>(a.nb_max)) * sizeof(int);
}

void launch_qg_basis(const QgBasisArgs& a, int n_workgroups, hipStream_t st) {
    if (a.n_jobs <= 0) return;
    const std::size_t lds = qg_basis_lds_bytes(a);
    if (lds > 64 * 1024)
        STP_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(qg_basis_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    qg_basis_kernel<<<static_cast<unsigned>(n_workgroups), kThreads, lds, st>>>(a);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
