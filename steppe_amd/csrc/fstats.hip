// steppe_amd/csrc/fstats.hip — standalone f-statistics over the resident f2 tensor:
// f4 / f3 per statistic, the f4-ratio, and the direct form of the all-combination
// sweeps.
//
// One wavefront owns one statistic; its lanes own surviving blocks b = lane + 64 i, so
// every f2 series is read with coalesced 512-byte loads from the survivor-transposed
// tensor f2T[(i + P*j) * nbs + s].  The block series X_b stays in registers (kPer
// values per lane) through the leave-one-out and jackknife passes; tensors with more
// than 64 * 32 surviving blocks take a streaming variant that re-reads the series
// each pass.  The arithmetic follows the reference's CPU backend:
//   f4 block value            src/device/cpu/cpu_backend.cpp:2154-2156
//   f3 block value            src/device/cpu/cpu_backend.cpp:2204-2205
//   compute_loo_and_total     src/device/cpu/cpu_backend.cpp:2674-2723
//   jackknife_diag            src/device/cpu/cpu_backend.cpp:2273-2303
//   read-out est / se / z / p src/core/qpadm/point_estimates.hpp:28-44, f4.cpp:94-97
//   ratio_block_jackknife     src/device/cpu/cpu_backend.cpp:2309-2378 (tot_mode 0) with
//                             f4ratio_blocks_jackknife's inputs (2468-2497)
// with double-double sums where the reference accumulates in long double.  This file
// is built with FP contraction off, so every product is rounded before it is added,
// as in the reference.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "common.hpp"
#include "fstats.hpp"
#include "wave_linalg.hpp"

namespace stp {
namespace {

using namespace wl;

constexpr int kWaves = 4;                 // statistics per workgroup
constexpr int kThreads = 64 * kWaves;
constexpr int kMaxPer = 32;               // cached series: nbs <= 64 * kMaxPer

// Sum over the wavefront (butterfly), then lane 0's value everywhere.
__device__ inline dd wave_sum_dd(dd v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const dd o{__shfl_xor(v.hi, off), __shfl_xor(v.lo, off)};
        v = dd_add_dd(v, o);
    }
    return dd{uni(v.hi), uni(v.lo)};
}

struct Est {
    double est, se, z;
};

__device__ inline Est read_out(double xt, double var) {
    const double se = var > 0.0 ? sqrt(var) : __builtin_nan("");
    return Est{xt, se, xt / se};
}

// The f4 / f3 block value from the statistic's f2 rows.
template <int kW>
struct Rows {
    const double* r[kW];
    __device__ double x(int b) const {
        if constexpr (kW == 4) return 0.5 * (((r[0][b] + r[1][b]) - r[2][b]) - r[3][b]);
        else return 0.5 * ((r[0][b] + r[1][b]) - r[2][b]);
    }
};

__device__ inline const double* row(const FstatTensor& t, int i, int j) {
    return t.f2T + (static_cast<std::int64_t>(i) + static_cast<std::int64_t>(t.P) * j) * t.nbs;
}

// f4 of (p1, p2, p3, p4): 0.5 (f2(p2,p3) + f2(p1,p4) - f2(p1,p3) - f2(p2,p4)).
__device__ inline Rows<4> f4_rows(const FstatTensor& t, int p1, int p2, int p3, int p4) {
    return Rows<4>{{row(t, p2, p3), row(t, p1, p4), row(t, p1, p3), row(t, p2, p4)}};
}
// f3 of (C; A, B): 0.5 (f2(C,A) + f2(C,B) - f2(A,B)).
__device__ inline Rows<3> f3_rows(const FstatTensor& t, int c, int a, int b) {
    return Rows<3>{{row(t, c, a), row(t, c, b), row(t, a, b)}};
}

// Leave-one-out values, total and jackknife variance of one statistic's block series
// (every lane of the wavefront calls it; the result is uniform).
template <int kPer, class R>
__device__ Est jackknife_stat(const R& rows, const FstatTensor& t) {
    const int lane = lane_id();
    const int nbs = t.nbs;
    const double n = t.n_total;
    const std::int32_t* bsz = t.bsz;
    if constexpr (kPer > 0) {
        double x[kPer];
        dd s{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = lane + 64 * i;
            x[i] = 0.0;
            if (b < nbs) {
                x[i] = rows.x(b);
                s = dd_fma(s, x[i], static_cast<double>(bsz[b]));
            }
        }
        const double tot = uni(dd_div_d(wave_sum_dd(s), n));
        dd wn{0.0, 0.0}, wd{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = lane + 64 * i;
            if (b < nbs) {
                const double bl = static_cast<double>(bsz[b]);
                const double rel = bl / n;
                x[i] = (tot - x[i] * rel) / (1.0 - rel);
                const double w = 1.0 - bl / n;
                wn = dd_fma(wn, x[i], w);
                wd = dd_add(wd, w);
            }
        }
        const double tl = uni(dd_div_dd(wave_sum_dd(wn), wave_sum_dd(wd)));
        dd ds{0.0, 0.0}, wb{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = lane + 64 * i;
            if (b < nbs) {
                ds = dd_add(dd_add(ds, tl), -x[i]);
                wb = dd_fma(wb, x[i], static_cast<double>(bsz[b]));
            }
        }
        const double term1 = dd_div_d(wave_sum_dd(ds), static_cast<double>(nbs)) * static_cast<double>(nbs);
        const double xt = uni(term1 + dd_div_d(wave_sum_dd(wb), n));
        dd acc{0.0, 0.0};
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int b = lane + 64 * i;
            if (b < nbs) {
                const double h = n / static_cast<double>(bsz[b]);
                const double v = (xt * h - x[i] * (h - 1.0) - tl) / sqrt(h - 1.0);
                acc = dd_fma(acc, v, v);
            }
        }
        return read_out(xt, uni(dd_div_d(wave_sum_dd(acc), static_cast<double>(nbs))));
    } else {
        // Streaming: the series is re-read (and the LOO values recomputed) each pass.
        dd s{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) s = dd_fma(s, rows.x(b), static_cast<double>(bsz[b]));
        const double tot = uni(dd_div_d(wave_sum_dd(s), n));
        auto loo = [&](int b) {
            const double rel = static_cast<double>(bsz[b]) / n;
            return (tot - rows.x(b) * rel) / (1.0 - rel);
        };
        dd wn{0.0, 0.0}, wd{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) {
            const double w = 1.0 - static_cast<double>(bsz[b]) / n;
            wn = dd_fma(wn, loo(b), w);
            wd = dd_add(wd, w);
        }
        const double tl = uni(dd_div_dd(wave_sum_dd(wn), wave_sum_dd(wd)));
        dd ds{0.0, 0.0}, wb{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) {
            const double l = loo(b);
            ds = dd_add(dd_add(ds, tl), -l);
            wb = dd_fma(wb, l, static_cast<double>(bsz[b]));
        }
        const double term1 = dd_div_d(wave_sum_dd(ds), static_cast<double>(nbs)) * static_cast<double>(nbs);
        const double xt = uni(term1 + dd_div_d(wave_sum_dd(wb), n));
        dd acc{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) {
            const double h = n / static_cast<double>(bsz[b]);
            const double v = (xt * h - loo(b) * (h - 1.0) - tl) / sqrt(h - 1.0);
            acc = dd_fma(acc, v, v);
        }
        return read_out(xt, uni(dd_div_d(wave_sum_dd(acc), static_cast<double>(nbs))));
    }
}

// Workgroups are spread so that each of the eight dies (workgroup i runs on die i % 8)
// takes one contiguous range of statistics: neighbouring statistics share f2 rows, which
// then meet in the same L2.  Returns -1 for the few padding workgroups.
__device__ inline std::int64_t die_major_group(std::int64_t n_groups) {
    const std::int64_t per = (n_groups + 7) / 8;
    const std::int64_t g = static_cast<std::int64_t>(blockIdx.x % 8) * per + blockIdx.x / 8;
    return g < n_groups ? g : -1;
}

__device__ inline std::int64_t my_item(std::int64_t n_items) {
    const std::int64_t g = die_major_group((n_items + kWaves - 1) / kWaves);
    if (g < 0) return -1;
    const std::int64_t k = g * kWaves + __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    return k < n_items ? k : -1;
}

__device__ inline double two_sided_p(double z) { return erfc(fabs(z) * 0.70710678118654752440); }

template <int kW, int kPer>
__global__ __launch_bounds__(kThreads) void fstat_kernel(FstatArgs a) {
    const std::int64_t k = my_item(a.N);
    if (k < 0) return;
    const std::int32_t* q = a.idx + kW * k;
    Est e;
    if constexpr (kW == 4) {
        const auto rows = f4_rows(a.t, __builtin_amdgcn_readfirstlane(q[0]), __builtin_amdgcn_readfirstlane(q[1]),
                                  __builtin_amdgcn_readfirstlane(q[2]), __builtin_amdgcn_readfirstlane(q[3]));
        e = jackknife_stat<kPer>(rows, a.t);
    } else {
        const auto rows = f3_rows(a.t, __builtin_amdgcn_readfirstlane(q[0]), __builtin_amdgcn_readfirstlane(q[1]),
                                  __builtin_amdgcn_readfirstlane(q[2]));
        e = jackknife_stat<kPer>(rows, a.t);
    }
    if (lane_id() == 0) {
        a.est[k] = e.est;
        a.se[k] = e.se;
        a.z[k] = e.z;
        a.p[k] = two_sided_p(e.z);
    }
}

// f4-ratio of (p1, p2, p3, p4, p5): numerator f4(p1,p2;p3,p4), denominator f4(p1,p2;p5,p4).
// Blocks whose denominator block value is below 1e-6 in magnitude are set missing for the
// tuple; the per-block ratio is of the leave-one-out values (taken over all surviving
// blocks, as the reference's assemble step computes them).
__global__ __launch_bounds__(kThreads) void f4ratio_kernel(F4RatioArgs a) {
    constexpr double kSetmiss = 1e-6;
    const std::int64_t k = my_item(a.N);
    if (k < 0) return;
    const FstatTensor& t = a.t;
    const std::int32_t* q = a.tuples + 5 * k;
    const int p1 = __builtin_amdgcn_readfirstlane(q[0]), p2 = __builtin_amdgcn_readfirstlane(q[1]);
    const int p3 = __builtin_amdgcn_readfirstlane(q[2]), p4 = __builtin_amdgcn_readfirstlane(q[3]);
    const int p5 = __builtin_amdgcn_readfirstlane(q[4]);
    const auto num = f4_rows(t, p1, p2, p3, p4);
    const auto den = f4_rows(t, p1, p2, p5, p4);
    const int lane = lane_id();
    const int nbs = t.nbs;
    const double n_all = t.n_total;
    dd sn{0.0, 0.0}, sd{0.0, 0.0};
    for (int b = lane; b < nbs; b += 64) {
        const double bl = static_cast<double>(t.bsz[b]);
        sn = dd_fma(sn, num.x(b), bl);
        sd = dd_fma(sd, den.x(b), bl);
    }
    const double tot_n = uni(dd_div_d(wave_sum_dd(sn), n_all));
    const double tot_d = uni(dd_div_d(wave_sum_dd(sd), n_all));
    auto loo = [&](double tot, double x, int b) {
        const double rel = static_cast<double>(t.bsz[b]) / n_all;
        return (tot - x * rel) / (1.0 - rel);
    };
    // kept blocks: count, size and the weighted block totals
    int kept = 0;
    double nk = 0.0;
    dd tn{0.0, 0.0}, td{0.0, 0.0};
    for (int b = lane; b < nbs; b += 64) {
        const double xd = den.x(b);
        if (fabs(xd) < kSetmiss) continue;
        const double bl = static_cast<double>(t.bsz[b]);
        ++kept;
        nk += bl;
        tn = dd_fma(tn, num.x(b), bl);
        td = dd_fma(td, xd, bl);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        kept += __shfl_xor(kept, off);
        nk += __shfl_xor(nk, off);   // integer-valued sums below 2^53: exact in any order
    }
    kept = __builtin_amdgcn_readfirstlane(kept);
    nk = uni(nk);
    const dd totnum = wave_sum_dd(tn), totden = wave_sum_dd(td);
    double alpha = __builtin_nan(""), se = __builtin_nan(""), z = __builtin_nan("");
    if (kept > 1 && nk > 0.0 && totden.hi != 0.0) {
        const double tot = uni(dd_div_dd(totnum, totden));
        dd diff{0.0, 0.0}, wnum{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) {
            const double xd = den.x(b);
            if (fabs(xd) < kSetmiss) continue;
            const double bl = static_cast<double>(t.bsz[b]);
            const double R = loo(tot_n, num.x(b), b) / loo(tot_d, xd, b);
            diff = dd_add(dd_add(diff, tot), -R);
            wnum = dd_fma(wnum, R, bl);
        }
        const double kd = static_cast<double>(kept);
        const double est = uni(dd_div_d(wave_sum_dd(diff), kd) * kd + dd_div_d(wave_sum_dd(wnum), nk));
        dd acc{0.0, 0.0};
        for (int b = lane; b < nbs; b += 64) {
            const double xd = den.x(b);
            if (fabs(xd) < kSetmiss) continue;
            const double bl = static_cast<double>(t.bsz[b]);
            const double R = loo(tot_n, num.x(b), b) / loo(tot_d, xd, b);
            const double h = nk / bl;
            const double tau = h * tot - (h - 1.0) * R;
            const double v = (tau - est) / sqrt(h - 1.0);
            acc = dd_fma(acc, v, v);
        }
        const Est e = read_out(est, uni(dd_div_d(wave_sum_dd(acc), kd)));
        alpha = e.est;
        se = e.se;
        z = e.z;
    }
    if (lane == 0) {
        a.alpha[k] = alpha;
        a.se[k] = se;
        a.z[k] = z;
    }
}

// Combinatorial-number-system unrank: r -> c[0] < ... < c[k-1] with r = sum_i C(c[i], i+1);
// each digit is the largest v with C(v, i+1) <= the remaining rank (binary search).
template <int kK>
__device__ inline void unrank(std::int64_t r, int range, const std::int64_t* binom, int (&c)[kK]) {
    int hi = range - 1;
#pragma unroll
    for (int pos = kK - 1; pos >= 0; --pos) {
        const int kk = pos + 1;
        const std::int64_t* C = binom + static_cast<std::int64_t>(kk) * (range + 1);
        int lo = kk - 1;   // C(kk-1, kk) = 0 <= r
        int h = hi;
        while (lo < h) {
            const int mid = (lo + h + 1) >> 1;
            if (C[mid] <= r) lo = mid;
            else h = mid - 1;
        }
        c[pos] = lo;
        r -= C[lo];
        hi = lo - 1;
    }
}

template <int kK, int kPer>
__global__ __launch_bounds__(kThreads) void sweep_direct_kernel(SweepArgs a) {
    const std::int64_t i = my_item(a.count);
    if (i < 0) return;
    const std::int64_t r = a.ranks ? a.ranks[i] : a.r0 + i;
    int c[kK];
    unrank<kK>(r, a.range, a.binom, c);
#pragma unroll
    for (int j = 0; j < kK; ++j) {
        c[j] = a.subset ? a.subset[c[j]] : c[j];
        c[j] = __builtin_amdgcn_readfirstlane(c[j]);
    }
    Est e;
    if constexpr (kK == 4) e = jackknife_stat<kPer>(f4_rows(a.t, c[0], c[1], c[2], c[3]), a.t);
    else e = jackknife_stat<kPer>(f3_rows(a.t, c[0], c[1], c[2]), a.t);
    if (lane_id() == 0 && fabs(e.z) >= a.tau) {
        const unsigned long long slot = atomicAdd(a.n_hits, 1ULL);
        if (static_cast<std::int64_t>(slot) < a.cap) a.hits[slot] = SweepHit{r, e.est, e.se, e.z};
    }
}

// ---------------------------------------------------------------------------------
// Weighted block means of every f2 series (one wavefront per series).
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void f2_means_kernel(FstatTensor t, double* __restrict__ mean) {
    const std::int64_t k = my_item(static_cast<std::int64_t>(t.P) * t.P);
    if (k < 0) return;
    const double* r = t.f2T + k * t.nbs;
    dd s{0.0, 0.0};
    for (int b = lane_id(); b < t.nbs; b += 64) s = dd_fma(s, r[b], static_cast<double>(t.bsz[b]));
    const double m = dd_div_d(wave_sum_dd(s), t.n_total);
    if (lane_id() == 0) mean[k] = m;
}

// ---------------------------------------------------------------------------------
// Gram form of the f4 sweep: one workgroup per item (c2, c3, row block bi, row block bj).
//
// The centred, weighted series e_b(p) of the item's rows are staged 16 blocks at a time
// into LDS as sE[k][slot] (slot < 128: rows of block bi, slot >= 128: rows of block bj);
// each wavefront owns up to 8 16x16 tiles of S = E E^T and accumulates them with
// v_mfma_f64_16x16x4f64 (A = rows of bi, B = rows of bj, both read from the same LDS
// chunk).  The staging threads also accumulate the diagonal D(p) = sum_b e_b(p)^2.  The
// next chunk's f2 values are loaded into registers before the current chunk's MFMAs.
// Each f4 (c0 < c1) of a tile is then read out.  Relative to the reference's operation
// order the estimate differs by the rounding of the four f2 means it is formed from (a few
// ulps of their magnitude, the level of the reference's own block-value rounding) and the
// variance by the Gram sums' rounding amplified by the cancellation (D0 + D1) / den.  A
// statistic goes to the recheck list for the direct kernel instead when that cancellation
// reaches max_cancel, when the f2 means exceed max_level times |est| + se (copied or nearly
// copied populations), or when the variance is 0 / 0; the others are then within 1e-9 of
// the reference's se and 1e-9 se (plus the means' rounding) of its estimate.
// ---------------------------------------------------------------------------------
constexpr int kGramWaves = 8;
constexpr int kGramThreads = 64 * kGramWaves;
constexpr int kGramK = 16;                                  // blocks per staged chunk
constexpr int kGramLd = 2 * kGramRows + 1;                  // LDS stride of one block column
constexpr int kGramSlots = 2 * kGramRows * kGramK / kGramThreads;   // staged rows per thread (8)
constexpr int kGramMaxTiles = 8;                            // 64 tiles of a 128 x 128 block / 8 waves

typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ inline void wave_append(bool pred, unsigned long long* counter, std::int64_t cap, int lane,
                                   unsigned long long& slot_out, bool& ok) {
    const unsigned long long bal = __ballot(pred);
    ok = false;
    if (bal == 0) return;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(counter, static_cast<unsigned long long>(__popcll(bal)));
    base = __shfl(base, 0);
    const unsigned long long off = __popcll(bal & ((1ULL << lane) - 1ULL));
    slot_out = base + off;
    ok = pred && static_cast<std::int64_t>(slot_out) < cap;
}

__global__ __launch_bounds__(kGramThreads) void f4_gram_kernel(GramArgs a) {
    __shared__ double sE[kGramK * kGramLd];
    __shared__ double sG[2 * kGramRows], sD[2 * kGramRows], sL[2 * kGramRows];
    const std::int64_t item = die_major_group(a.n_items);
    if (item < 0) return;
    const GramItem it = a.items[item];
    const FstatTensor& t = a.t;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int P = t.P, nbs = t.nbs;
    const int m = it.c2;
    const int r0i = it.bi * kGramRows, r0j = it.bj * kGramRows;
    const int ni = min(kGramRows, m - r0i), nj = min(kGramRows, m - r0j);
    const bool same = it.bi == it.bj;
    const int pc2 = a.subset ? a.subset[it.c2] : it.c2;
    const int pc3 = a.subset ? a.subset[it.c3] : it.c3;

    // ---- staging slots of this thread ----
    const int kcol = tid % kGramK;
    std::int64_t offA[kGramSlots], offB[kGramSlots];
    double Gs[kGramSlots], Ds[kGramSlots];
    bool live[kGramSlots];
#pragma unroll
    for (int s = 0; s < kGramSlots; ++s) {
        const int slot = tid / kGramK + (kGramThreads / kGramK) * s;
        const bool lo = slot < kGramRows;
        const int p = lo ? r0i + slot : r0j + (slot - kGramRows);
        live[s] = lo ? slot < ni : (!same && slot - kGramRows < nj);
        const int pp = live[s] ? (a.subset ? a.subset[p] : p) : 0;
        const std::int64_t ra = static_cast<std::int64_t>(pp) + static_cast<std::int64_t>(P) * pc2;
        const std::int64_t rb = static_cast<std::int64_t>(pp) + static_cast<std::int64_t>(P) * pc3;
        offA[s] = ra * nbs;
        offB[s] = rb * nbs;
        const double ma = live[s] ? a.mean[ra] : 0.0, mb = live[s] ? a.mean[rb] : 0.0;
        Gs[s] = ma - mb;
        Ds[s] = 0.0;
        if (kcol == 0) {
            sG[slot] = Gs[s];
            sL[slot] = fabs(ma) + fabs(mb);
        }
    }
    double va[kGramSlots], vb[kGramSlots];
    auto fetch = [&](int c) {
        const int b = c * kGramK + kcol;
#pragma unroll
        for (int s = 0; s < kGramSlots; ++s) {
            const bool ok = live[s] && b < nbs;
            va[s] = ok ? t.f2T[offA[s] + b] : 0.0;
            vb[s] = ok ? t.f2T[offB[s] + b] : 0.0;
        }
    };
    auto stage = [&](int c) {
        const int b = c * kGramK + kcol;
        const double w = b < nbs ? a.wroot[b] : 0.0;
#pragma unroll
        for (int s = 0; s < kGramSlots; ++s) {
            const int slot = tid / kGramK + (kGramThreads / kGramK) * s;
            const double e = (live[s] && b < nbs) ? (va[s] - vb[s] - Gs[s]) * w : 0.0;
            Ds[s] += e * e;
            sE[kcol * kGramLd + slot] = e;
        }
    };

    // ---- this wavefront's tiles ----
    const int nti = (ni + 15) / 16, ntj = (nj + 15) / 16;
    const int n_tiles = same ? nti * (nti + 1) / 2 : nti * ntj;
    int colA[kGramMaxTiles], colB[kGramMaxTiles];
    int my_tiles = 0;
#pragma unroll
    for (int u = 0; u < kGramMaxTiles; ++u) {
        int tl = wave + kGramWaves * u;
        colA[u] = colB[u] = 0;
        if (tl < n_tiles) {
            int ti = 0, tj = 0;
            if (same) {
                for (ti = 0; ti < nti; ++ti) {
                    const int len = nti - ti;
                    if (tl < len) { tj = ti + tl; break; }
                    tl -= len;
                }
            } else {
                ti = tl / ntj;
                tj = tl - ti * ntj;
            }
            colA[u] = ti * 16;
            colB[u] = (same ? 0 : kGramRows) + tj * 16;
            my_tiles = u + 1;
        }
    }
    dbl4 acc[kGramMaxTiles];
#pragma unroll
    for (int u = 0; u < kGramMaxTiles; ++u) acc[u] = dbl4{0.0, 0.0, 0.0, 0.0};

    // ---- stream the blocks ----
    const int nch = (nbs + kGramK - 1) / kGramK;
    fetch(0);
    stage(0);
    __syncthreads();
    const int krow = lane >> 4, kc = lane & 15;
    for (int c = 0; c < nch; ++c) {
        if (c + 1 < nch) fetch(c + 1);
#pragma unroll
        for (int u = 0; u < kGramMaxTiles; ++u) {
            if (u < my_tiles) {
#pragma unroll
                for (int ks = 0; ks < kGramK / 4; ++ks) {
                    const double* row = sE + (4 * ks + krow) * kGramLd + kc;
                    acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(row[colA[u]], row[colB[u]], acc[u], 0, 0, 0);
                }
            }
        }
        __syncthreads();
        if (c + 1 < nch) {
            stage(c + 1);
            __syncthreads();
        }
    }

    // ---- diagonal: sum over the 16 lanes that staged each slot ----
#pragma unroll
    for (int s = 0; s < kGramSlots; ++s) {
        double d = Ds[s];
        d += __shfl_xor(d, 8);
        d += __shfl_xor(d, 4);
        d += __shfl_xor(d, 2);
        d += __shfl_xor(d, 1);
        if (kcol == 0) sD[tid / kGramK + (kGramThreads / kGramK) * s] = d;
    }
    __syncthreads();

    // ---- read-out: lane holds S(row (lane/16) + 4r, col lane%16) of each tile ----
    const double inv4nb = 1.0 / (4.0 * static_cast<double>(nbs));
#pragma unroll
    for (int u = 0; u < kGramMaxTiles; ++u) {
        if (u >= my_tiles) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int sp = colA[u] + krow + 4 * r;          // slot of c0 (block bi rows)
            const int sq = colB[u] + kc;                    // slot of c1
            const int p = r0i + sp;
            const int q = r0j + (sq - (same ? 0 : kGramRows));
            const bool valid = sp < ni && q < m && p < q;
            const double D0 = sD[sp], D1 = sD[sq];
            const double sum = D0 + D1;
            const double den = sum - 2.0 * acc[u][r];
            const double var = den * inv4nb;
            const double est = 0.5 * (sG[sq] - sG[sp]);
            const double se = var > 0.0 ? sqrt(var) : __builtin_nan("");
            // false for 0 / 0 and NaN
            const bool precise = den * a.max_cancel > sum && sL[sp] + sL[sq] <= a.max_level * (fabs(est) + se);
            const double z = est / se;
            const std::int64_t rank = it.base + static_cast<std::int64_t>(q) * (q - 1) / 2 + p;
            unsigned long long slot;
            bool ok;
            wave_append(valid && precise && fabs(z) >= a.tau, a.n_hits, a.cap, lane, slot, ok);
            if (ok) a.hits[slot] = SweepHit{rank, est, se, z};
            wave_append(valid && !precise, a.n_recheck, a.cap, lane, slot, ok);
            if (ok) a.recheck[slot] = rank;
        }
    }
}

// Smallest cached width that holds ceil(nbs / 64) values per lane; 0 = streaming.
int per_lane(int nbs) {
    const int need = (nbs + 63) / 64;
    if (need > kMaxPer) return 0;
    for (int k : {2, 4, 8, 16, kMaxPer})
        if (need <= k) return k;
    return 0;
}

template <class F>
void dispatch_per(int nbs, F&& f) {
    switch (per_lane(nbs)) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        default: f(std::integral_constant<int, 0>{}); break;
    }
}

dim3 grid_for(std::int64_t n_items) {
    const std::int64_t groups = (n_items + kWaves - 1) / kWaves;
    return dim3(static_cast<unsigned>(round_up(groups, 8)));
}

}  // namespace

void launch_fstat(int width, const FstatArgs& a, hipStream_t st) {
    if (a.N <= 0) return;
    STP_REQUIRE(width == 3 || width == 4, "fstat: width must be 3 or 4");
    STP_REQUIRE(a.t.nbs >= 1, "fstat: no surviving block");
    STP_REQUIRE((a.N + kWaves - 1) / kWaves <= 0x7fffffffLL, "fstat: too many statistics for one launch");
    dispatch_per(a.t.nbs, [&](auto per) {
        constexpr int kPer = decltype(per)::value;
        if (width == 4) fstat_kernel<4, kPer><<<grid_for(a.N), kThreads, 0, st>>>(a);
        else fstat_kernel<3, kPer><<<grid_for(a.N), kThreads, 0, st>>>(a);
    });
    STP_LAUNCH_CHECK();
}

void launch_f4ratio(const F4RatioArgs& a, hipStream_t st) {
    if (a.N <= 0) return;
    STP_REQUIRE(a.t.nbs >= 1, "f4ratio: no surviving block");
    STP_REQUIRE((a.N + kWaves - 1) / kWaves <= 0x7fffffffLL, "f4ratio: too many tuples for one launch");
    f4ratio_kernel<<<grid_for(a.N), kThreads, 0, st>>>(a);
    STP_LAUNCH_CHECK();
}

void launch_f2_means(const FstatTensor& t, double* mean, hipStream_t st) {
    const std::int64_t rows = static_cast<std::int64_t>(t.P) * t.P;
    if (rows <= 0) return;
    STP_REQUIRE(t.nbs >= 1, "f2 means: no surviving block");
    f2_means_kernel<<<grid_for(rows), kThreads, 0, st>>>(t, mean);
    STP_LAUNCH_CHECK();
}

void launch_f4_gram(const GramArgs& a, hipStream_t st) {
    if (a.n_items <= 0) return;
    STP_REQUIRE(a.t.nbs >= 1, "f4 Gram sweep: no surviving block");
    STP_REQUIRE(a.n_items <= 0x7ffffff0LL, "f4 Gram sweep: too many items for one launch");
    f4_gram_kernel<<<dim3(static_cast<unsigned>(round_up(a.n_items, 8))), kGramThreads, 0, st>>>(a);
    STP_LAUNCH_CHECK();
}

void launch_sweep_direct(const SweepArgs& a, hipStream_t st) {
    if (a.count <= 0) return;
    STP_REQUIRE(a.k == 3 || a.k == 4, "sweep: k must be 3 or 4");
    STP_REQUIRE(a.t.nbs >= 1, "sweep: no surviving block");
    STP_REQUIRE(a.cap >= a.count, "sweep: hit buffer smaller than the launch");
    STP_REQUIRE((a.count + kWaves - 1) / kWaves <= 0x7fffffffLL, "sweep: launch too large");
    dispatch_per(a.t.nbs, [&](auto per) {
        constexpr int kPer = decltype(per)::value;
        if (a.k == 4) sweep_direct_kernel<4, kPer><<<grid_for(a.count), kThreads, 0, st>>>(a);
        else sweep_direct_kernel<3, kPer><<<grid_for(a.count), kThreads, 0, st>>>(a);
    });
    STP_LAUNCH_CHECK();
}

}  // namespace stp
