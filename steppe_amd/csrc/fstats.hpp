// steppe_amd/csrc/fstats.hpp — launch interface of the standalone f-statistic kernels
// (f4 / f3 per statistic, f4-ratio, and the all-combination f3 / f4 sweeps) over the
// survivor-transposed resident f2 tensor f2T[(i + P*j) * nbs + s].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace stp {

// The resident tensor as every f-statistic kernel reads it.
struct FstatTensor {
    const double* f2T = nullptr;    // [(i + P*j) * nbs + s], surviving blocks only
    const std::int32_t* bsz = nullptr;  // survivor block sizes [nbs]
    int P = 0, nbs = 0;
    double n_total = 0.0;           // sum of the survivor block sizes
};

// One statistic per wavefront: width 4 (f4 of (p1, p2, p3, p4)) or 3 (f3 of (C; A, B)).
struct FstatArgs {
    FstatTensor t;
    const std::int32_t* idx = nullptr;  // N * width population indices
    std::int64_t N = 0;
    double *est = nullptr, *se = nullptr, *z = nullptr, *p = nullptr;
};
void launch_fstat(int width, const FstatArgs& a, hipStream_t st);

// One (p1, p2, p3, p4, p5) tuple per wavefront: alpha = f4(p1,p2;p3,p4) / f4(p1,p2;p5,p4).
struct F4RatioArgs {
    FstatTensor t;
    const std::int32_t* tuples = nullptr;  // N * 5
    std::int64_t N = 0;
    double *alpha = nullptr, *se = nullptr, *z = nullptr;
};
void launch_f4ratio(const F4RatioArgs& a, hipStream_t st);

// A sweep candidate: combination rank (combinatorial number system) and its estimate.
struct SweepHit {
    std::int64_t rank;
    double est, se, z;
};

// Direct sweep: statistics of the combinations with ranks r0 .. r0+count-1 (or of the
// listed ranks), one per wavefront; those with |z| >= tau are appended to hits.
struct SweepArgs {
    FstatTensor t;
    int k = 4, range = 0;
    const std::int32_t* subset = nullptr;   // position -> population (null: identity)
    const std::int64_t* binom = nullptr;    // C(v, kk) at [kk * (range + 1) + v], kk <= 4
    std::int64_t r0 = 0, count = 0;
    const std::int64_t* ranks = nullptr;    // explicit ranks (count of them) instead of r0..
    double tau = 0.0;
    SweepHit* hits = nullptr;
    unsigned long long* n_hits = nullptr;
    std::int64_t cap = 0;                   // hits capacity; the host sizes launches to fit
};
void launch_sweep_direct(const SweepArgs& a, hipStream_t st);

// Weighted block means of every f2 series: mean[i + P*j] = sum_s bsz[s] f2T[(i+P*j)*nbs+s] / n.
void launch_f2_means(const FstatTensor& t, double* mean, hipStream_t st);

// Gram form of the f4 sweep.  For positions c2 < c3 every f4 of (c0 < c1 < c2, c2, c3) is
// 0.5 (g(c1) - g(c0)) per block with g(p) = f2(p,c2) - f2(p,c3), so its jackknife variance
// is (D(c0) + D(c1) - 2 S(c0,c1)) / (4 nbs) with S the Gram matrix of the centred,
// weighted series e_b(p) = (g_b(p) - G(p)) sqrt(w_b / (n - w_b)) and D its diagonal.
// One item is one 128 x 128 block (bi <= bj) of S for one (c2, c3).
constexpr int kGramRows = 128;       // rows of S per item block
struct GramItem {
    int c2, c3, bi, bj;
    std::int64_t base;               // C(c2, 3) + C(c3, 4)
};
struct GramArgs {
    FstatTensor t;
    const double* mean = nullptr;    // launch_f2_means output
    const double* wroot = nullptr;   // sqrt(w_b / (n - w_b)) per surviving block
    const std::int32_t* subset = nullptr;
    const GramItem* items = nullptr;
    std::int64_t n_items = 0;
    double tau = 0.0;
    double max_cancel = 1e4;         // (D0 + D1) / (D0 + D1 - 2 S01) above this: recomputed directly
    double max_level = 1e6;          // the four f2 means above this times (|est| + se): recomputed
    SweepHit* hits = nullptr;
    unsigned long long* n_hits = nullptr;
    std::int64_t* recheck = nullptr;
    unsigned long long* n_recheck = nullptr;
    std::int64_t cap = 0;            // capacity of hits and of recheck
};
void launch_f4_gram(const GramArgs& a, hipStream_t st);

}  // namespace stp
