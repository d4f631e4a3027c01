// oracle/cpu_ref_fstats.cpp — CPU restatement of the reference's standalone
// f-statistics over a resident f2 tensor: f4 / qpDstat (f2 path), f3, f4-ratio and
// the all-combination f3 / f4 sweeps.
//
// TEST INFRASTRUCTURE ONLY.  The parity tests, smoke() and bench.py's CPU baseline
// call it as the checker; the product library never links or calls it.
//
// Restated from (paths relative to the reference repository root):
//   run_f4 body                  src/core/qpadm/f4.cpp:38-91
//   run_f3 body                  src/core/qpadm/f3.cpp:37-95
//   run_f4ratio body             src/core/qpadm/f4ratio.cpp:34-83 (kSetmissThresh 1e-6)
//   assemble_f4_quartets         src/device/cpu/cpu_backend.cpp:2120-2167
//   assemble_f3_triples          src/device/cpu/cpu_backend.cpp:2170-2217
//   survivor blocks              src/device/cpu/cpu_backend.cpp:125-150
//   compute_loo_and_total        src/device/cpu/cpu_backend.cpp:2674-2723
//   jackknife_diag               src/device/cpu/cpu_backend.cpp:2273-2303
//   fill_point_estimates         src/core/qpadm/point_estimates.hpp:28-44, f4_two_sided_p f4.cpp:94-97
//   f4ratio_blocks_jackknife     src/device/cpu/cpu_backend.cpp:2468-2497
//   ratio_block_jackknife (tot_mode 0)  src/device/cpu/cpu_backend.cpp:2309-2378
//   sweep driver                 src/core/qpadm/fstat_sweep.cpp:31-97 (cap kFstatMaxComb = 1e8,
//                                default reservoir kFstatDefaultSweepTopK = 1e6, include/steppe/config.hpp:20-21)
//   sweep enumeration            src/device/cuda/sweep_unrank.cuh:25-45 (combinatorial number
//                                system: rank r <-> c0 < c1 < ... with r = sum_i C(c_i, i+1))
//   sweep filter / reservoir     src/device/cuda/cuda_backend_fstats_assemble.cu:49-327 and
//                                sweep_zfilter_tau_kernel (src/device/cuda/qpadm_fit_kernels.cu:549-570):
//                                |z| >= tau with tau = min_z, rising to the K-th largest |z|
//                                in top-K mode; the survivors come back ordered by |z|
//                                descending (a stable radix sort, so ties keep rank order).
//
// The sweep has no CPU implementation in the reference (the CPU backend throws); the
// restatement here builds it from the pieces above and is pinned in the tests against
// f4 / f3 over the explicitly enumerated combinations.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "cref_common.hpp"

namespace {

using cref::F4;
using cref::Vec;

double two_sided_p(double z) {
    static const double kInvSqrt2 = 1.0 / std::sqrt(2.0);
    return std::erfc(std::fabs(z) * kInvSqrt2);
}

// The per-statistic block series over surviving blocks, then LOO / total.
// kind 4: f4 quartets (p1, p2, p3, p4); kind 3: f3 triples (C, A, B).
F4 assemble(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* idx,
            int N, int kind) {
    F4 x;
    x.nl = N;
    x.nr = 1;
    if (N <= 0 || nb <= 0) return x;
    const std::vector<int> surv = cref::survivors(vpair, P, nb);
    x.nb = static_cast<int>(surv.size());
    x.bsize.resize(x.nb);
    for (int s = 0; s < x.nb; ++s) x.bsize[s] = bsz[surv[s]];
    const std::size_t m = static_cast<std::size_t>(N);
    x.x_blocks.assign(m * x.nb, 0.0);
    if (x.nb <= 0) return x;
    const std::size_t slab = static_cast<std::size_t>(P) * P;
    auto F = [&](int i, int j, int b) { return f2[i + static_cast<std::size_t>(P) * j + slab * b]; };
    for (int k = 0; k < N; ++k) {
        const std::int32_t* q = idx + static_cast<std::size_t>(kind) * k;
        for (int s = 0; s < x.nb; ++s) {
            const int b = surv[s];
            double v;
            if (kind == 4)
                v = 0.5 * (F(q[1], q[2], b) + F(q[0], q[3], b) - F(q[0], q[2], b) - F(q[1], q[3], b));
            else
                v = 0.5 * (F(q[0], q[1], b) + F(q[0], q[2], b) - F(q[1], q[2], b));
            x.x_blocks[k + m * s] = v;
        }
    }
    cref::loo_and_total(x);
    return x;
}

// Diagonal of the block-jackknife covariance (long-double sums as the reference).
Vec jackknife_diag(const F4& x) {
    const int m = x.nl * x.nr, nb = x.nb;
    Vec var(m, 0.0);
    long double n_ld = 0.0L;
    for (int b = 0; b < nb; ++b) n_ld += static_cast<long double>(x.bsize[b]);
    const double n = static_cast<double>(n_ld);
    for (int k = 0; k < m; ++k) {
        long double acc = 0.0L;
        const double est = x.x_total[k], tl = x.tot_line[k];
        for (int b = 0; b < nb; ++b) {
            const double h = n / static_cast<double>(x.bsize[b]);
            const double sq = std::sqrt(h - 1.0);
            const double loo = x.x_loo[k + static_cast<std::size_t>(m) * b];
            const double xt = (est * h - loo * (h - 1.0) - tl) / sq;
            acc += static_cast<long double>(xt) * static_cast<long double>(xt);
        }
        var[k] = static_cast<double>(acc / static_cast<long double>(nb));
    }
    return var;
}

void point_estimates(const F4& x, const Vec& var, int N, double* est, double* se, double* z, double* p) {
    for (int k = 0; k < N; ++k) {
        const double e = x.x_total[k];
        const double s = var[k] > 0.0 ? std::sqrt(var[k]) : std::nan("");
        const double zz = e / s;
        est[k] = e;
        se[k] = s;
        z[k] = zz;
        p[k] = two_sided_p(zz);
    }
}

// run_f4 / run_f3 body: assemble, diagonal jackknife, read-out.
void fstat(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* idx,
           int N, int kind, double* est, double* se, double* z, double* p) {
    if (N <= 0) return;
    const F4 x = assemble(f2, vpair, bsz, P, nb, idx, N, kind);
    if (x.nb <= 0) {
        for (int k = 0; k < N; ++k) est[k] = se[k] = z[k] = p[k] = std::nan("");
        return;
    }
    point_estimates(x, jackknife_diag(x), N, est, se, z, p);
}

// C(n, k) in 64-bit, saturating at UINT64_MAX (the reference's choose_saturating).
std::uint64_t choose_sat(std::int64_t n, int k) {
    if (k < 0 || n < k) return 0;
    std::uint64_t r = 1;
    for (int i = 1; i <= k; ++i) {
        const std::uint64_t num = static_cast<std::uint64_t>(n - k + i);
        if (r > ~0ULL / num) return ~0ULL;
        r = r * num / static_cast<std::uint64_t>(i);
    }
    return r;
}

}  // namespace

extern "C" {

int cref_f4(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* quartets,
            int N, double* est, double* se, double* z, double* p) {
    fstat(f2, vpair, bsz, P, nb, quartets, N, 4, est, se, z, p);
    return 0;
}

int cref_f3(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* triples,
            int N, double* est, double* se, double* z, double* p) {
    fstat(f2, vpair, bsz, P, nb, triples, N, 3, est, se, z, p);
    return 0;
}

// f4-ratio alpha = f4(p1,p2;p3,p4) / f4(p1,p2;p5,p4) with the ratio block jackknife
// (tot_mode 0): blocks where |den block| < 1e-6 are set missing per tuple; the per-block
// ratio is of the leave-one-out numerator and denominator.
int cref_f4ratio(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb,
                 const std::int32_t* tuples, int N, double* alpha, double* se, double* z) {
    constexpr double kSetmiss = 1e-6;
    if (N <= 0) return 0;
    std::vector<std::int32_t> flat(static_cast<std::size_t>(8) * N);
    for (int k = 0; k < N; ++k) {
        const std::int32_t* t = tuples + 5 * static_cast<std::size_t>(k);
        std::int32_t* a = flat.data() + 4 * static_cast<std::size_t>(k);
        std::int32_t* d = flat.data() + 4 * (static_cast<std::size_t>(N) + k);
        a[0] = t[0], a[1] = t[1], a[2] = t[2], a[3] = t[3];
        d[0] = t[0], d[1] = t[1], d[2] = t[4], d[3] = t[3];
    }
    const F4 X = assemble(f2, vpair, bsz, P, nb, flat.data(), 2 * N, 4);
    for (int k = 0; k < N; ++k) alpha[k] = se[k] = z[k] = std::nan("");
    const int nbs = X.nb;
    if (nbs <= 0) return 0;
    const std::size_t m = static_cast<std::size_t>(2) * N;
    for (int k = 0; k < N; ++k) {
        auto xnum = [&](int b) { return X.x_blocks[k + m * b]; };
        auto xden = [&](int b) { return X.x_blocks[N + k + m * b]; };
        auto lnum = [&](int b) { return X.x_loo[k + m * b]; };
        auto lden = [&](int b) { return X.x_loo[N + k + m * b]; };
        long double n_ld = 0.0L, totnum = 0.0L, totden = 0.0L;
        int kept = 0;
        for (int b = 0; b < nbs; ++b) {
            if (std::fabs(xden(b)) < kSetmiss) continue;
            const double bl = static_cast<double>(X.bsize[b]);
            n_ld += bl;
            ++kept;
            totnum += static_cast<long double>(xnum(b)) * static_cast<long double>(bl);
            totden += static_cast<long double>(xden(b)) * static_cast<long double>(bl);
        }
        if (kept <= 1 || n_ld <= 0.0L || totden == 0.0L) continue;
        const double n = static_cast<double>(n_ld);
        const double tot = static_cast<double>(totnum / totden);
        long double diffsum = 0.0L, wnum = 0.0L;
        for (int b = 0; b < nbs; ++b) {
            if (std::fabs(xden(b)) < kSetmiss) continue;
            const double bl = static_cast<double>(X.bsize[b]);
            const double Rb = lnum(b) / lden(b);
            diffsum += static_cast<long double>(tot) - static_cast<long double>(Rb);
            wnum += static_cast<long double>(Rb) * static_cast<long double>(bl);
        }
        const double est = static_cast<double>(diffsum / static_cast<long double>(kept)) * static_cast<double>(kept) +
                           static_cast<double>(wnum / n_ld);
        long double vacc = 0.0L;
        for (int b = 0; b < nbs; ++b) {
            if (std::fabs(xden(b)) < kSetmiss) continue;
            const double bl = static_cast<double>(X.bsize[b]);
            const double Rb = lnum(b) / lden(b);
            const double h = n / bl;
            const double tau = h * tot - (h - 1.0) * Rb;
            const double xt = (tau - est) / std::sqrt(h - 1.0);
            vacc += static_cast<long double>(xt) * static_cast<long double>(xt);
        }
        const double var = static_cast<double>(vacc / static_cast<long double>(kept));
        alpha[k] = est;
        se[k] = var > 0.0 ? std::sqrt(var) : std::nan("");
        z[k] = est / se[k];
    }
    return 0;
}

// Unrank r in the combinatorial number system: c[0] < ... < c[k-1] with
// r = sum_i C(c[i], i + 1).
void cref_unrank(std::int64_t r, int k, std::int32_t* c) {
    for (int pos = k - 1; pos >= 0; --pos) {
        const int kk = pos + 1;
        std::int64_t v = kk - 1;
        while (static_cast<std::int64_t>(choose_sat(v + 1, kk)) <= r) ++v;
        c[pos] = static_cast<std::int32_t>(v);
        r -= static_cast<std::int64_t>(choose_sat(v, kk));
    }
}

// All-combination sweep over the C(range, k) k-subsets of the population range
// (range = P, or the n_subset positions of subset[] mapped through it), k = 4 (f4 of
// (c0, c1, c2, c3)) or 3 (f3 of (c0; c1, c2)).  filter 0: keep |z| >= min_z; filter 1:
// the top_k largest |z| among those (top_k <= 0 -> 1e6).  Results are ordered by |z|
// descending then rank ascending, at most the reservoir size K (top_k in mode 1, 1e6
// in mode 0).  keys rows hold 4 ints (the fourth is 0 for k = 3).  Returns the status
// ordinal (0 ok, 5 invalid_config when C(range, k) > 1e8 and !sure).
int cref_fstat_sweep(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, int k,
                     const std::int32_t* subset, int n_subset, int filter, double min_z, std::int64_t top_k, int sure,
                     std::int64_t cap, std::int32_t* keys, double* est, double* se, double* z, double* p,
                     std::int64_t* n_out, std::uint64_t* enumerated, int* capped) {
    constexpr std::uint64_t kMaxComb = 100000000ULL;
    constexpr std::int64_t kDefaultK = 1000000;
    *n_out = 0;
    *capped = 0;
    const int range = n_subset > 0 ? n_subset : P;
    const std::uint64_t total = choose_sat(range, k);
    *enumerated = total;
    if (range < k || total == 0) return 0;
    if (total > kMaxComb && !sure) {
        *capped = 1;
        return 5;
    }
    std::int64_t K = (filter == 1 && top_k > 0) ? top_k : kDefaultK;
    K = std::min<std::int64_t>(K, static_cast<std::int64_t>(total));
    const std::int64_t C = static_cast<std::int64_t>(total);
    std::vector<std::int32_t> items(static_cast<std::size_t>(k) * C);
    for (std::int64_t r = 0; r < C; ++r) {
        std::int32_t c[4];
        cref_unrank(r, k, c);
        for (int i = 0; i < k; ++i) items[k * r + i] = subset && n_subset > 0 ? subset[c[i]] : c[i];
    }
    Vec e(C), s(C), zz(C), pp(C);
    fstat(f2, vpair, bsz, P, nb, items.data(), static_cast<int>(C), k, e.data(), s.data(), zz.data(), pp.data());
    std::vector<std::int64_t> sel;
    for (std::int64_t r = 0; r < C; ++r)
        if (std::fabs(zz[r]) >= min_z) sel.push_back(r);
    std::stable_sort(sel.begin(), sel.end(),
                     [&](std::int64_t a, std::int64_t b) { return std::fabs(zz[a]) > std::fabs(zz[b]); });
    if (static_cast<std::int64_t>(sel.size()) > K) sel.resize(K);
    const std::int64_t n = std::min<std::int64_t>(static_cast<std::int64_t>(sel.size()), cap);
    for (std::int64_t i = 0; i < n; ++i) {
        const std::int64_t r = sel[i];
        for (int j = 0; j < 4; ++j) keys[4 * i + j] = j < k ? items[k * r + j] : 0;
        est[i] = e[r];
        se[i] = s[r];
        z[i] = zz[r];
        p[i] = two_sided_p(zz[r]);
    }
    *n_out = static_cast<std::int64_t>(sel.size());
    return 0;
}

}  // extern "C"
