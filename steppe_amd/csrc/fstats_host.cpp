// steppe_amd/csrc/fstats_host.cpp — host side of the standalone f-statistics on a
// resident f2 tensor (HipBackend::fstat / f4ratio / fstat_sweep).
//
// Mirrors the reference's drivers: run_f4 / run_f3 (src/core/qpadm/f4.cpp:38-91,
// f3.cpp:37-95: assemble over surviving blocks, diagonal jackknife, est/se/z/p),
// run_f4ratio (src/core/qpadm/f4ratio.cpp:34-83) and the sweep driver
// (src/core/qpadm/fstat_sweep.cpp:31-97 with the device pipeline of
// src/device/cuda/cuda_backend_fstats_assemble.cu:49-327): the C(range, k) cap
// (1e8 unless sure), the reservoir of size K (top_k in top-K mode, 1e6 otherwise),
// the |z| >= tau filter with tau rising to the K-th largest |z| in top-K mode, and
// survivors ordered by |z| descending with ties in combination-rank order.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#include "engine.hpp"
#include "fstats.hpp"
#include "qpadm.hpp"

namespace stp {

namespace {

constexpr std::uint64_t kMaxComb = 100000000ULL;   // include/steppe/config.hpp:20
constexpr std::int64_t kDefaultTopK = 1000000;      // include/steppe/config.hpp:21
constexpr std::int64_t kDirectChunk = 1 << 22;      // combinations per direct launch
constexpr std::int64_t kGramLaunch = 1 << 24;       // f4 statistics per Gram launch (about)

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

void unrank_host(std::int64_t r, int k, const std::vector<std::int64_t>& binom, int range, int* c) {
    int hi = range - 1;
    for (int pos = k - 1; pos >= 0; --pos) {
        const int kk = pos + 1;
        const std::int64_t* C = binom.data() + static_cast<std::size_t>(kk) * (range + 1);
        int lo = kk - 1, h = hi;
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

// Survivor order: |z| descending, then rank ascending (NaN |z| never reaches here).
bool hit_before(const SweepHit& a, const SweepHit& b) {
    const double za = std::fabs(a.z), zb = std::fabs(b.z);
    if (za != zb) return za > zb;
    return a.rank < b.rank;
}

// Keep the K best hits; in top-K mode raise tau to the K-th |z| (the reference's
// compact_and_raise, cuda_backend_fstats_assemble.cu:181-209).
void keep_top(std::vector<SweepHit>& res, std::int64_t K, int filter, double& tau) {
    if (static_cast<std::int64_t>(res.size()) <= K) return;
    std::nth_element(res.begin(), res.begin() + (K - 1), res.end(), hit_before);
    const double kth = std::fabs(res[K - 1].z);
    res.resize(K);
    if (filter == 1 && kth > tau) tau = kth;
}

}  // namespace

void HipBackend::survivor_list(const F2Blocks& h, std::vector<std::int32_t>& surv, std::vector<std::int32_t>& bsz,
                               double& n_total) {
    STP_REQUIRE(h.tier == kTierResident,
                "f-statistics: the f2 tensor is held in host memory or on disk; fits need it resident in HBM");
    const int P = h.P, nb = h.n_block;
    surv.clear();
    bsz.clear();
    n_total = 0.0;
    if (nb <= 0) return;
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;
    qp_keep_.ensure(static_cast<std::size_t>(nb));
    launch_survivors(h.vpair.get(), slab, nb, qp_keep_.get(), stream_);
    std::vector<std::int32_t> keep(nb);
    STP_HIP(hipMemcpyAsync(keep.data(), qp_keep_.get(), nb * sizeof(std::int32_t), hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipStreamSynchronize(stream_));
    for (int b = 0; b < nb; ++b)
        if (keep[b]) {
            surv.push_back(b);
            bsz.push_back(h.block_sizes[b]);
            n_total += static_cast<double>(h.block_sizes[b]);
        }
}

FstatTensor HipBackend::survivor_tensor(const F2Blocks& h, std::vector<std::int32_t>* bsz_out) {
    std::vector<std::int32_t> surv, bsz;
    double n_total = 0.0;
    survivor_list(h, surv, bsz, n_total);
    if (bsz_out) *bsz_out = bsz;
    FstatTensor t;
    t.P = h.P;
    t.nbs = static_cast<int>(surv.size());
    t.n_total = n_total;
    if (t.nbs == 0) return t;
    for (int s : bsz) STP_REQUIRE(s > 0, "fstat: surviving blocks must have positive sizes");
    std::vector<std::int32_t> arena(surv);
    arena.insert(arena.end(), bsz.begin(), bsz.end());
    fs_i32_.ensure(arena.size());
    STP_HIP(hipMemcpyAsync(fs_i32_.get(), arena.data(), arena.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                           stream_));
    const std::int64_t slab = static_cast<std::int64_t>(h.P) * h.P;
    qpT_.ensure(static_cast<std::size_t>(slab) * t.nbs);
    launch_transpose_survivors(h.f2.get(), slab, fs_i32_.get(), t.nbs, qpT_.get(), stream_);
    // The arena is pageable host memory that goes out of scope here.
    STP_HIP(hipStreamSynchronize(stream_));
    t.f2T = qpT_.get();
    t.bsz = fs_i32_.get() + t.nbs;
    return t;
}

void HipBackend::fstat(const F2Blocks& h, int width, std::int64_t n, const std::int32_t* idx, double* est, double* se,
                       double* z, double* p) {
    activate();
    STP_REQUIRE(width == 3 || width == 4, "fstat: width must be 3 or 4");
    STP_REQUIRE(n >= 0, "fstat: negative count");
    if (n == 0) return;
    STP_REQUIRE(idx && est && se && z && p, "fstat: null argument");
    for (std::int64_t k = 0; k < n * width; ++k)
        STP_REQUIRE(idx[k] >= 0 && idx[k] < h.P,
                    "fstat: population index out of range in tuple " + std::to_string(k / width));
    const FstatTensor t = survivor_tensor(h);
    if (t.nbs == 0) {
        std::fill(est, est + n, std::nan(""));
        std::fill(se, se + n, std::nan(""));
        std::fill(z, z + n, std::nan(""));
        std::fill(p, p + n, std::nan(""));
        return;
    }
    fs_idx_.ensure(static_cast<std::size_t>(n) * width);
    STP_HIP(hipMemcpyAsync(fs_idx_.get(), idx, static_cast<std::size_t>(n) * width * sizeof(std::int32_t),
                           hipMemcpyHostToDevice, stream_));
    fs_out_.ensure(static_cast<std::size_t>(4) * n);
    FstatArgs a;
    a.t = t;
    a.idx = fs_idx_.get();
    a.N = n;
    a.est = fs_out_.get();
    a.se = a.est + n;
    a.z = a.se + n;
    a.p = a.z + n;
    mark_begin(kFstat);
    launch_fstat(width, a, stream_);
    mark_end(kFstat);
    std::vector<double> host(static_cast<std::size_t>(4) * n);
    STP_HIP(hipMemcpyAsync(host.data(), fs_out_.get(), host.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    std::copy(host.begin(), host.begin() + n, est);
    std::copy(host.begin() + n, host.begin() + 2 * n, se);
    std::copy(host.begin() + 2 * n, host.begin() + 3 * n, z);
    std::copy(host.begin() + 3 * n, host.end(), p);
}

void HipBackend::f4ratio(const F2Blocks& h, std::int64_t n, const std::int32_t* tuples, double* alpha, double* se,
                         double* z) {
    activate();
    STP_REQUIRE(n >= 0, "f4ratio: negative count");
    if (n == 0) return;
    STP_REQUIRE(tuples && alpha && se && z, "f4ratio: null argument");
    for (std::int64_t k = 0; k < n * 5; ++k)
        STP_REQUIRE(tuples[k] >= 0 && tuples[k] < h.P,
                    "f4ratio: population index out of range in tuple " + std::to_string(k / 5));
    const FstatTensor t = survivor_tensor(h);
    if (t.nbs == 0) {
        std::fill(alpha, alpha + n, std::nan(""));
        std::fill(se, se + n, std::nan(""));
        std::fill(z, z + n, std::nan(""));
        return;
    }
    fs_idx_.ensure(static_cast<std::size_t>(n) * 5);
    STP_HIP(hipMemcpyAsync(fs_idx_.get(), tuples, static_cast<std::size_t>(n) * 5 * sizeof(std::int32_t),
                           hipMemcpyHostToDevice, stream_));
    fs_out_.ensure(static_cast<std::size_t>(3) * n);
    F4RatioArgs a;
    a.t = t;
    a.tuples = fs_idx_.get();
    a.N = n;
    a.alpha = fs_out_.get();
    a.se = a.alpha + n;
    a.z = a.se + n;
    mark_begin(kFstat);
    launch_f4ratio(a, stream_);
    mark_end(kFstat);
    std::vector<double> host(static_cast<std::size_t>(3) * n);
    STP_HIP(hipMemcpyAsync(host.data(), fs_out_.get(), host.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    std::copy(host.begin(), host.begin() + n, alpha);
    std::copy(host.begin() + n, host.begin() + 2 * n, se);
    std::copy(host.begin() + 2 * n, host.end(), z);
}

// The Gram form: items are 128 x 128 row blocks of S for each (c2, c3), grouped into
// launches of about kGramLaunch statistics (all items of a pair in one launch).
// Statistics whose variance would be formed with too much cancellation come back as
// ranks and are recomputed by the direct kernel in the same launch's hit buffer.
void HipBackend::f4_sweep_gram(const FstatTensor& t, const std::vector<std::int32_t>& bsz, int range,
                               const std::int32_t* d_subset, const std::vector<std::int64_t>& binom, std::int64_t K,
                               int filter, double& tau, std::vector<SweepHit>& res, std::int64_t& rechecked) {
    const std::int64_t rows = static_cast<std::int64_t>(t.P) * t.P;
    fs_mean_.ensure(static_cast<std::size_t>(rows));
    launch_f2_means(t, fs_mean_.get(), stream_);
    std::vector<double> wroot(t.nbs);
    for (int b = 0; b < t.nbs; ++b) {
        const double w = static_cast<double>(bsz[b]);
        wroot[b] = std::sqrt(w / (t.n_total - w));
    }
    fs_wroot_.ensure(wroot.size());
    STP_HIP(hipMemcpyAsync(fs_wroot_.get(), wroot.data(), wroot.size() * sizeof(double), hipMemcpyHostToDevice,
                           stream_));
    auto C = [&](int v, int kk) { return binom[static_cast<std::size_t>(kk) * (range + 1) + v]; };
    // launches: [item range, statistic count]
    std::vector<GramItem> items;
    std::vector<std::pair<std::size_t, std::int64_t>> cuts;   // (end item, statistics)
    std::int64_t acc = 0, cap = 0;
    for (int c3 = 3; c3 < range; ++c3)
        for (int c2 = 2; c2 < c3; ++c2) {
            const int nbk = (c2 + kGramRows - 1) / kGramRows;
            const std::int64_t base = C(c2, 3) + C(c3, 4);
            for (int bi = 0; bi < nbk; ++bi)
                for (int bj = bi; bj < nbk; ++bj) items.push_back(GramItem{c2, c3, bi, bj, base});
            acc += C(c2, 2);
            if (acc >= kGramLaunch) {
                cuts.emplace_back(items.size(), acc);
                cap = std::max(cap, acc);
                acc = 0;
            }
        }
    if (acc > 0) {
        cuts.emplace_back(items.size(), acc);
        cap = std::max(cap, acc);
    }
    fs_items_.ensure(items.size());
    STP_HIP(hipMemcpyAsync(fs_items_.get(), items.data(), items.size() * sizeof(GramItem), hipMemcpyHostToDevice,
                           stream_));
    fs_hits_.ensure(static_cast<std::size_t>(cap));
    fs_recheck_.ensure(static_cast<std::size_t>(cap));
    fs_count_.ensure(2);
    std::vector<SweepHit> buf;
    std::size_t i0 = 0;
    for (const auto& [i1, nstat] : cuts) {
        STP_HIP(hipMemsetAsync(fs_count_.get(), 0, 2 * sizeof(unsigned long long), stream_));
        GramArgs g;
        g.t = t;
        g.mean = fs_mean_.get();
        g.wroot = fs_wroot_.get();
        g.subset = d_subset;
        g.items = fs_items_.get() + i0;
        g.n_items = static_cast<std::int64_t>(i1 - i0);
        g.tau = tau;
        g.hits = fs_hits_.get();
        g.n_hits = fs_count_.get();
        g.recheck = fs_recheck_.get();
        g.n_recheck = fs_count_.get() + 1;
        g.cap = cap;
        launch_f4_gram(g, stream_);
        unsigned long long cnt[2] = {0, 0};
        STP_HIP(hipMemcpyAsync(cnt, fs_count_.get(), sizeof(cnt), hipMemcpyDeviceToHost, stream_));
        STP_HIP(hipStreamSynchronize(stream_));
        STP_REQUIRE(cnt[0] <= static_cast<unsigned long long>(nstat) &&
                        cnt[1] <= static_cast<unsigned long long>(nstat),
                    "fstat_sweep: Gram launch counts exceed its statistics");
        if (cnt[1] > 0) {
            SweepArgs d;
            d.t = t;
            d.k = 4;
            d.range = range;
            d.subset = d_subset;
            d.binom = fs_i64_.get();
            d.ranks = fs_recheck_.get();
            d.count = static_cast<std::int64_t>(cnt[1]);
            d.tau = tau;
            d.hits = fs_hits_.get();
            d.n_hits = fs_count_.get();
            d.cap = cap;
            launch_sweep_direct(d, stream_);
            STP_HIP(hipMemcpyAsync(cnt, fs_count_.get(), sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                   stream_));
            STP_HIP(hipStreamSynchronize(stream_));
            STP_REQUIRE(cnt[0] <= static_cast<unsigned long long>(nstat), "fstat_sweep: hit count exceeds the launch");
            rechecked += static_cast<std::int64_t>(cnt[1]);
        }
        if (cnt[0] > 0) {
            buf.resize(cnt[0]);
            STP_HIP(hipMemcpyAsync(buf.data(), fs_hits_.get(), cnt[0] * sizeof(SweepHit), hipMemcpyDeviceToHost,
                                   stream_));
            STP_HIP(hipStreamSynchronize(stream_));
            res.insert(res.end(), buf.begin(), buf.end());
            keep_top(res, K, filter, tau);
        }
        i0 = i1;
    }
}

void HipBackend::fstat_sweep(const F2Blocks& h, const FstatSweepRequest& req, FstatSweepOut& out) {
    activate();
    out = FstatSweepOut{};
    const int k = req.k;
    STP_REQUIRE(k == 3 || k == 4, "fstat_sweep: k must be 3 or 4");
    STP_REQUIRE(req.filter == 0 || req.filter == 1, "fstat_sweep: filter must be 0 (min_z) or 1 (top_k)");
    STP_REQUIRE(req.method >= 0 && req.method <= 2, "fstat_sweep: method must be 0, 1 or 2");
    STP_REQUIRE(!(req.method == 2 && k != 4), "fstat_sweep: the Gram method applies to f4 sweeps only");
    for (std::size_t i = 0; i < req.subset.size(); ++i)
        STP_REQUIRE(req.subset[i] >= 0 && req.subset[i] < h.P, "fstat_sweep: subset index out of range");
    const int range = req.subset.empty() ? h.P : static_cast<int>(req.subset.size());
    const std::uint64_t total = choose_sat(range, k);
    out.enumerated = total;
    if (range < k || total == 0) return;
    if (total > kMaxComb && !req.sure) {
        out.capped = true;
        out.status = 5;
        return;
    }
    STP_REQUIRE(total < (1ULL << 62), "fstat_sweep: too many combinations to enumerate");
    std::vector<std::int32_t> host_bsz;
    const FstatTensor t = survivor_tensor(h, &host_bsz);
    if (t.nbs == 0) return;
    std::int64_t K = (req.filter == 1 && req.top_k > 0) ? req.top_k : kDefaultTopK;
    K = std::min<std::int64_t>(K, static_cast<std::int64_t>(total));

    // C(v, kk) for v <= range, kk <= 4, and the subset.
    std::vector<std::int64_t> binom(static_cast<std::size_t>(5) * (range + 1), 0);
    for (int kk = 0; kk <= 4; ++kk)
        for (int v = 0; v <= range; ++v) binom[kk * (range + 1) + v] = static_cast<std::int64_t>(choose_sat(v, kk));
    fs_i64_.ensure(binom.size());
    STP_HIP(hipMemcpyAsync(fs_i64_.get(), binom.data(), binom.size() * sizeof(std::int64_t), hipMemcpyHostToDevice,
                           stream_));
    fs_idx_.ensure(std::max<std::size_t>(req.subset.size(), 1));
    if (!req.subset.empty())
        STP_HIP(hipMemcpyAsync(fs_idx_.get(), req.subset.data(), req.subset.size() * sizeof(std::int32_t),
                               hipMemcpyHostToDevice, stream_));

    double tau = req.min_z;
    std::vector<SweepHit> res;
    const bool gram = k == 4 && t.nbs >= 2 && (req.method == 2 || (req.method == 0 && range >= 24));
    const std::int32_t* d_subset = req.subset.empty() ? nullptr : fs_idx_.get();
    mark_begin(kFstat);
    if (gram) {
        f4_sweep_gram(t, host_bsz, range, d_subset, binom, K, req.filter, tau, res, out.rechecked);
    } else {
        // Combinations per launch; STEPPE_FSTAT_CHUNK overrides it, as it does the
        // reference's sweep chunk (cuda_backend_fstats_assemble.cu:128-131).
        std::int64_t chunk = kDirectChunk;
        if (const char* env = std::getenv("STEPPE_FSTAT_CHUNK")) {
            const long long v = std::strtoll(env, nullptr, 10);
            if (v > 0) chunk = v;
        }
        chunk = std::min<std::int64_t>(static_cast<std::int64_t>(total), chunk);
        fs_hits_.ensure(static_cast<std::size_t>(chunk));
        fs_count_.ensure(1);
        SweepArgs a;
        a.t = t;
        a.k = k;
        a.range = range;
        a.subset = d_subset;
        a.binom = fs_i64_.get();
        a.hits = fs_hits_.get();
        a.n_hits = fs_count_.get();
        a.cap = chunk;
        std::vector<SweepHit> buf;
        for (std::int64_t r0 = 0; r0 < static_cast<std::int64_t>(total); r0 += chunk) {
            a.r0 = r0;
            a.count = std::min<std::int64_t>(chunk, static_cast<std::int64_t>(total) - r0);
            a.tau = tau;
            STP_HIP(hipMemsetAsync(fs_count_.get(), 0, sizeof(unsigned long long), stream_));
            launch_sweep_direct(a, stream_);
            unsigned long long nh = 0;
            STP_HIP(hipMemcpyAsync(&nh, fs_count_.get(), sizeof(nh), hipMemcpyDeviceToHost, stream_));
            STP_HIP(hipStreamSynchronize(stream_));
            STP_REQUIRE(nh <= static_cast<unsigned long long>(a.count), "fstat_sweep: hit count exceeds the launch");
            if (nh == 0) continue;
            buf.resize(nh);
            STP_HIP(hipMemcpyAsync(buf.data(), fs_hits_.get(), nh * sizeof(SweepHit), hipMemcpyDeviceToHost,
                                   stream_));
            STP_HIP(hipStreamSynchronize(stream_));
            res.insert(res.end(), buf.begin(), buf.end());
            keep_top(res, K, req.filter, tau);
        }
    }
    mark_end(kFstat);
    sync();
    std::sort(res.begin(), res.end(), hit_before);
    if (static_cast<std::int64_t>(res.size()) > K) res.resize(K);

    const std::size_t ns = res.size();
    out.keys.assign(4 * ns, 0);
    out.est.resize(ns);
    out.se.resize(ns);
    out.z.resize(ns);
    out.p.resize(ns);
    out.rank.resize(ns);
    static const double kInvSqrt2 = 1.0 / std::sqrt(2.0);
    for (std::size_t i = 0; i < ns; ++i) {
        int c[4] = {0, 0, 0, 0};
        unrank_host(res[i].rank, k, binom, range, c);
        for (int j = 0; j < k; ++j) out.keys[4 * i + j] = req.subset.empty() ? c[j] : req.subset[c[j]];
        out.est[i] = res[i].est;
        out.se[i] = res[i].se;
        out.z[i] = res[i].z;
        out.p[i] = std::erfc(std::fabs(res[i].z) * kInvSqrt2);
        out.rank[i] = res[i].rank;
    }
}

}  // namespace stp
