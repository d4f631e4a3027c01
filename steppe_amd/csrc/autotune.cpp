// steppe_amd/csrc/autotune.cpp — launch configurations of the f2 contraction, chosen
// by timing them on the workload.
//
// The reference sizes its f2 launches with fixed rules: power-of-two size buckets of
// blocks (src/device/cuda/cuda_backend_f2_blocks.cu:55-74) and a VRAM chunk budget
// (src/device/vram_budget.hpp:63).  Here the contraction's launch shape is a small
// configuration (F2Tune) that the plan carries: the work-item schedule over the dies,
// the FP64 kernel's wavefront tile (64 x 64 or 32 x 64 outputs), whether it prefetches
// the next K-step, the FP64 and pair-count grids.  The split kernel's tile is fixed by
// its LDS ring (one 128 x 128 workgroup per CU); its schedule, the count grid and the
// placement of its global-to-LDS loads (a burst after each phase barrier, or one at a
// time between the phase's MFMAs) are tuned.  f2_autotune times every candidate on the caller's device inputs (every
// candidate writes bit-identical tensors: the same sums per entry, in the same order),
// keeps the fastest per key — device name, precision, P, and the block-size histogram
// (blocks per power-of-two padded length) — and saves the table to a text file
// (STEPPE_TUNE_CACHE, else $XDG_CACHE_HOME or ~/.cache, steppe_amd/f2_tune.tsv) that
// later processes read, so plans built for the same workload use the winner.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "engine.hpp"
#include "steppe_hip.h"

namespace stp {

std::string F2Tune::str() const {
    std::ostringstream o;
    o << "sched=" << sched << ",fp64_rows=" << fp64_rows << ",fp64_prefetch=" << fp64_prefetch
      << ",fp64_wg=" << fp64_wg_per_cu << ",count_wg=" << count_wg_per_cu << ",split_ring=" << split_ring;
    return o.str();
}

bool F2Tune::parse(const std::string& s, F2Tune& out) {
    F2Tune t;
    int found = 0;
    std::stringstream ss(s);
    std::string kv;
    while (std::getline(ss, kv, ',')) {
        const auto eq = kv.find('=');
        if (eq == std::string::npos) return false;
        const std::string k = kv.substr(0, eq);
        const int v = std::atoi(kv.c_str() + eq + 1);
        if (k == "sched") t.sched = v;
        else if (k == "fp64_rows") t.fp64_rows = v;
        else if (k == "fp64_prefetch") t.fp64_prefetch = v;
        else if (k == "fp64_wg") t.fp64_wg_per_cu = v;
        else if (k == "count_wg") t.count_wg_per_cu = v;
        else if (k == "split_ring") t.split_ring = v;
        else return false;
        ++found;
    }
    if (found != 6 || (t.split_ring != 0 && t.split_ring != 1) || (t.sched != 0 && t.sched != 1) || (t.fp64_rows != 64 && t.fp64_rows != 32) ||
        t.fp64_wg_per_cu < 1 || t.fp64_wg_per_cu > 8 || t.count_wg_per_cu < 1 || t.count_wg_per_cu > 8)
        return false;
    out = t;
    return true;
}

namespace {

std::string cache_path() {
    if (const char* p = std::getenv("STEPPE_TUNE_CACHE")) return p;
    std::string base;
    if (const char* x = std::getenv("XDG_CACHE_HOME")) base = x;
    else if (const char* h = std::getenv("HOME")) base = std::string(h) + "/.cache";
    else return "";
    return base + "/steppe_amd/f2_tune.tsv";
}

// Candidates for one precision: the defaults first, then every other combination.
std::vector<F2Tune> candidates(int precision) {
    std::vector<F2Tune> c;
    if (precision == STP_PREC_SPLIT_BF16) {
        for (int ring : {0, 1})
            for (int sched : {0, 1})
                for (int cw : {2, 1, 4}) {
                    F2Tune t;
                    t.sched = sched;
                    t.count_wg_per_cu = cw;
                    t.split_ring = ring;
                    c.push_back(t);
                }
    } else {
        for (int sched : {0, 1})
            for (int rows : {64, 32})
                for (int pf : {1, 0})
                    for (int wg : {2, 1, 3}) {
                        if (rows == 64 && wg == 3) continue;   // two wavefronts per SIMD at most
                        F2Tune t;
                        t.sched = sched;
                        t.fp64_rows = rows;
                        t.fp64_prefetch = pf;
                        t.fp64_wg_per_cu = wg;
                        c.push_back(t);
                    }
    }
    return c;
}

}  // namespace

std::string HipBackend::tune_key(const F2Plan& plan, int precision) const {
    // blocks per power-of-two padded length (in 64-SNP columns)
    std::vector<int> hist;
    for (int b = 0; b < plan.n_block; ++b) {
        int k = 0;
        for (std::int64_t c = plan.kpad[b] / kKPad; c > 1; c >>= 1) ++k;
        if (static_cast<int>(hist.size()) <= k) hist.resize(k + 1, 0);
        ++hist[k];
    }
    std::ostringstream o;
    o << name_ << '|' << (precision == STP_PREC_FP64 ? "fp64" : "split_bf16") << "|P=" << plan.P << "|h=";
    for (std::size_t i = 0; i < hist.size(); ++i) o << (i ? ":" : "") << hist[i];
    return o.str();
}

void HipBackend::tune_load() {
    tune_loaded_ = true;
    const std::string path = cache_path();
    if (path.empty()) return;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const auto tab = line.find('\t');
        if (tab == std::string::npos) continue;
        F2Tune t;
        if (F2Tune::parse(line.substr(tab + 1), t)) tuned_[line.substr(0, tab)] = t;
    }
}

void HipBackend::tune_save() const {
    const std::string path = cache_path();
    if (path.empty()) return;
    // create the directory chain, then replace the file
    for (std::size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') (void)::mkdir(path.substr(0, i).c_str(), 0755);
    const std::string tmp = path + ".tmp";
    {
        std::ofstream f(tmp, std::ios::trunc);
        if (!f) return;
        for (const auto& [k, t] : tuned_) f << k << '\t' << t.str() << '\n';
        if (!f) return;
    }
    (void)std::rename(tmp.c_str(), path.c_str());
}

F2Tune HipBackend::tune_for(const F2Plan& plan, int precision) {
    // STEPPE_F2_TUNE forces one configuration (tests and diagnostics)
    if (const char* env = std::getenv("STEPPE_F2_TUNE")) {
        F2Tune t;
        if (F2Tune::parse(env, t)) return t;
    }
    if (!tune_loaded_) tune_load();
    const auto it = tuned_.find(tune_key(plan, precision));
    return it == tuned_.end() ? F2Tune{} : it->second;
}

std::string HipBackend::f2_autotune(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                                    std::int64_t M, const std::int32_t* block_id, int n_block, int precision,
                                    double* f2, double* vpair) {
    activate();
    STP_REQUIRE(precision == STP_PREC_FP64 || precision == STP_PREC_SPLIT_BF16,
                "autotune: unknown precision form " + std::to_string(precision));
    STP_REQUIRE(a2 && n && f2 && vpair, "autotune: null device buffer");
    STP_REQUIRE(ld >= M, "autotune: ld must be >= M");
    if (!tune_loaded_) tune_load();
    F2Source src;
    src.a2 = a2;
    src.n = n;
    src.ld = ld;
    hipEvent_t e0, e1;
    STP_HIP(hipEventCreate(&e0));
    STP_HIP(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    const bool saved_timing = timing_;
    timing_ = false;
    const std::vector<F2Tune> cands = candidates(precision);
    std::vector<double> ms(cands.size(), 0.0);
    std::string key;
    for (std::size_t i = 0; i < cands.size(); ++i) {
        const F2Plan plan = make_f2_plan(P, M, block_id, n_block, cands[i]);
        if (key.empty()) key = tune_key(plan, precision);
        upload_plan(plan, work_.plan);
        plan_valid_ = false;   // the backend's cached plan no longer matches work_.plan
        run_f2(plan, precision, src, f2, vpair, work_);   // warm-up
        sync();
        double best = 1e300;
        for (int r = 0; r < 2; ++r) {
            STP_HIP(hipEventRecord(e0, stream_));
            run_f2(plan, precision, src, f2, vpair, work_);
            STP_HIP(hipEventRecord(e1, stream_));
            STP_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            STP_HIP(hipEventElapsedTime(&t, e0, e1));
            best = std::min(best, static_cast<double>(t));
        }
        ms[i] = best;
    }
    timing_ = saved_timing;
    const std::size_t win = static_cast<std::size_t>(std::min_element(ms.begin(), ms.end()) - ms.begin());
    tuned_[key] = cands[win];
    tune_save();
    // leave the winner's result in f2 / vpair
    const F2Plan plan = make_f2_plan(P, M, block_id, n_block, cands[win]);
    upload_plan(plan, work_.plan);
    run_f2(plan, precision, src, f2, vpair, work_);
    sync();
    std::ostringstream o;
    o << "{\"key\": \"" << key << "\", \"chosen\": " << win << ", \"cache\": \"" << cache_path()
      << "\", \"candidates\": [";
    for (std::size_t i = 0; i < cands.size(); ++i)
        o << (i ? ", " : "") << "{\"config\": \"" << cands[i].str() << "\", \"ms\": " << ms[i] << "}";
    o << "]}";
    return o.str();
}

}  // namespace stp
