// steppe_amd/csrc/qpadm_host.cpp — host side of the batched qpAdm fit
// (HipBackend::qpadm_batch).
//
// Mirrors the reference's batched entry (ComputeBackend::fit_models_batched,
// src/device/cuda/cuda_backend_qpadm_fit.cu:623-960 and run_qpadm_search,
// src/core/qpadm/model_search.cpp:30-218): survivor blocks are found once for the
// tensor, models are validated and grouped by the size of their largest normal
// matrix, and each group runs as one persistent launch; results come back in input
// order.  The f2 slab is first transposed to pair-major, survivor-only order so a
// model's f4 basis reads each f2 series contiguously.
#include <algorithm>
#include <numeric>
#include <vector>

#include "engine.hpp"
#include "qpadm.hpp"

namespace stp {

namespace {

struct ModelShape {
    int nl, nr, m, T;
    double cost;
};

// The normal-matrix bucket of a model: rank sweeps and LOO fits run at ranks up to
// nl - 1, so the largest systems are nl*(nl-1) (A-step) and (nl-1)*nr (B-step).
// Models the batched kernel fits: one f4 entry per lane, one normal-matrix row per lane.
bool fits_batched(int nl, int nr) { return nl * nr <= kMaxQpadmM && nr <= kMaxQpadmNr; }

// Largest normal system of a large model's fits (rank sweep and refits up to nl - 1).
int large_t(int nl, int nr, int rank) {
    const int rmax = std::max(nl - 1, rank);
    return std::max({nl * rmax, rmax * nr, nl});
}

int bucket_of(int nl, int nr, int rank) {
    const int rmax = std::max(nl - 1, rank);
    const int t = std::max({nl * rmax, rmax * nr, nl});
    if (t <= 16) return 16;
    if (t <= 32) return 32;
    if (t <= 48) return 48;
    return 64;
}

}  // namespace

void HipBackend::qpadm_batch(const F2Blocks& h, int n_models, const std::int32_t* target,
                             const std::int32_t* left_off, const std::int32_t* left, const std::int32_t* right_off,
                             const std::int32_t* right, const stp_qpadm_options& opt, stp_qpadm_result* out) {
    activate();
    STP_REQUIRE(n_models >= 0, "qpadm: n_models must be non-negative");
    if (n_models == 0) return;
    STP_REQUIRE(target && left_off && left && right_off && right && out, "qpadm: null model arena or output");
    STP_REQUIRE(opt.als_iterations >= 0, "qpadm: als_iterations must be non-negative");
    STP_REQUIRE(opt.jackknife >= 0 && opt.jackknife <= 2, "qpadm: jackknife must be 0, 1 or 2");
    STP_REQUIRE(opt.rank >= -1 && opt.rank < STP_QPADM_MAX_RANK, "qpadm: rank out of range");
    const int P = h.P, nb = h.n_block;

    // ---- validate models ----
    std::vector<ModelShape> shape(n_models);
    int total_left = 0, total_right = 0;
    for (int i = 0; i < n_models; ++i) {
        const int nl = left_off[i + 1] - left_off[i];
        const int nr = right_off[i + 1] - right_off[i] - 1;
        const std::string tag = "qpadm: model " + std::to_string(i) + ": ";
        STP_REQUIRE(left_off[i] == total_left && right_off[i] == total_right, tag + "offsets are not cumulative");
        STP_REQUIRE(nl >= 1, tag + "needs at least one source population");
        STP_REQUIRE(nr >= 1, tag + "needs at least two right populations");
        STP_REQUIRE(nl <= STP_QPADM_MAX_NL, tag + "more than " + std::to_string(STP_QPADM_MAX_NL) + " sources");
        // The reference seeds the rank nl - 1 fit from nl - 1 right singular vectors of the
        // nl x nr f4 matrix (seed_AB, cpu_backend.cpp:2738-2744), which exist for nl <= nr + 1.
        STP_REQUIRE(nl <= nr + 1, tag + "needs at least sources - 1 right populations beyond the first");
        STP_REQUIRE(opt.rank < nl, tag + "rank must be below the number of sources");
        if (!fits_batched(nl, nr)) {
            STP_REQUIRE(nl * nr <= kQpLargeMaxM && nr <= kQpLargeMaxNr && large_t(nl, nr, opt.rank) <= kQpLargeMaxT,
                        tag + "model too large (sources x (right - 1) must be <= " + std::to_string(kQpLargeMaxM) +
                            ", right - 1 <= " + std::to_string(kQpLargeMaxNr) + ")");
        }
        STP_REQUIRE(target[i] >= 0 && target[i] < P, tag + "target index out of range");
        for (int k = left_off[i]; k < left_off[i + 1]; ++k)
            STP_REQUIRE(left[k] >= 0 && left[k] < P, tag + "left index out of range");
        for (int k = right_off[i]; k < right_off[i + 1]; ++k)
            STP_REQUIRE(right[k] >= 0 && right[k] < P, tag + "right index out of range");
        total_left += nl;
        total_right += nr + 1;
        const int T = bucket_of(nl, nr, opt.rank);
        const double t = static_cast<double>(T);
        shape[i] = ModelShape{nl, nr, nl * nr, T, (nb + 8.0 * nl) * t * t * t};
    }
    STP_REQUIRE(left_off[n_models] == total_left && right_off[n_models] == total_right,
                "qpadm: offsets are not cumulative");

    // ---- survivor blocks ----
    STP_REQUIRE(h.tier == kTierResident,
                "qpadm: the f2 tensor is held in host memory or on disk; fits need it resident in HBM");
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;
    qp_keep_.ensure(static_cast<std::size_t>(nb));
    launch_survivors(h.vpair.get(), slab, nb, qp_keep_.get(), stream_);
    std::vector<std::int32_t> keep(nb);
    STP_HIP(hipMemcpyAsync(keep.data(), qp_keep_.get(), nb * sizeof(std::int32_t), hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipStreamSynchronize(stream_));
    std::vector<std::int32_t> surv, bsz;
    double n_total = 0.0;
    for (int b = 0; b < nb; ++b)
        if (keep[b]) {
            surv.push_back(b);
            bsz.push_back(h.block_sizes[b]);
            n_total += static_cast<double>(h.block_sizes[b]);
        }
    const int nbs = static_cast<int>(surv.size());
    STP_REQUIRE(nbs >= 1, "qpadm: no jackknife block survives (every block has partially missing pairs)");
    for (int s : bsz) STP_REQUIRE(s > 0, "qpadm: surviving blocks must have positive sizes");

    // ---- device arena: survivors, sizes, models, per-bucket claim orders ----
    std::vector<int> order, large;
    for (int i = 0; i < n_models; ++i) (fits_batched(shape[i].nl, shape[i].nr) ? order : large).push_back(i);
    const int n_small = static_cast<int>(order.size());
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        if (shape[a].T != shape[b].T) return shape[a].T < shape[b].T;
        return shape[a].cost > shape[b].cost;
    });
    std::vector<std::int32_t> arena;
    auto put = [&](const std::int32_t* p, std::size_t n) {
        const std::size_t off = arena.size();
        arena.insert(arena.end(), p, p + n);
        return off;
    };
    const std::size_t o_surv = put(surv.data(), surv.size());
    const std::size_t o_bsz = put(bsz.data(), bsz.size());
    const std::size_t o_tgt = put(target, n_models);
    const std::size_t o_loff = put(left_off, n_models + 1);
    const std::size_t o_left = put(left, total_left);
    const std::size_t o_roff = put(right_off, n_models + 1);
    const std::size_t o_right = put(right, total_right);
    const std::size_t o_order = put(order.data(), order.size());
    qp_i32_.ensure(arena.size());
    STP_HIP(hipMemcpyAsync(qp_i32_.get(), arena.data(), arena.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                           stream_));

    // ---- survivor-transposed f2 ----
    qpT_.ensure(static_cast<std::size_t>(slab) * nbs);
    launch_transpose_survivors(h.f2.get(), slab, qp_i32_.get() + o_surv, nbs, qpT_.get(), stream_);

    // ---- launches per bucket ----
    qp_out_.ensure(static_cast<std::size_t>(n_models));
    STP_HIP(hipMemsetAsync(qp_out_.get(), 0, static_cast<std::size_t>(n_models) * sizeof(stp_qpadm_result), stream_));
    // Each launch group fills every CU to the kernel's occupancy; every workgroup has its
    // own scratch slice (X, LOO and the LOO weights), sized for the largest group.
    std::size_t scratch_elems = 0;
    for (int slot = 0; slot < n_small;) {
        const int T = shape[order[slot]].T;
        int end = slot, m_max = 0, nr_max = 0;
        for (; end < n_small && shape[order[end]].T == T; ++end) {
            m_max = std::max(m_max, shape[order[end]].m);
            nr_max = std::max(nr_max, shape[order[end]].nr);
        }
        const std::size_t lds = qpadm_lds_bytes(m_max, m_max + 1, qpadm_wave_slice(nr_max));
        const std::int64_t n_wg = std::min<std::int64_t>(end - slot, qpadm_blocks_per_cu(T, lds) * n_cu_);
        scratch_elems = std::max(scratch_elems, static_cast<std::size_t>((2 * m_max + STP_QPADM_MAX_NL) * nbs * n_wg));
        slot = end;
    }
    qp_scratch_.ensure(scratch_elems);
    counter_.ensure(64);
    mark_begin(kQpadm);
    int slot = 0;
    int bucket_id = 0;
    while (slot < n_small) {
        const int T = shape[order[slot]].T;
        int end = slot;
        int m_max = 0, nr_max = 0;
        while (end < n_small && shape[order[end]].T == T) {
            m_max = std::max(m_max, shape[order[end]].m);
            nr_max = std::max(nr_max, shape[order[end]].nr);
            ++end;
        }
        QpadmArgs a{};
        a.f2T = qpT_.get();
        a.bsz = qp_i32_.get() + o_bsz;
        a.P = P;
        a.nbs = nbs;
        a.n_total = n_total;
        a.target = qp_i32_.get() + o_tgt;
        a.left_off = qp_i32_.get() + o_loff;
        a.left = qp_i32_.get() + o_left;
        a.right_off = qp_i32_.get() + o_roff;
        a.right = qp_i32_.get() + o_right;
        a.order = qp_i32_.get() + o_order + slot;
        a.n_order = end - slot;
        a.counter = counter_.get() + 32 + bucket_id;
        a.fudge = opt.fudge;
        a.iters = opt.als_iterations;
        a.rank = opt.rank;
        a.alpha = opt.rank_alpha;
        a.jackknife = opt.jackknife;
        a.se_require_p = opt.se_require_p;
        a.p_se_threshold = opt.p_se_threshold;
        a.scratch = qp_scratch_.get();
        a.scratch_stride = static_cast<std::int64_t>(2 * m_max + STP_QPADM_MAX_NL) * nbs;
        a.out = qp_out_.get();
        a.m_max = m_max;
        a.ldq = m_max + 1;
        a.nr_max = nr_max;
        a.wave_slice = qpadm_wave_slice(nr_max);
        const std::size_t lds = qpadm_lds_bytes(a.m_max, a.ldq, a.wave_slice);
        STP_REQUIRE(lds <= 96 * 1024, "qpadm: model sizes need more LDS than a workgroup has");
        launch_zero_u32(a.counter, 1, stream_);
        launch_qpadm(T, a, std::min(qpadm_blocks_per_cu(T, lds) * n_cu_, a.n_order), stream_);
        slot = end;
        ++bucket_id;
        STP_REQUIRE(bucket_id < 32, "qpadm: too many launch groups");
    }
    if (!large.empty()) {
        QpLargeArgs a{};
        a.f2T = qpT_.get();
        a.bsz = qp_i32_.get() + o_bsz;
        a.P = P;
        a.nbs = nbs;
        a.n_total = n_total;
        a.target = qp_i32_.get() + o_tgt;
        a.left = qp_i32_.get() + o_left;
        a.right = qp_i32_.get() + o_right;
        a.out = qp_out_.get();
        a.fudge = opt.fudge;
        a.iters = opt.als_iterations;
        a.alpha = opt.rank_alpha;
        a.jackknife = opt.jackknife;
        a.se_require_p = opt.se_require_p;
        a.p_se_threshold = opt.p_se_threshold;
        std::vector<QpLargeModel> lm;
        for (int mi : large) {
            QpLargeModel md{};
            md.mi = mi;
            md.l0 = left_off[mi];
            md.nl = shape[mi].nl;
            md.r0 = right_off[mi];
            md.nr = shape[mi].nr;
            md.m = md.nl * md.nr;
            md.rank = opt.rank < 0 ? md.nl - 1 : opt.rank;
            lm.push_back(md);
        }
        qpadm_large(a, lm, opt.rank);
    }
    mark_end(kQpadm);
    STP_HIP(hipMemcpyAsync(out, qp_out_.get(), static_cast<std::size_t>(n_models) * sizeof(stp_qpadm_result),
                           hipMemcpyDeviceToHost, stream_));
    sync();
}

// Models beyond the batched kernel: chunks of models whose setup scratch fits a fixed
// budget; per chunk the setup kernel, the fit-table fits, the leave-one-block-out refits
// and the record kernel, each one launch.
void HipBackend::qpadm_large(QpLargeArgs a, const std::vector<QpLargeModel>& large, int rank_opt) {
    constexpr std::size_t kSetupBudget = std::size_t{1} << 29;   // doubles (4 GiB)
    constexpr std::size_t kFitBudget = std::size_t{1} << 28;     // doubles (2 GiB)
    const int nbs = a.nbs;
    std::size_t i0 = 0;
    while (i0 < large.size()) {
        // ---- the chunk's models and their setup scratch ----
        std::vector<QpLargeModel> models;
        std::size_t setup = 0;
        int m_max = 0, nr_max = 0, t_max = 0;
        for (; i0 < large.size(); ++i0) {
            QpLargeModel md = large[i0];
            const int nl = md.nl, nr = md.nr, m = md.m;
            const std::size_t need = qpadm_large_setup_doubles(m, nbs);
            if (!models.empty() && (setup + need > kSetupBudget || models.size() >= 4096)) break;
            md.base = static_cast<std::int64_t>(setup);
            models.push_back(md);
            setup += need;
            m_max = std::max(m_max, m);
            nr_max = std::max(nr_max, nr);
            t_max = std::max(t_max, large_t(nl, nr, rank_opt));
        }
        const int nm = static_cast<int>(models.size());
        // ---- fit tasks: every model's fit table, then its leave-one-block-out refits ----
        std::vector<QpLargeTask> tasks;
        for (int sl = 0; sl < nm; ++sl) {
            const int nl = models[sl].nl, nr = models[sl].nr, rank = models[sl].rank;
            const int nrows = nl >= 2 ? nl + 1 : 1;
            for (int k = 0; k < nrows; ++k) {
                const int nlr = k == 0 ? nl : nl - 1;
                const int rmax = std::min(nlr, nr) - 1, rf = nlr - 1;
                for (int rr = 0; rr <= rmax; ++rr) tasks.push_back({sl, k, rr, -1});
                if (rf > rmax) tasks.push_back({sl, k, rf, -1});
                if (k == 0 && rank > rmax && rank != rf) tasks.push_back({sl, 0, rank, -1});
            }
        }
        const int n_table = static_cast<int>(tasks.size());
        if (a.jackknife != 0 && nbs >= 2)
            for (int sl = 0; sl < nm; ++sl)
                for (int b = 0; b < nbs; ++b) tasks.push_back({sl, 0, 0, b});
        const int n_loo = static_cast<int>(tasks.size()) - n_table;

        ql_models_.ensure(static_cast<std::size_t>(nm));
        STP_HIP(hipMemcpyAsync(ql_models_.get(), models.data(), nm * sizeof(QpLargeModel), hipMemcpyHostToDevice,
                               stream_));
        ql_tasks_.ensure(tasks.size());
        STP_HIP(hipMemcpyAsync(ql_tasks_.get(), tasks.data(), tasks.size() * sizeof(QpLargeTask),
                               hipMemcpyHostToDevice, stream_));
        ql_tables_.ensure(static_cast<std::size_t>(nm));
        launch_zero_u32(reinterpret_cast<unsigned*>(ql_tables_.get()),
                        static_cast<int>(nm * sizeof(QpLargeTables) / sizeof(unsigned)), stream_);
        ql_setup_.ensure(setup);
        ql_wm_.ensure(static_cast<std::size_t>(nm) * nbs * STP_QPADM_MAX_NL);
        const std::size_t per_wave = qpadm_large_fit_doubles(m_max, nr_max, t_max);
        const int waves_per_wg = qpadm_large_fit_waves();
        int n_wg = std::min<int>(2 * n_cu_, static_cast<int>(ceil_div(std::max(n_table, n_loo), waves_per_wg)));
        n_wg = std::max(1, std::min<int>(n_wg, static_cast<int>(kFitBudget / (per_wave * waves_per_wg))));
        ql_fit_.ensure(per_wave * waves_per_wg * n_wg);
        counter_.ensure(64);

        a.models = ql_models_.get();
        a.n_models = nm;
        a.tables = ql_tables_.get();
        a.setup = ql_setup_.get();
        a.fit = ql_fit_.get();
        a.fit_stride = static_cast<std::int64_t>(per_wave);
        a.m_max = m_max;
        a.nr_max = nr_max;
        a.t_max = t_max;
        a.wm = ql_wm_.get();
        a.counter = counter_.get() + 30;
        launch_qpadm_large_setup(a, stream_);
        a.tasks = ql_tasks_.get();
        a.n_tasks = n_table;
        launch_zero_u32(a.counter, 1, stream_);
        launch_qpadm_large_fits(a, n_wg, stream_);
        if (n_loo > 0) {
            a.tasks = ql_tasks_.get() + n_table;
            a.n_tasks = n_loo;
            launch_zero_u32(a.counter, 1, stream_);
            launch_qpadm_large_fits(a, n_wg, stream_);
        }
        launch_qpadm_large_finish(a, stream_);
        // the chunk's task list (host memory) must outlive its copy
        sync();
    }
}

}  // namespace stp
