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
        STP_REQUIRE(nl <= nr, tag + "needs at least as many right populations beyond the first as sources");
        STP_REQUIRE(nl * nr <= kMaxQpadmM && nr <= kMaxQpadmNr,
                    tag + "model too large for the device fit (sources x (right - 1) must be <= 64)");
        STP_REQUIRE(opt.rank < nl, tag + "rank must be below the number of sources");
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
    std::vector<int> order(n_models);
    std::iota(order.begin(), order.end(), 0);
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
    for (int slot = 0; slot < n_models;) {
        const int T = shape[order[slot]].T;
        int end = slot, m_max = 0, nr_max = 0;
        for (; end < n_models && shape[order[end]].T == T; ++end) {
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
    while (slot < n_models) {
        const int T = shape[order[slot]].T;
        int end = slot;
        int m_max = 0, nr_max = 0;
        while (end < n_models && shape[order[end]].T == T) {
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
    mark_end(kQpadm);
    STP_HIP(hipMemcpyAsync(out, qp_out_.get(), static_cast<std::size_t>(n_models) * sizeof(stp_qpadm_result),
                           hipMemcpyDeviceToHost, stream_));
    sync();
}

}  // namespace stp
