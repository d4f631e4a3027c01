// steppe_amd/csrc/f2_graph.cpp — one resident f2 extraction captured as a HIP graph.
//
// The reference drives an extraction from the host, chunk by chunk: per size bucket a
// decode, the operand staging and the batched GEMMs, with a synchronisation after each
// chunk (src/device/cuda/cuda_backend_f2_blocks.cu:301-347).  Here the whole resident
// extraction of a fixed shape (tile geometry, block partition, precision, output
// buffers) is a handful of kernels on one stream: ploidy detection, decode, operand
// prep, pair counts, the persistent contraction and the fix-up.  They are captured once
// into a HIP graph and replayed with a single launch, so a repeated extraction (a
// benchmark step, a service re-reading new genotypes into the same buffers) costs one
// submission.  The graph owns every buffer its kernels were captured with; clearing
// counters is done by kernels (no memset nodes).
#include <cstring>
#include <memory>
#include <vector>

#include "engine.hpp"
#include "steppe_hip.h"

namespace stp {

F2Graph::~F2Graph() {
    (void)hipSetDevice(device);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

void HipBackend::record_f2_graph(F2Graph& g) {
    if (g.ploidy_mode == 0)
        launch_detect_ploidy(g.packed, static_cast<std::int64_t>(g.bpr), g.M, static_cast<std::int64_t>(g.n_ind),
                             g.ploidy.get(), stream_);
    mark_begin(kDecode);
    launch_decode_counts(g.packed, static_cast<std::int64_t>(g.bpr), g.M, g.pop_offsets.get(), g.ploidy.get(), g.P,
                         g.a2, g.n, g.ld, stream_);
    mark_end(kDecode);
    F2Source src;
    src.a2 = g.a2;
    src.n = g.n;
    src.ld = g.ld;
    run_f2(g.plan, g.precision, src, g.f2, g.vpair, g.work);
}

F2Graph* HipBackend::f2_graph_create(const stp_tile& t, const std::int32_t* block_id, int n_block, int precision,
                                     std::uint16_t* a2, std::uint16_t* n, std::int64_t ld, double* f2,
                                     double* vpair) {
    activate();
    STP_REQUIRE(t.packed_on_device, "f2 graph: the tile must be resident on the device (packed_on_device)");
    STP_REQUIRE(t.n_pop > 0 && t.pop_offsets != nullptr, "f2 graph: at least one population is required");
    STP_REQUIRE(t.n_individuals > 0 && t.packed != nullptr, "f2 graph: packed records are required");
    STP_REQUIRE(t.bytes_per_record >= (t.n_snp + 3) / 4, "f2 graph: bytes_per_record < ceil(n_snp/4)");
    STP_REQUIRE(f2 != nullptr && vpair != nullptr, "f2 graph: f2 and vpair outputs are required");
    STP_REQUIRE(precision == STP_PREC_FP64 || precision == STP_PREC_SPLIT_BF16,
                "f2 graph: unknown precision form " + std::to_string(precision));
    STP_REQUIRE((a2 == nullptr) == (n == nullptr), "f2 graph: give both count buffers or neither");
    const std::int64_t M = static_cast<std::int64_t>(t.n_snp);
    for (int p = 0; p < t.n_pop; ++p) {
        STP_REQUIRE(t.pop_offsets[p] <= t.pop_offsets[p + 1], "f2 graph: pop_offsets must be non-decreasing");
        STP_REQUIRE(t.pop_offsets[p + 1] - t.pop_offsets[p] <= static_cast<std::uint64_t>(kMaxPopForU16),
                    "f2 graph: a population exceeds 16383 individuals (uint16 count limit)");
    }
    STP_REQUIRE(t.pop_offsets[t.n_pop] <= t.n_individuals, "f2 graph: pop_offsets exceed n_individuals");

    auto g = std::make_unique<F2Graph>();
    g->device = device_;
    g->precision = precision;
    g->P = t.n_pop;
    g->M = M;
    g->bpr = t.bytes_per_record;
    g->n_ind = t.n_individuals;
    g->packed = t.packed;
    g->f2 = f2;
    g->vpair = vpair;
    g->plan = make_f2_plan(t.n_pop, M, block_id, n_block);
    const F2Tune tn = tune_for(g->plan, precision);
    if (tn.str() != g->plan.tune.str()) g->plan = make_f2_plan(t.n_pop, M, block_id, n_block, tn);
    upload_plan(g->plan, g->work.plan);
    if (a2) {
        STP_REQUIRE(ld >= M && ld % 16 == 0, "f2 graph: ld must be >= n_snp and a multiple of 16");
        g->a2 = a2;
        g->n = n;
        g->ld = ld;
    } else {
        g->ld = round_up(std::max<std::int64_t>(M, 1), 64);
        g->own_a2.alloc(static_cast<std::size_t>(g->P) * g->ld);
        g->own_n.alloc(static_cast<std::size_t>(g->P) * g->ld);
        g->a2 = g->own_a2.get();
        g->n = g->own_n.get();
    }
    std::vector<std::uint64_t> offs(t.pop_offsets, t.pop_offsets + t.n_pop + 1);
    g->pop_offsets.alloc(offs.size());
    STP_HIP(hipMemcpyAsync(g->pop_offsets.get(), offs.data(), offs.size() * sizeof(std::uint64_t),
                           hipMemcpyHostToDevice, stream_));
    g->ploidy.alloc(t.n_individuals);
    std::vector<std::int8_t> pl;
    if (t.sample_ploidy) {
        g->ploidy_mode = 1;
        pl.resize(t.n_individuals);
        for (std::size_t i = 0; i < t.n_individuals; ++i) {
            const int v = t.sample_ploidy[i];
            pl[i] = static_cast<std::int8_t>((v == 1 || v == 2) ? v : 0);
        }
        STP_HIP(hipMemcpyAsync(g->ploidy.get(), pl.data(), pl.size(), hipMemcpyHostToDevice, stream_));
    } else if (t.detect_ploidy_on_device) {
        g->ploidy_mode = 0;
    } else {
        g->ploidy_mode = 2;
        g->fixed_ploidy = (t.ploidy == 1 || t.ploidy == 2) ? t.ploidy : 0;
        launch_fill_ploidy(g->ploidy.get(), static_cast<std::int64_t>(t.n_individuals), g->fixed_ploidy, stream_);
    }

    // One eager pass sizes every workspace buffer (so that nothing is allocated while
    // capturing) and leaves the outputs complete.
    record_f2_graph(*g);
    sync();

    // Capture.  Stage events are captured as event-record nodes when timing is on; if the
    // runtime refuses them the graph is captured without them.
    auto capture = [&](bool marks) {
        const bool saved_timing = timing_;
        bool saved_pending[kStages];
        std::memcpy(saved_pending, pending_, sizeof(pending_));
        timing_ = saved_timing && marks;
        hipGraph_t gr = nullptr;
        STP_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
            record_f2_graph(*g);
        } catch (...) {
            (void)hipStreamEndCapture(stream_, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            (void)hipGetLastError();
            timing_ = saved_timing;
            std::memcpy(pending_, saved_pending, sizeof(pending_));
            throw;
        }
        const hipError_t e = hipStreamEndCapture(stream_, &gr);
        const bool captured_marks = timing_;
        timing_ = saved_timing;
        std::memcpy(pending_, saved_pending, sizeof(pending_));
        if (e != hipSuccess) {
            if (gr) (void)hipGraphDestroy(gr);
            hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        g->graph = gr;
        const hipError_t ei = hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(gr);
            g->graph = nullptr;
            hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        g->timed = captured_marks;
    };
    if (timing_) {
        try {
            capture(true);
        } catch (const DeviceError&) {
            capture(false);
        }
    } else {
        capture(false);
    }
    return g.release();
}

void HipBackend::f2_graph_launch(F2Graph& g) {
    activate();
    STP_REQUIRE(g.device == device_, "f2 graph: launched on another device's context");
    STP_HIP(hipGraphLaunch(g.exec, stream_));
    if (g.timed && timing_) {
        pending_[kDecode] = pending_[kPrep] = pending_[kCount] = pending_[kF2] = pending_[kFixup] = true;
    }
    sync();
}

}  // namespace stp
