// steppe_amd/csrc/engine.cpp — HipBackend: host orchestration of the hot path.
//
// decode   : tile staging (H2D when the tile is host-resident, as the reference's
//            cuda_backend_decode.cu:64-69), ploidy resolution, one decode launch.
// f2       : block plan (ranges, padded K layout, largest-block-first work items),
//            one prep launch and one persistent contraction launch, no per-block or
//            per-bucket host loop (the reference loops buckets x chunks with a sync
//            per chunk, cuda_backend_f2_blocks.cu:319-341).
#include <cstdlib>

#include "engine.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>

#include "steppe_hip.h"

namespace stp {

// ---------------------------------------------------------------------------------------
F2Plan make_f2_plan(int P, std::int64_t M, const std::int32_t* block_id, int n_block, const F2Tune& tune) {
    STP_REQUIRE(P > 0, "f2: P must be positive");
    STP_REQUIRE(M >= 0, "f2: M must be non-negative");
    STP_REQUIRE(n_block > 0, "f2: n_block must be positive");
    STP_REQUIRE(M == 0 || block_id != nullptr, "f2: block_id is required");
    F2Plan pl;
    pl.tune = tune;
    pl.P = P;
    pl.M = M;
    pl.n_block = n_block;
    pl.G = static_cast<int>(ceil_div(P, kRowGroup));
    pl.T = static_cast<int>(ceil_div(pl.G, kTileRows / kRowGroup));
    pl.s0.assign(n_block, 0);
    pl.len.assign(n_block, 0);
    int prev = -1;
    std::int64_t s = 0;
    while (s < M) {
        const int b = block_id[s];
        if (b < 0 || b >= n_block)
            throw ArgError("block_id[" + std::to_string(s) + "] = " + std::to_string(b) +
                           " is out of range [0, " + std::to_string(n_block) + ")");
        if (b < prev)
            throw ArgError("block_id is not non-decreasing at column " + std::to_string(s) +
                           "; the partition is not contiguous");
        std::int64_t e = s;
        while (e < M && block_id[e] == b) ++e;
        pl.s0[b] = s;
        pl.len[b] = static_cast<std::int32_t>(e - s);
        prev = b;
        s = e;
    }
    pl.k0.assign(n_block, 0);
    pl.kpad.assign(n_block, 0);
    std::int64_t k = 0;
    for (int b = 0; b < n_block; ++b) {
        pl.k0[b] = k;
        pl.kpad[b] = static_cast<std::int32_t>(round_up(pl.len[b], kKPad));
        k += pl.kpad[b];
    }
    pl.K_total = k;
    pl.chunk_block.resize(static_cast<std::size_t>(k / kKPad));
    for (int b = 0; b < n_block; ++b)
        for (std::int64_t c = pl.k0[b] / kKPad; c < (pl.k0[b] + pl.kpad[b]) / kKPad; ++c)
            pl.chunk_block[static_cast<std::size_t>(c)] = b;
    // Work items.  Every block goes to one die's queue (greedy longest-first onto the
    // least-loaded die, so the dies finish together), and each queue is block-major:
    // all tile pairs of a block sit together, so the workgroups of one die run a
    // block's tiles concurrently and share its operand panels in that die's L2.
    std::vector<int> order(n_block);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pl.len[a] > pl.len[b]; });
    std::vector<std::vector<int>> die_blocks(kXcds);
    if (tune.sched == 1) {
        // one shared queue: every worker drains it (the claim falls through to it)
        die_blocks[0] = order;
    } else {
        std::vector<std::int64_t> load(kXcds, 0);
        for (int b : order) {
            const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            load[x] += pl.kpad[b];
            die_blocks[x].push_back(b);
        }
    }
    auto build = [&](int T, std::vector<F2Item>& items, ItemQueues& q) {
        items.clear();
        items.reserve(static_cast<std::size_t>(n_block) * T * (T + 1) / 2);
        for (int x = 0; x < kXcds; ++x) {
            q.off[x] = static_cast<std::int32_t>(items.size());
            for (int b : die_blocks[x])
                for (int ti = 0; ti < T; ++ti)
                    for (int tj = ti; tj < T; ++tj)
                        items.push_back(F2Item{b, static_cast<std::int16_t>(ti), static_cast<std::int16_t>(tj)});
        }
        q.off[kXcds] = static_cast<std::int32_t>(items.size());
    };
    build(pl.T, pl.items, pl.q64);
    pl.G32 = static_cast<int>(ceil_div(P, 32));
    pl.T128 = static_cast<int>(ceil_div(pl.G32, 4));
    build(pl.T128, pl.items128, pl.q128);
    return pl;
}

// ---------------------------------------------------------------------------------------
HipBackend::HipBackend(int device) : device_(device) {
    int n = 0;
    STP_HIP(hipGetDeviceCount(&n));
    STP_REQUIRE(device >= 0 && device < n, "no HIP device " + std::to_string(device) + " (visible: " +
                                               std::to_string(n) + ")");
    STP_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    STP_HIP(hipGetDeviceProperties(&prop, device));
    n_cu_ = prop.multiProcessorCount;
    name_ = prop.gcnArchName;
    STP_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    counter_.alloc(64);
    noalloc_ = resolve_noalloc_scratch();
}

HipBackend::~HipBackend() {
    for (int i = 0; i < kStages; ++i) {
        if (ev_begin_[i]) (void)hipEventDestroy(ev_begin_[i]);
        if (ev_end_[i]) (void)hipEventDestroy(ev_end_[i]);
    }
    if (stream_) {
        (void)hipSetDevice(device_);
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
}

void HipBackend::activate() const { STP_HIP(hipSetDevice(device_)); }

void HipBackend::sync() {
    STP_HIP(hipStreamSynchronize(stream_));
    if (timing_) collect_timing();
}

void HipBackend::set_timing(bool on) {
    activate();
    if (on && !ev_begin_[0]) {
        for (int i = 0; i < kStages; ++i) {
            STP_HIP(hipEventCreate(&ev_begin_[i]));
            STP_HIP(hipEventCreate(&ev_end_[i]));
        }
    }
    timing_ = on;
}

void HipBackend::mark_begin(Stage s) {
    if (timing_) STP_HIP(hipEventRecord(ev_begin_[s], stream_));
}

void HipBackend::mark_end(Stage s) {
    if (!timing_) return;
    STP_HIP(hipEventRecord(ev_end_[s], stream_));
    pending_[s] = true;
}

void HipBackend::collect_timing() {
    for (int i = 0; i < kStages; ++i) {
        if (!pending_[i]) continue;
        float ms = 0.f;
        STP_HIP(hipEventElapsedTime(&ms, ev_begin_[i], ev_end_[i]));
        total_ms_[i] += ms;
        ++count_[i];
        pending_[i] = false;
    }
}

void HipBackend::read_timing(double* ms, std::int64_t* count, bool reset) {
    for (int i = 0; i < kStages; ++i) {
        if (ms) ms[i] = total_ms_[i];
        if (count) count[i] = count_[i];
        if (reset) { total_ms_[i] = 0.0; count_[i] = 0; }
    }
}

BlockTable F2Work::PlanDev::table() const {
    return BlockTable{s0.get(), len.get(), k0.get(), kpad.get()};
}

template <class T>
static void h2d(DevBuf<T>& buf, const std::vector<T>& v, hipStream_t st) {
    buf.ensure(std::max<std::size_t>(v.size(), 1));
    if (!v.empty())
        STP_HIP(hipMemcpyAsync(buf.get(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

void HipBackend::upload_plan(const F2Plan& plan, F2Work::PlanDev& pd) {
    h2d(pd.s0, plan.s0, stream_);
    h2d(pd.k0, plan.k0, stream_);
    h2d(pd.len, plan.len, stream_);
    h2d(pd.kpad, plan.kpad, stream_);
    h2d(pd.chunk_block, plan.chunk_block, stream_);
    h2d(pd.items, plan.items, stream_);
    h2d(pd.items128, plan.items128, stream_);
}

const F2Plan& HipBackend::plan_for(int P, std::int64_t M, const std::int32_t* block_id, int n_block, int precision) {
    const bool same = plan_valid_ && P == plan_.P && M == plan_.M && n_block == plan_.n_block &&
                      (M == 0 || std::memcmp(plan_bid_.data(), block_id, static_cast<std::size_t>(M) * 4) == 0);
    if (same) {
        const F2Tune t = tune_for(plan_, precision);
        if (t.str() == plan_.tune.str()) return plan_;
    }
    plan_valid_ = false;
    plan_ = make_f2_plan(P, M, block_id, n_block);
    const F2Tune t = tune_for(plan_, precision);
    if (t.str() != plan_.tune.str()) plan_ = make_f2_plan(P, M, block_id, n_block, t);
    plan_bid_.assign(block_id, block_id + M);
    upload_plan(plan_, work_.plan);
    sync();
    plan_valid_ = true;
    return plan_;
}

// ---------------------------------------------------------------------------------------
void HipBackend::prepare_tile(const stp_tile& t) {
    STP_REQUIRE(t.n_pop >= 0, "decode: n_pop must be non-negative");
    STP_REQUIRE(t.n_pop == 0 || t.pop_offsets != nullptr, "decode: pop_offsets is required");
    STP_REQUIRE(t.bytes_per_record >= (t.n_snp + 3) / 4, "decode: bytes_per_record < ceil(n_snp/4)");
    STP_REQUIRE(t.n_individuals == 0 || t.packed != nullptr, "decode: packed is required");
    std::vector<std::uint64_t> offs(t.pop_offsets, t.pop_offsets + t.n_pop + 1);
    for (int p = 0; p < t.n_pop; ++p) {
        STP_REQUIRE(offs[p] <= offs[p + 1], "decode: pop_offsets must be non-decreasing");
        STP_REQUIRE(offs[p + 1] - offs[p] <= static_cast<std::uint64_t>(kMaxPopForU16),
                    "decode: a population exceeds 16383 individuals (uint16 count limit)");
    }
    if (t.n_pop > 0)
        STP_REQUIRE(offs[t.n_pop] <= t.n_individuals, "decode: pop_offsets exceed n_individuals");

    const std::size_t bytes = t.n_individuals * t.bytes_per_record;
    if (t.packed_on_device) {
        d_packed_ = t.packed;
    } else {
        packed_stage_.ensure(std::max<std::size_t>(bytes, 4));
        if (bytes)
            STP_HIP(hipMemcpyAsync(packed_stage_.get(), t.packed, bytes, hipMemcpyHostToDevice, stream_));
        d_packed_ = packed_stage_.get();
    }
    h2d(pop_offsets_, offs, stream_);
    ploidy_.ensure(std::max<std::size_t>(t.n_individuals, 1));
    if (t.sample_ploidy) {
        std::vector<std::int8_t> pl(t.n_individuals);
        for (std::size_t g = 0; g < t.n_individuals; ++g) {
            const int v = t.sample_ploidy[g];
            pl[g] = static_cast<std::int8_t>((v == 1 || v == 2) ? v : 0);
        }
        STP_HIP(hipMemcpyAsync(ploidy_.get(), pl.data(), pl.size(), hipMemcpyHostToDevice, stream_));
        // pl is released at scope exit: finish the copy first.
        sync();
    } else if (t.detect_ploidy_on_device) {
        launch_detect_ploidy(d_packed_, static_cast<std::int64_t>(t.bytes_per_record),
                             static_cast<std::int64_t>(t.n_snp), static_cast<std::int64_t>(t.n_individuals),
                             ploidy_.get(), stream_);
    } else {
        const int v = (t.ploidy == 1 || t.ploidy == 2) ? t.ploidy : 0;
        launch_fill_ploidy(ploidy_.get(), static_cast<std::int64_t>(t.n_individuals), v, stream_);
    }
    // offs is pageable and local: complete its copy before it goes out of scope.
    sync();
}

std::vector<int> HipBackend::detect_ploidy(const stp_tile& t) {
    activate();
    STP_REQUIRE(t.bytes_per_record >= (t.n_snp + 3) / 4, "detect_ploidy: bytes_per_record < ceil(n_snp/4)");
    const std::size_t bytes = t.n_individuals * t.bytes_per_record;
    const std::uint8_t* d = t.packed;
    if (!t.packed_on_device) {
        packed_stage_.ensure(std::max<std::size_t>(bytes, 4));
        if (bytes) STP_HIP(hipMemcpyAsync(packed_stage_.get(), t.packed, bytes, hipMemcpyHostToDevice, stream_));
        d = packed_stage_.get();
    }
    ploidy_.ensure(std::max<std::size_t>(t.n_individuals, 1));
    launch_detect_ploidy(d, static_cast<std::int64_t>(t.bytes_per_record), static_cast<std::int64_t>(t.n_snp),
                         static_cast<std::int64_t>(t.n_individuals), ploidy_.get(), stream_);
    std::vector<std::int8_t> h(t.n_individuals);
    if (!h.empty())
        STP_HIP(hipMemcpyAsync(h.data(), ploidy_.get(), h.size(), hipMemcpyDeviceToHost, stream_));
    sync();
    return std::vector<int>(h.begin(), h.end());
}

void HipBackend::decode_counts(const stp_tile& t, std::uint16_t* d_a2, std::uint16_t* d_n, std::int64_t ld) {
    activate();
    prepare_tile(t);
    mark_begin(kDecode);
    launch_decode_counts(d_packed_, static_cast<std::int64_t>(t.bytes_per_record),
                         static_cast<std::int64_t>(t.n_snp), pop_offsets_.get(), ploidy_.get(), t.n_pop, d_a2,
                         d_n, ld, stream_);
    mark_end(kDecode);
    sync();
}

void HipBackend::decode_counts_noalloc(const stp_tile& t, std::uint16_t* d_a2, std::uint16_t* d_n, std::int64_t ld) {
    activate();
    STP_REQUIRE(t.packed_on_device, "decode (no allocation): the tile must be resident on the device");
    STP_REQUIRE(t.n_pop >= 0 && t.n_pop <= kNoallocMaxPops,
                "decode (no allocation): at most " + std::to_string(kNoallocMaxPops) + " populations");
    STP_REQUIRE(t.n_individuals <= static_cast<std::size_t>(kNoallocMaxIndividuals),
                "decode (no allocation): at most " + std::to_string(kNoallocMaxIndividuals) + " individuals");
    STP_REQUIRE(t.n_pop == 0 || t.pop_offsets != nullptr, "decode: pop_offsets is required");
    STP_REQUIRE(t.bytes_per_record >= (t.n_snp + 3) / 4, "decode: bytes_per_record < ceil(n_snp/4)");
    STP_REQUIRE(t.n_individuals == 0 || t.packed != nullptr, "decode: packed is required");
    for (int p = 0; p < t.n_pop; ++p) {
        STP_REQUIRE(t.pop_offsets[p] <= t.pop_offsets[p + 1], "decode: pop_offsets must be non-decreasing");
        STP_REQUIRE(t.pop_offsets[p + 1] - t.pop_offsets[p] <= static_cast<std::uint64_t>(kMaxPopForU16),
                    "decode: a population exceeds 16383 individuals (uint16 count limit)");
    }
    if (t.n_pop > 0)
        STP_REQUIRE(t.pop_offsets[t.n_pop] <= t.n_individuals, "decode: pop_offsets exceed n_individuals");
    const std::int64_t n_ind = static_cast<std::int64_t>(t.n_individuals);
    if (t.n_pop > 0)
        STP_HIP(hipMemcpyAsync(noalloc_.offs, t.pop_offsets, (t.n_pop + 1) * sizeof(std::uint64_t),
                               hipMemcpyHostToDevice, stream_));
    if (t.sample_ploidy) {
        if (n_ind > 0)
            STP_HIP(hipMemcpyAsync(noalloc_.ploidy_i32, t.sample_ploidy, n_ind * sizeof(std::int32_t),
                                   hipMemcpyHostToDevice, stream_));
        launch_ploidy_from_i32(noalloc_.ploidy_i32, n_ind, noalloc_.ploidy, stream_);
    } else if (t.detect_ploidy_on_device) {
        launch_detect_ploidy(t.packed, static_cast<std::int64_t>(t.bytes_per_record),
                             static_cast<std::int64_t>(t.n_snp), n_ind, noalloc_.ploidy, stream_);
    } else {
        launch_fill_ploidy(noalloc_.ploidy, n_ind, (t.ploidy == 1 || t.ploidy == 2) ? t.ploidy : 0, stream_);
    }
    mark_begin(kDecode);
    launch_decode_counts(t.packed, static_cast<std::int64_t>(t.bytes_per_record), static_cast<std::int64_t>(t.n_snp),
                         noalloc_.offs, noalloc_.ploidy, t.n_pop, d_a2, d_n, ld, stream_);
    mark_end(kDecode);
    // the host arrays copied above belong to the caller: complete before returning
    sync();
}

void HipBackend::decode_af(const stp_tile& t, double* q, double* v, double* n) {
    activate();
    const std::int64_t M = static_cast<std::int64_t>(t.n_snp);
    const std::int64_t ld = round_up(std::max<std::int64_t>(M, 1), 64);
    a2_.ensure(static_cast<std::size_t>(t.n_pop) * ld);
    n_.ensure(static_cast<std::size_t>(t.n_pop) * ld);
    decode_counts(t, a2_.get(), n_.get(), ld);
    const std::size_t pm = static_cast<std::size_t>(t.n_pop) * static_cast<std::size_t>(M);
    if (pm == 0) return;
    qvn_.ensure(3 * pm);
    launch_counts_to_qvn(a2_.get(), n_.get(), ld, t.n_pop, M, qvn_.get(), qvn_.get() + pm, qvn_.get() + 2 * pm,
                         stream_);
    STP_HIP(hipMemcpyAsync(q, qvn_.get(), pm * sizeof(double), hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipMemcpyAsync(v, qvn_.get() + pm, pm * sizeof(double), hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipMemcpyAsync(n, qvn_.get() + 2 * pm, pm * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
}

// ---------------------------------------------------------------------------------------
void HipBackend::run_f2(const F2Plan& plan, int precision, const F2Source& src, double* d_f2, double* d_vpair,
                        F2Work& w) {
    STP_REQUIRE(precision == STP_PREC_FP64 || precision == STP_PREC_SPLIT_BF16,
                "f2: unknown precision form " + std::to_string(precision));
    const std::int64_t rows16 = static_cast<std::int64_t>(plan.G) * kRowGroup;
    const std::size_t plane = static_cast<std::size_t>(rows16) * static_cast<std::size_t>(plan.K_total);
    const std::int64_t n_chunks = plan.K_total / kKPad;
    const BlockTable bt = w.plan.table();
    const int n_items = static_cast<int>(plan.items.size());
    const F2Tune& tn = plan.tune;
    // Persistent grids are whole multiples of the die count so every die queue has
    // its own workers.
    auto grid = [&](int per_cu, std::int64_t units) {
        return static_cast<int>(
            round_up(std::max<std::int64_t>(1, std::min<std::int64_t>(static_cast<std::int64_t>(per_cu) * n_cu_,
                                                                      ceil_div(units, 4))),
                     kXcds));
    };
    // The pair-count kernel runs one 128x128 item per 4-wavefront workgroup.
    const int n_wg_count = grid(tn.count_wg_per_cu, 4 * static_cast<std::int64_t>(plan.items128.size()));
    w.planeV8.ensure(std::max<std::size_t>(plane, 16));
    // queue cursors of the two kernels, then one item word per split workgroup
    w.counters.ensure(2 * kXcds + kMaxSplitWorkgroups);
    if (precision == STP_PREC_FP64) {
        w.planeA.ensure(std::max<std::size_t>(plane, 1));
        w.planeV.ensure(std::max<std::size_t>(plane, 1));
        w.planeQ.ensure(std::max<std::size_t>(plane, 1));
        const std::int64_t KS = plan.K_total / 4;
        mark_begin(kPrep);
        if (src.a2)
            launch_prep_fp64_counts(src.a2, src.n, src.ld, plan.P, plan.G, bt, w.plan.chunk_block.get(), n_chunks,
                                    KS, w.planeA.get(), w.planeV.get(), w.planeQ.get(), w.planeV8.get(), stream_);
        else
            launch_prep_fp64_qvn(src.q, src.v, src.nq, plan.P, plan.G, bt, w.plan.chunk_block.get(), n_chunks, KS,
                                 w.planeA.get(), w.planeV.get(), w.planeQ.get(), w.planeV8.get(), stream_);
        mark_end(kPrep);
        mark_begin(kCount);
        launch_count(w.planeV8.get(), bt, w.plan.items128.get(), plan.q128, w.counters.get(), plan.P, plan.G,
                     plan.K_total / 64, d_vpair, n_wg_count, stream_);
        mark_end(kCount);
        const FixupList fl = fixup_list(plan, kFixupTauFp64, w);
        const int units = n_items * (tn.fp64_rows == 32 ? 2 : 1);
        mark_begin(kF2);
        launch_f2_fp64(w.planeA.get(), w.planeV.get(), w.planeQ.get(), bt, w.plan.items.get(), plan.q64,
                       w.counters.get() + kXcds, plan.P, plan.G, KS, d_f2, d_vpair, fl,
                       grid(tn.fp64_wg_per_cu, 4 * static_cast<std::int64_t>(units)), tn.fp64_rows,
                       tn.fp64_prefetch != 0, stream_);
        mark_end(kF2);
        run_fixup(plan, src, fl, d_f2, d_vpair, w);
        return;
    }
    // Split-BF16: 13 digit planes of 2 bytes per (row, SNP), rows padded to 32.
    const std::int64_t KS16 = plan.K_total / 16;
    const std::size_t split_bytes = static_cast<std::size_t>(plan.G32) * 32 * plan.K_total * 2 * kSplitPlanes;
    w.planeSplit.ensure(std::max<std::size_t>(split_bytes, 16));
    mark_begin(kPrep);
    if (src.a2)
        launch_prep_split_counts(src.a2, src.n, src.ld, plan.P, plan.G32, plan.G, bt, w.plan.chunk_block.get(),
                                 n_chunks, KS16, w.planeSplit.get(), w.planeV8.get(), stream_);
    else
        launch_prep_split_qvn(src.q, src.v, src.nq, plan.P, plan.G32, plan.G, bt, w.plan.chunk_block.get(), n_chunks,
                              KS16, w.planeSplit.get(), w.planeV8.get(), stream_);
    mark_end(kPrep);
    mark_begin(kCount);
    launch_count(w.planeV8.get(), bt, w.plan.items128.get(), plan.q128, w.counters.get(), plan.P, plan.G,
                 plan.K_total / 64, d_vpair, n_wg_count, stream_);
    mark_end(kCount);
    const int n_items128 = static_cast<int>(plan.items128.size());
    const FixupList fl = fixup_list(plan, kFixupTauSplit, w);
    mark_begin(kF2);
    launch_f2_split(w.planeSplit.get(), bt, w.plan.items128.get(), plan.q128, w.counters.get() + kXcds, plan.P,
                    plan.G32, KS16, d_vpair, d_f2, fl,
                    std::min(kMaxSplitWorkgroups,
                             static_cast<int>(round_up(std::max(1, std::min(n_cu_, n_items128)), kXcds))),
                    tn.split_ring, reinterpret_cast<int*>(w.counters.get() + 2 * kXcds), stream_);
    mark_end(kF2);
    run_fixup(plan, src, fl, d_f2, d_vpair, w);
}

// The flag list of one contraction: capacity of 1/16 of the upper-triangle entries
// (at least 64 Ki, at most 16 Mi entries); the count is cleared on the stream.
// STEPPE_F2_FIXUP_CAP lowers the capacity (tests use it to exercise the overflow scan).
FixupList HipBackend::fixup_list(const F2Plan& plan, double tau, F2Work& w) {
    const std::int64_t upper = static_cast<std::int64_t>(plan.P) * (plan.P + 1) / 2 * plan.n_block;
    std::int64_t cap = std::min<std::int64_t>(std::max<std::int64_t>(upper / 16, 1 << 16), 1 << 24);
    w.fix_count.ensure(1);
    w.fix_entries.ensure(static_cast<std::size_t>(3 * cap));
    if (const char* env = std::getenv("STEPPE_F2_FIXUP_CAP")) {
        const long long v = std::strtoll(env, nullptr, 10);
        if (v > 0) cap = std::min<std::int64_t>(cap, v);
    }
    launch_zero_u32(w.fix_count.get(), 1, stream_);
    FixupList fl;
    fl.count = w.fix_count.get();
    fl.entries = w.fix_entries.get();
    fl.cap = static_cast<unsigned>(cap);
    fl.tau = tau;
    return fl;
}

void HipBackend::run_fixup(const F2Plan& plan, const F2Source& src, const FixupList& fl, double* d_f2,
                           const double* d_vpair, const F2Work& w) {
    const BlockTable bt = w.plan.table();
    const int n_wg = 2 * n_cu_;
    mark_begin(kFixup);
    if (src.a2)
        launch_f2_fixup_counts(src.a2, src.n, src.ld, bt, fl, plan.P, plan.n_block, d_vpair, d_f2, n_wg, stream_);
    else
        launch_f2_fixup_qvn(src.q, src.v, src.nq, bt, fl, plan.P, plan.n_block, d_vpair, d_f2, n_wg, stream_);
    mark_end(kFixup);
}

void HipBackend::f2_from_counts(const std::uint16_t* d_a2, const std::uint16_t* d_n, std::int64_t ld, int P,
                                std::int64_t M, const std::int32_t* block_id, int n_block, int precision,
                                double* d_f2, double* d_vpair) {
    activate();
    STP_REQUIRE(ld >= M, "f2: ld must be >= M");
    const F2Plan& plan = plan_for(P, M, block_id, n_block, precision);
    F2Source src;
    src.a2 = d_a2;
    src.n = d_n;
    src.ld = ld;
    run_f2(plan, precision, src, d_f2, d_vpair);
    sync();
}

void HipBackend::f2_from_qvn(const double* Q, const double* V, const double* N, int P, std::int64_t M,
                             const std::int32_t* block_id, int n_block, int precision, double* f2,
                             double* vpair, std::int32_t* block_sizes) {
    activate();
    const F2Plan& plan = plan_for(P, M, block_id, n_block, precision);
    const std::size_t pm = static_cast<std::size_t>(P) * static_cast<std::size_t>(M);
    qvn_.ensure(std::max<std::size_t>(3 * pm, 1));
    if (pm) {
        STP_HIP(hipMemcpyAsync(qvn_.get(), Q, pm * sizeof(double), hipMemcpyHostToDevice, stream_));
        STP_HIP(hipMemcpyAsync(qvn_.get() + pm, V, pm * sizeof(double), hipMemcpyHostToDevice, stream_));
        STP_HIP(hipMemcpyAsync(qvn_.get() + 2 * pm, N, pm * sizeof(double), hipMemcpyHostToDevice, stream_));
    }
    const std::size_t tot = static_cast<std::size_t>(P) * P * n_block;
    DevBuf<double> df2(tot), dvp(tot);
    const double* q = qvn_.get();
    F2Source src;
    src.q = q;
    src.v = q + pm;
    src.nq = q + 2 * pm;
    run_f2(plan, precision, src, df2.get(), dvp.get());
    STP_HIP(hipMemcpyAsync(f2, df2.get(), tot * sizeof(double), hipMemcpyDeviceToHost, stream_));
    STP_HIP(hipMemcpyAsync(vpair, dvp.get(), tot * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    for (int b = 0; b < n_block; ++b) block_sizes[b] = plan.len[b];
}

F2Blocks* HipBackend::extract_f2(const stp_tile& t, const std::int32_t* block_id, int n_block, int precision,
                                  const TierRequest& tier) {
    activate();
    const std::int64_t M = static_cast<std::int64_t>(t.n_snp);
    STP_REQUIRE(t.n_pop > 0, "extract_f2: at least one population is required");
    (void)make_f2_plan(t.n_pop, M, block_id, n_block);   // validate the partition before decoding
    const std::int64_t ld = round_up(std::max<std::int64_t>(M, 1), 64);
    a2_.ensure(static_cast<std::size_t>(t.n_pop) * ld);
    n_.ensure(static_cast<std::size_t>(t.n_pop) * ld);
    prepare_tile(t);
    mark_begin(kDecode);
    launch_decode_counts(d_packed_, static_cast<std::int64_t>(t.bytes_per_record), M, pop_offsets_.get(),
                         ploidy_.get(), t.n_pop, a2_.get(), n_.get(), ld, stream_);
    mark_end(kDecode);
    return f2_counts_tiered(a2_.get(), n_.get(), ld, t.n_pop, M, block_id, n_block, precision, tier);
}

F2Blocks* HipBackend::upload(const double* f2, const double* vpair, const std::int32_t* block_sizes, int P,
                             int n_block) {
    activate();
    STP_REQUIRE(P > 0 && n_block > 0, "upload: P and n_block must be positive");
    STP_REQUIRE(f2 != nullptr && block_sizes != nullptr, "upload: f2 and block_sizes are required");
    auto out = std::make_unique<F2Blocks>();
    out->P = P;
    out->n_block = n_block;
    const std::size_t tot = static_cast<std::size_t>(P) * P * n_block;
    out->f2.alloc(tot);
    out->vpair.alloc(tot);
    out->block_sizes.assign(block_sizes, block_sizes + n_block);
    STP_HIP(hipMemcpyAsync(out->f2.get(), f2, tot * sizeof(double), hipMemcpyHostToDevice, stream_));
    if (vpair) {
        STP_HIP(hipMemcpyAsync(out->vpair.get(), vpair, tot * sizeof(double), hipMemcpyHostToDevice, stream_));
        sync();
    } else {
        std::vector<double> ones(tot, 1.0);
        STP_HIP(hipMemcpyAsync(out->vpair.get(), ones.data(), tot * sizeof(double), hipMemcpyHostToDevice, stream_));
        sync();
    }
    return out.release();
}

void HipBackend::synth_panel(std::uint64_t seed, int P, std::int64_t M, std::int64_t n_ind, std::int64_t bpr,
                             const float* fst, const float* pop_missing, const std::int32_t* ind_pop,
                             const std::int8_t* ind_haploid, std::uint8_t* d_packed) {
    activate();
    STP_REQUIRE(P > 0 && M > 0 && n_ind > 0, "synth: P, M and n_ind must be positive");
    for (std::int64_t g = 0; g < n_ind; ++g)
        STP_REQUIRE(ind_pop[g] >= 0 && ind_pop[g] < P, "synth: ind_pop out of range");
    DevBuf<float> dfst(P), dmiss(P);
    DevBuf<std::int32_t> dpop(n_ind);
    DevBuf<std::int8_t> dhap(n_ind);
    STP_HIP(hipMemcpyAsync(dfst.get(), fst, P * sizeof(float), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(dmiss.get(), pop_missing, P * sizeof(float), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(dpop.get(), ind_pop, n_ind * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_));
    STP_HIP(hipMemcpyAsync(dhap.get(), ind_haploid, n_ind, hipMemcpyHostToDevice, stream_));
    synth_freq_.ensure(static_cast<std::size_t>(P) * M);
    launch_synth_panel(seed, P, M, n_ind, bpr, dfst.get(), dmiss.get(), dpop.get(), dhap.get(), synth_freq_.get(),
                       d_packed, stream_);
    sync();
    synth_freq_.release();
}

}  // namespace stp
