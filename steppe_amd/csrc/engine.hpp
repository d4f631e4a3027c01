// steppe_amd/csrc/engine.hpp — the host-side engine behind the C ABI.
//
// HipBackend mirrors the part of the reference's ComputeBackend seam
// (src/device/backend.hpp:673-1656) that the f-statistics hot path uses:
// decode_af / detect_sample_ploidy_device / compute_f2_blocks(_device), and owns
// the per-device resources the reference's CudaBackend owns (one stream, pooled
// device workspaces; src/device/cuda/cuda_backend.cuh:576-602).  F2Blocks mirrors
// DeviceF2Blocks (src/device/device_f2_blocks.hpp:18-53): a move-only owner of the
// resident [P x P x n_block] f2 and vpair tensors.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "f2_layout.hpp"
#include "frontend.hpp"
#include "fstats.hpp"
#include "geno_io.hpp"
#include "qpadm.hpp"
#include "qpgraph.hpp"
#include "qpgraph_enum.hpp"
#include "steppe_hip.h"

namespace stp {

// ---- launchers (defined in the .hip translation units) -----------------------------
void launch_detect_ploidy(const std::uint8_t* packed, std::int64_t bpr, std::int64_t n_snp,
                          std::int64_t n_ind, std::int8_t* out, hipStream_t st);
void launch_fill_ploidy(std::int8_t* out, std::int64_t n, int value, hipStream_t st);
void launch_decode_counts(const std::uint8_t* packed, std::int64_t bpr, std::int64_t n_snp,
                          const std::uint64_t* pop_offsets, const std::int8_t* ploidy, int n_pop,
                          std::uint16_t* a2, std::uint16_t* n, std::int64_t ld, hipStream_t st);
void launch_counts_to_qvn(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                          std::int64_t M, double* q, double* v, double* nout, hipStream_t st);
void launch_zero_u32(unsigned* p, int n, hipStream_t st);
void launch_prep_fp64_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                             int G, BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks,
                             std::int64_t KS, double* pA, double* pV, double* pQ, void* pV8,
                             hipStream_t st);
void launch_prep_fp64_qvn(const double* q, const double* v, const double* n, int P, int G, BlockTable bt,
                          const std::int32_t* chunk_block, std::int64_t n_chunks, std::int64_t KS,
                          double* pA, double* pV, double* pQ, void* pV8, hipStream_t st);
void launch_f2_fp64(const double* pA, const double* pV, const double* pQ, BlockTable bt, const F2Item* items,
                    const ItemQueues& queues, unsigned* counters, int P, int G, std::int64_t KS, double* f2,
                    const double* vpair, const FixupList& fl, int n_workgroups, int wave_rows, bool prefetch,
                    hipStream_t st);
void launch_prep_split_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, int G32,
                              int G16, BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks,
                              std::int64_t KS16, std::uint8_t* planes, void* pV8, hipStream_t st);
void launch_prep_split_qvn(const double* q, const double* v, const double* n, int P, int G32, int G16,
                           BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks, std::int64_t KS16,
                           std::uint8_t* planes, void* pV8, hipStream_t st);
void launch_count(const void* pV8, BlockTable bt, const F2Item* items, const ItemQueues& queues,
                  unsigned* counters, int P, int G16, std::int64_t KS64, double* vpair, int n_workgroups,
                  hipStream_t st);
void launch_f2_split(const std::uint8_t* planes, BlockTable bt, const F2Item* items, const ItemQueues& queues,
                     unsigned* counters, int P, int G32, std::int64_t KS16, const double* vpair, double* f2,
                     const FixupList& fl, int n_workgroups, int ring, int* item_slot, hipStream_t st);
void launch_f2_fixup_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, BlockTable bt,
                            const FixupList& fl, int P, int n_block, const double* vpair, double* f2,
                            int n_workgroups, hipStream_t st);
void launch_f2_fixup_qvn(const double* q, const double* v, const double* n, BlockTable bt, const FixupList& fl,
                         int P, int n_block, const double* vpair, double* f2, int n_workgroups, hipStream_t st);
void launch_synth_panel(std::uint64_t seed, int P, std::int64_t M, std::int64_t n_ind, std::int64_t bpr,
                        const float* fst, const float* pop_missing, const std::int32_t* ind_pop,
                        const std::int8_t* ind_haploid, float* freq_scratch, std::uint8_t* packed,
                        hipStream_t st);

// Largest population (individuals) the 16-bit counts hold: a2 <= 4 * size fits uint16.
constexpr int kMaxPopForU16 = 16383;

// The allocation-free decode's fixed device scratch (decode.hip): at most this many
// populations and individuals per call.
constexpr int kNoallocMaxPops = 65536;
constexpr std::int64_t kNoallocMaxIndividuals = std::int64_t{1} << 20;
struct NoallocScratch {
    std::uint64_t* offs = nullptr;       // population offsets (device address)
    std::int8_t* ploidy = nullptr;       // per-sample ploidy
    std::int32_t* ploidy_i32 = nullptr;  // staged caller ploidy values
    std::size_t bytes = 0;
};
// This device's address of the scratch array (the device must be current).
NoallocScratch resolve_noalloc_scratch();
void launch_ploidy_from_i32(const std::int32_t* in, std::int64_t n, std::int8_t* out, hipStream_t st);

// Upper bound on the split contraction's grid (one workgroup per CU; the per-workgroup
// item words of its second ring layout live after the queue cursors).
constexpr int kMaxSplitWorkgroups = 4096;

// ---- launch configuration chosen by the autotuner (autotune.cpp) -----------------------
struct F2Tune {
    int sched = 0;            // work items: 0 per-die queues (blocks longest-first onto the
                              // least-loaded die), 1 one shared queue (longest blocks first)
    int fp64_rows = 64;       // FP64 wavefront tile: 64 x 64 or 32 x 64 outputs
    int fp64_prefetch = 1;    // FP64: load the next K-step while this one's MFMAs run
    int fp64_wg_per_cu = 2;   // FP64 contraction workgroups (4 wavefronts) per CU
    int count_wg_per_cu = 2;  // pair-count kernel workgroups per CU
    int split_ring = 0;       // split contraction LDS ring: 0 (A, V double-buffered, 152 KB) or
                              // 1 (A, V, B double-buffered, 160 KB)
    std::string str() const;
    static bool parse(const std::string& s, F2Tune& out);
};

// ---- block plan ----------------------------------------------------------------------
// Host description of one f2 launch: block ranges, padded K layout, LPT work items.
struct F2Plan {
    int P = 0, G = 0, T = 0, n_block = 0;
    std::int64_t M = 0, K_total = 0;
    std::vector<std::int64_t> s0, k0;
    std::vector<std::int32_t> len, kpad, chunk_block;
    std::vector<F2Item> items;     // 64-row tile pairs (FP64 contraction, pair counts)
    std::vector<F2Item> items128;  // 128-row tile pairs (split contraction)
    ItemQueues q64{}, q128{};      // per-die segments of items / items128
    int G32 = 0, T128 = 0;
    F2Tune tune;                   // launch configuration (its sched built the queues)
};

// Validates block_id (contiguous, non-decreasing, in [0, n_block); the reference's
// core::block_ranges throws on the same conditions, block_partition_rule.hpp:54-96).
F2Plan make_f2_plan(int P, std::int64_t M, const std::int32_t* block_id, int n_block, const F2Tune& tune = F2Tune{});

// Device workspace of one f2 contraction: the block plan, the operand planes, the work
// queue cursors and the fix-up list.  The backend keeps one for its eager calls; a
// captured f2 graph owns its own, so that later calls cannot move the buffers its
// kernels were captured with.
struct F2Work {
    struct PlanDev {
        DevBuf<std::int64_t> s0, k0;
        DevBuf<std::int32_t> len, kpad, chunk_block;
        DevBuf<F2Item> items, items128;
        BlockTable table() const;
    };
    PlanDev plan;
    DevBuf<double> planeA, planeV, planeQ;      // FP64 operand planes
    DevBuf<std::uint8_t> planeV8;               // int8 V plane (pair counts)
    DevBuf<std::uint8_t> planeSplit;            // split-BF16 digit planes
    DevBuf<unsigned> counters;                  // work-queue cursors: pair counts, contraction
    DevBuf<unsigned> fix_count;                 // cancellation-prone entries flagged
    DevBuf<std::int32_t> fix_entries;
};

// A resident extraction of one fixed shape captured as a HIP graph: ploidy detection,
// decode, operand prep, pair counts, contraction and fix-up of a device-resident tile
// into caller-owned f2 / vpair buffers, replayed with one launch.  The graph owns its
// workspace (plan, planes, queue cursors, fix-up list, population offsets, ploidy);
// the caller keeps the tile's packed records and the output buffers alive.
struct F2Graph {
    int precision = 0;
    int P = 0;
    std::int64_t M = 0;
    std::size_t bpr = 0, n_ind = 0;
    const std::uint8_t* packed = nullptr;
    int ploidy_mode = 0;                     // 0 detect on device, 1 given per sample, 2 fixed
    int fixed_ploidy = 0;
    DevBuf<std::uint64_t> pop_offsets;
    DevBuf<std::int8_t> ploidy;
    std::uint16_t* a2 = nullptr;
    std::uint16_t* n = nullptr;
    std::int64_t ld = 0;
    DevBuf<std::uint16_t> own_a2, own_n;
    double* f2 = nullptr;
    double* vpair = nullptr;
    F2Plan plan;
    F2Work work;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool timed = false;                      // the stage events are part of the graph
    int device = 0;
    ~F2Graph();
};

// Where an f2 result lives (the reference's OutputTier, src/device/tier_select.hpp:12-17):
// in HBM, in host memory, or in an STPF2BK1 f2.bin file written block range by block range.
enum F2Tier { kTierResident = 0, kTierHost = 1, kTierDisk = 2 };

// Host memory of the host tier: 2 MiB aligned and advised for transparent huge pages (a
// result of many GB would otherwise spend most of its copy time in 4 KiB page faults);
// left uninitialised, since the streamed contraction writes every block.
class HostArray {
public:
    void alloc(std::size_t n);
    double* get() const { return p_.get(); }
    std::size_t size() const { return n_; }

private:
    struct Free {
        void operator()(double* p) const;
    };
    std::unique_ptr<double, Free> p_;
    std::size_t n_ = 0;
};

class F2Blocks {
public:
    int P = 0, n_block = 0;
    int tier = kTierResident;
    DevBuf<double> f2, vpair;                 // resident tier
    HostArray host_f2, host_vpair;            // host tier
    std::string disk_path;                    // disk tier
    std::vector<std::int32_t> block_sizes;
    std::vector<std::string> labels;   // population labels when built from files
};

// Device bytes the contraction of `plan` needs besides its result (operand planes of the
// precision, the int8 V plane, the fix-up list), and the automatic tier choice from the
// result size, that working set and the free device / host memory (tiered.cpp).
std::size_t f2_working_set_bytes(const F2Plan& plan, int precision);
int choose_f2_tier(std::size_t result_bytes, std::size_t working_bytes, std::size_t free_vram, std::size_t free_host);

// Tier of an f2 computation: -1 chooses from free memory as the reference's
// select_output_tier does (src/device/tier_select.hpp:57-80: resident when the result
// fits 70% of free HBM, else host when it fits 60% of free host memory, else disk);
// chunk_blocks (0: from free HBM) bounds the blocks computed per streamed chunk.
struct TierRequest {
    int tier = -1;
    std::string disk_path;
    std::int64_t chunk_blocks = 0;
};

// Genotype files -> f2 (the reference's run_extract_f2 arguments,
// src/app/extract_f2_core.cpp:96-104 and bindings/bind_fstats.cpp:131-186).
struct ExtractRequest {
    std::string geno, snp, ind;
    PopSelect pops;
    double maf_min = 0.0, maxmiss = 1.0;
    bool autosomes_only = false, drop_monomorphic = false, transversions_only = false;
    int strand_mode = 0;          // 0 drop strand-ambiguous SNPs, 1 keep, 2 flip (= keep)
    TierRequest tier;
    double blgsize = 0.05;        // jackknife block size in Morgans
    int ploidy = 0;               // 0 detect per sample, 1 pseudo-haploid, 2 diploid
    int precision = 0;
};
struct ExtractInfo {
    std::int64_t n_snp_total = 0, n_snp_kept = 0, n_pseudo_haploid = 0, n_diploid = 0;
};

// A resident set of qpGraph graphs: topology arena plus each graph's f3 basis (f_obs),
// unridged Q diagonal and Qinv on the device, computed once; evaluations then move only
// theta (and edge lengths) up and scores / gradients down.
struct QgSet {
    int n = 0, np_max = 0, ne_max = 0, npop_max = 0, npath_max = 0, dmax = 0;
    std::vector<QgTopology> topo;
    std::vector<std::int32_t> slot;   // graph -> its basis slot in fobs / qdiag / qinv
    std::vector<std::int32_t> host_status;
    DevBuf<QgTopoDev> d_topo;
    DevBuf<std::int32_t> d_iarena, status;
    DevBuf<double> d_darena, fobs, qdiag, qinv;
    DevBuf<std::int32_t> i32;
    DevBuf<std::int64_t> i64;
    DevBuf<double> in, out;
};

// An all-combination f-statistic sweep request (the reference's SweepRequest,
// include/steppe/fstat_sweep.hpp:25-38): k = 4 (f4) or 3 (f3) over the C(range, k)
// combinations of the population range (every population, or the subset's).
struct FstatSweepRequest {
    int k = 4;
    std::vector<std::int32_t> subset;
    int filter = 0;          // 0: |z| >= min_z; 1: the top_k largest |z| (also >= min_z)
    double min_z = 3.0;
    std::int64_t top_k = 100;
    bool sure = false;       // lift the 1e8-combination cap
    int method = 0;          // 0: automatic, 1: direct per statistic, 2: Gram (f4 only)
};

// Survivors ordered by |z| descending then combination rank; keys hold 4 population
// indices per row (the fourth 0 for f3).  status: 0 ok, 5 invalid_config (capped).
struct FstatSweepOut {
    std::vector<std::int32_t> keys;
    std::vector<double> est, se, z, p;
    std::vector<std::int64_t> rank;
    std::uint64_t enumerated = 0;
    bool capped = false;
    int status = 0;
    std::int64_t rechecked = 0;   // Gram method: statistics recomputed directly
};

// Options of the device qpGraph fit (projected L-BFGS per instance).
struct QgFitOptions {
    int maxit = 200;
    double gtol = 1e-11, xtol = 1e-12, ftol = 1e-14;
    double fudge = 1e-4;
    int constrained = 1;
};

// Per-graph results of a multistart fit of every graph of a set (row strides: dmax for
// the theta arrays, ne_max for edge_length, np_max for f3_fit).
struct QgFleetOut {
    std::vector<double> score, restart_spread, theta, theta_lo, theta_hi, edge_length, f3_fit;
    std::vector<std::int32_t> status, iterations;
};

// Options of the topology search (the reference's QpGraphSearchOptions,
// include/steppe/qpgraph_search.hpp:30-36, with the fit options of QpGraphOptions).
struct QgSearchOptions {
    int max_nadmix = 1;          // 0: trees only; larger values act as 1 (the bounded space)
    int numstart = 10;
    double diag_f3 = 1e-5;
    QgFitOptions fit;
    bool run_heuristic = true;
    int heuristic_seeds = 8;
};

// Result of the topology search (the reference's QpGraphSearchResult,
// include/steppe/qpgraph_search.hpp:48-72).  cands is the whole enumeration; fitted[k]
// is the candidate index of fitted graph k, whose best score and restart spread are
// score[k] and restart_spread[k]; best indexes fitted graphs.  status: 0 ok, 3
// non_spd_covariance (the shared basis covariance or every fit failed).
struct QgSearchResult {
    int n_trees = 0, n_admix1 = 0, n_candidates = 0;
    std::vector<qgenum::Candidate> cands;
    std::vector<int> fitted;
    std::vector<double> score, restart_spread;
    int best = -1;
    double best_score = 0.0, second_best_score = 0.0;
    QgTopology best_topo;
    int best_fit_status = 3;
    std::vector<double> theta, theta_lo, theta_hi, edge_length, f3_fit, f_obs, qdiag;
    bool heuristic_recovered = false;
    std::vector<std::uint64_t> seed_hashes;
    double fit_all_wall_ms = 0.0, topologies_per_s = 0.0;
    int status = 0;
};

class HipBackend {
public:
    explicit HipBackend(int device);
    ~HipBackend();
    HipBackend(const HipBackend&) = delete;
    HipBackend& operator=(const HipBackend&) = delete;

    int device() const { return device_; }
    int n_cu() const { return n_cu_; }
    const std::string& name() const { return name_; }
    hipStream_t stream() const { return stream_; }
    void sync();
    void activate() const;

    std::vector<int> detect_ploidy(const stp_tile& t);
    // Decode into caller-owned device buffers.
    void decode_counts(const stp_tile& t, std::uint16_t* d_a2, std::uint16_t* d_n, std::int64_t ld);
    // The same with no device allocation: population offsets and ploidy go through the
    // fixed scratch array resolved at construction (device-resident tiles only).
    void decode_counts_noalloc(const stp_tile& t, std::uint16_t* d_a2, std::uint16_t* d_n, std::int64_t ld);
    void decode_af(const stp_tile& t, double* q, double* v, double* n);
    // f2 over device counts into caller-owned device outputs.
    void f2_from_counts(const std::uint16_t* d_a2, const std::uint16_t* d_n, std::int64_t ld, int P,
                        std::int64_t M, const std::int32_t* block_id, int n_block, int precision,
                        double* d_f2, double* d_vpair);
    // f2 over host Q/V/N into host outputs (the reference's compute_f2_blocks seam).
    void f2_from_qvn(const double* Q, const double* V, const double* N, int P, std::int64_t M,
                     const std::int32_t* block_id, int n_block, int precision, double* f2,
                     double* vpair, std::int32_t* block_sizes);
    F2Blocks* extract_f2(const stp_tile& t, const std::int32_t* block_id, int n_block, int precision,
                         const TierRequest& tier = TierRequest{kTierResident, "", 0});
    // Capture / replay of a fixed-shape resident extraction (f2_graph.cpp).  a2 / n may be
    // null (the graph then holds its own counts).
    F2Graph* f2_graph_create(const stp_tile& t, const std::int32_t* block_id, int n_block, int precision,
                             std::uint16_t* a2, std::uint16_t* n, std::int64_t ld, double* f2, double* vpair);
    void f2_graph_launch(F2Graph& g);
    // Autotuner (autotune.cpp): times every candidate launch configuration of the f2
    // contraction on these device inputs and outputs, keeps the fastest for the workload's
    // key (device, precision, P, block-size histogram) and writes the table to the tune
    // cache file.  Returns the sweep as JSON.
    std::string f2_autotune(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, std::int64_t M,
                            const std::int32_t* block_id, int n_block, int precision, double* f2, double* vpair);
    // The configuration for a plan's key: the tuned one, else the defaults.
    F2Tune tune_for(const F2Plan& plan, int precision);
    // Genotype files -> resident f2 (frontend_host.cpp): read the packed triple, build the
    // individual-major tile in HBM, decode, keep / compact SNPs, block, contract.
    F2Blocks* extract_f2_files(const ExtractRequest& req, ExtractInfo& info);
    // Host copies of any tier's tensors (f2 / vpair may be null).
    void download(const F2Blocks& h, double* f2, double* vpair);
    F2Blocks* upload(const double* f2, const double* vpair, const std::int32_t* block_sizes, int P,
                     int n_block);
    // Batched qpAdm over models on a resident tensor; results to host (qpadm_host.cpp).
    void qpadm_batch(const F2Blocks& h, int n_models, const std::int32_t* target, const std::int32_t* left_off,
                     const std::int32_t* left, const std::int32_t* right_off, const std::int32_t* right,
                     const stp_qpadm_options& opt, stp_qpadm_result* out);

    // Standalone f-statistics on a resident tensor (fstats_host.cpp): f4 (width 4) or
    // f3 (width 3) over n index tuples, the f4-ratio over n 5-tuples, and the
    // all-combination sweeps.  Results go to host arrays.
    void fstat(const F2Blocks& h, int width, std::int64_t n, const std::int32_t* idx, double* est, double* se,
               double* z, double* p);
    void f4ratio(const F2Blocks& h, std::int64_t n, const std::int32_t* tuples, double* alpha, double* se, double* z);
    void fstat_sweep(const F2Blocks& h, const FstatSweepRequest& req, FstatSweepOut& out);

    // qpGraph (qpgraph_host.cpp): build a resident set from topologies and f2 handles
    // (the f3 basis and covariance run on the device), evaluate scores / adjoints.
    QgSet* qpgraph_set_create(const std::vector<const QgTopology*>& topos, const F2Blocks* const* f2, double diag_f3);
    // A set from a given f3 basis and Qinv per graph (concatenated in graph order:
    // npair values, then npair x npair row-major), as the reference's fleet fit receives.
    QgSet* qpgraph_set_from_basis(const std::vector<const QgTopology*>& topos, const double* f_obs,
                                  const double* qinv);
    // Graphs that all use one f3 basis (one npair for all): one f_obs / Qinv / Q diagonal.
    QgSet* qpgraph_set_shared_basis(const std::vector<const QgTopology*>& topos, const double* f_obs,
                                    const double* qinv, const double* qdiag);
    // The bounded topology search over pops (qpgraph_search.cpp); names are the f2 labels.
    QgSearchResult qpgraph_search(const F2Blocks& f2, const std::vector<std::string>& names,
                                  const std::vector<std::string>& pops, const QgSearchOptions& o);
    void qpgraph_set_eval(QgSet& s, int n_inst, const std::int32_t* inst_graph, const double* theta,
                          const double* bl_in, int mode, double fudge, int constrained, double* score,
                          double* grad_theta, double* grad_bl, double* bl_out, double* fit_out);
    void qpgraph_set_basis(const QgSet& s, int g, double* f_obs, double* qdiag, double* qinv);
    // Device L-BFGS fits of instances from given start points (theta0[i*dmax ..]).
    void qpgraph_set_fit(QgSet& s, int n_inst, const std::int32_t* inst_graph, const double* theta0,
                         const QgFitOptions& o, double* theta_out, double* score_out, std::int32_t* iters_out);
    // numstart restarts per graph from the reference's start points, best per graph,
    // then its edge lengths and fitted f3 (the reference's qpgraph_fit_fleet per graph).
    void qpgraph_set_fit_fleet(QgSet& s, int numstart, const QgFitOptions& o, QgFleetOut& out);

    // Per-stage device timing with HIP events on this backend's stream (opt-in).
    enum Stage {
        kDecode = 0, kPrep = 1, kF2 = 2, kCount = 3, kQpadm = 4, kQpgraph = 5, kFstat = 6, kFront = 7, kFixup = 8,
        kStages = 9
    };
    void set_timing(bool on);
    void read_timing(double* ms, std::int64_t* count, bool reset);

    void synth_panel(std::uint64_t seed, int P, std::int64_t M, std::int64_t n_ind, std::int64_t bpr,
                     const float* fst, const float* pop_missing, const std::int32_t* ind_pop,
                     const std::int8_t* ind_haploid, std::uint8_t* d_packed);

private:
    // Surviving blocks of a resident tensor (blocks whose pairs are not partly missing),
    // their sizes and total; then the survivor-transposed tensor in qpT_.
    void survivor_list(const F2Blocks& h, std::vector<std::int32_t>& surv, std::vector<std::int32_t>& bsz,
                       double& n_total);
    FstatTensor survivor_tensor(const F2Blocks& h, std::vector<std::int32_t>* bsz_out = nullptr);
    // The Gram form of an f4 sweep (fstats_host.cpp); appends |z| >= tau survivors.
    void f4_sweep_gram(const FstatTensor& t, const std::vector<std::int32_t>& bsz, int range,
                       const std::int32_t* d_subset, const std::vector<std::int64_t>& binom, std::int64_t K,
                       int filter, double& tau, std::vector<SweepHit>& res, std::int64_t& rechecked);
    void qpgraph_upload_topologies(QgSet& s, const std::vector<const QgTopology*>& topos);
    void upload_plan(const F2Plan& plan, F2Work::PlanDev& pd);
    // The plan for (P, M, block_id, n_block), rebuilt and re-uploaded only when the
    // partition differs from the previous call's (compared element by element).
    const F2Plan& plan_for(int P, std::int64_t M, const std::int32_t* block_id, int n_block, int precision);
    F2Plan plan_;
    std::vector<std::int32_t> plan_bid_;
    bool plan_valid_ = false;
    // Tuned configurations by key, loaded from / saved to the tune cache file.
    std::string tune_key(const F2Plan& plan, int precision) const;
    void tune_load();
    void tune_save() const;
    std::map<std::string, F2Tune> tuned_;
    bool tune_loaded_ = false;
    void prepare_tile(const stp_tile& t);
    // The launches of one graph replay, on this backend's stream (captured or eager).
    void record_f2_graph(F2Graph& g);
    // Where the f2 prep reads frequencies from: device counts (a2, n, ld) or device
    // FP64 Q/V/N [P x M].
    struct F2Source {
        const std::uint16_t* a2 = nullptr;
        const std::uint16_t* n = nullptr;
        std::int64_t ld = 0;
        const double* q = nullptr;
        const double* v = nullptr;
        const double* nq = nullptr;
    };
    void run_f2(const F2Plan& plan, int precision, const F2Source& src, double* d_f2, double* d_vpair) {
        run_f2(plan, precision, src, d_f2, d_vpair, work_);
    }
    // The contraction on a given workspace (whose plan must already hold `plan`).
    void run_f2(const F2Plan& plan, int precision, const F2Source& src, double* d_f2, double* d_vpair, F2Work& w);
    FixupList fixup_list(const F2Plan& plan, double tau, F2Work& w);
    void run_fixup(const F2Plan& plan, const F2Source& src, const FixupList& fl, double* d_f2, const double* d_vpair,
                   const F2Work& w);
    // f2 over device counts into the requested tier (tiered.cpp): the whole tensor in HBM,
    // or block ranges computed into a HBM slab and streamed to host memory or an f2.bin.
    F2Blocks* f2_counts_tiered(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                               std::int64_t M, const std::int32_t* block_id, int n_block, int precision,
                               const TierRequest& tier);

    int device_ = 0;
    int n_cu_ = 0;
    NoallocScratch noalloc_;   // this device's address of the allocation-free decode's scratch
    std::string name_;
    hipStream_t stream_ = nullptr;

    // Pooled workspaces (grow-only).
    DevBuf<std::uint8_t> packed_stage_;
    DevBuf<std::uint64_t> pop_offsets_;
    DevBuf<std::int8_t> ploidy_;
    DevBuf<std::uint16_t> a2_, n_;
    DevBuf<unsigned> counter_;
    DevBuf<double> qvn_;
    DevBuf<float> synth_freq_;
    F2Work work_;

    // Stage timing: events bracket each stage launch; elapsed times are folded
    // into the totals at the next sync.
    void mark_begin(Stage s);
    void mark_end(Stage s);
    void collect_timing();
    bool timing_ = false;
    // qpAdm working set (grow-only): survivor-transposed f2, survivor sizes and list,
    // the model arena, per-workgroup scratch and the device result records.
    DevBuf<double> qpT_, qp_scratch_;
    DevBuf<std::int32_t> qp_i32_, qp_keep_;
    DevBuf<stp_qpadm_result> qp_out_;
    // Large-model qpAdm working set (grow-only): model and task lists, fit tables,
    // per-model setup scratch, per-wavefront fit scratch, leave-one-block-out weights.
    void qpadm_large(QpLargeArgs a, const std::vector<QpLargeModel>& large, int rank_opt);
    DevBuf<QpLargeModel> ql_models_;
    DevBuf<QpLargeTask> ql_tasks_;
    DevBuf<QpLargeTables> ql_tables_;
    DevBuf<double> ql_setup_, ql_fit_, ql_wm_;
    // f-statistics working set (grow-only): survivor sizes / index tuples, outputs,
    // sweep combinatorics and hit buffers.
    DevBuf<std::int32_t> fs_i32_, fs_idx_;
    DevBuf<double> fs_out_;
    DevBuf<std::int64_t> fs_i64_;
    DevBuf<SweepHit> fs_hits_;
    DevBuf<unsigned long long> fs_count_;
    DevBuf<double> fs_mean_, fs_wroot_;
    DevBuf<GramItem> fs_items_;
    DevBuf<std::int64_t> fs_recheck_;
    // Front-end working set: the canonical tile, a GENO staging block, the selection,
    // SNP metadata, keep flags and kept columns, and the compacted counts.
    DevBuf<std::uint8_t> fe_tile_, fe_stage_, fe_flags_;
    DevBuf<std::int32_t> fe_i32_;
    DevBuf<char> fe_alleles_;
    DevBuf<std::uint16_t> fe_a2_, fe_n_;
    hipEvent_t ev_begin_[kStages] = {}, ev_end_[kStages] = {};
    bool pending_[kStages] = {};
    double total_ms_[kStages] = {};
    std::int64_t count_[kStages] = {};

    // The tile currently staged on the device.
    const std::uint8_t* d_packed_ = nullptr;
};

}  // namespace stp
