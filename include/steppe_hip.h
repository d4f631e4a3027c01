/* steppe_hip.h — the C ABI of the MI355X f-statistics engine (libsteppe_hip.so).
 *
 * This is the drop-in boundary for the reference's f-statistics hot path: genotype
 * decode -> per-population counts -> per-jackknife-block f2 tensor held on the GPU
 * (an "f2 handle") -> fits.  Each entry point names the reference interface it
 * replaces (paths relative to the reference repository root).  Plain pointers and
 * sizes only; "d_" pointers are device (HBM) pointers, all others host pointers.
 *
 * Conventions
 *   - Every function returns STP_OK (0) or a negative status; stp_last_error()
 *     returns the calling thread's last error message (like the reference's
 *     exceptions, src/device/cuda/check.cuh, surfaced as ValueError in Python).
 *   - Calls are synchronous: a function returns after its device work finished,
 *     as the reference's CudaBackend syncs its stream before returning
 *     (src/device/cuda/cuda_backend.cuh:576-602).  A context is used by one
 *     thread at a time.
 *   - f2 tensors are [P x P x n_block] float64 at i + P*j + P*P*b
 *     (include/steppe/fstats.hpp:19-60); Q/V/N matrices are [P x M] at i + P*s.
 */
#ifndef STEPPE_HIP_H
#define STEPPE_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "stp_qpadm_result.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define STP_OK 0
#define STP_ERR_INVALID (-1)  /* invalid argument (reference: std::invalid_argument)  */
#define STP_ERR_DEVICE (-2)   /* HIP runtime / kernel failure (reference: CudaError)   */
#define STP_ERR_OOM (-3)      /* device allocation failed (reference: Status::DeviceOom) */
#define STP_ERR_IO (-4)       /* file could not be opened / read / written             */
#define STP_ERR_FORMAT (-5)   /* not a valid STPF2BK1 file                              */
#define STP_ERR_INTERNAL (-9)

/* f2 arithmetic forms (reference: steppe::Precision, include/steppe/config.hpp:74-110). */
#define STP_PREC_FP64 0        /* FP64 MFMA contraction                                 */
#define STP_PREC_SPLIT_BF16 1  /* fixed-point BF16 slices on BF16 MFMA, FP64 recombination */

typedef struct stp_context stp_context;
typedef struct stp_f2_handle stp_f2_handle;

/* Packed genotype tile (reference: DecodeTileView, src/device/backend.hpp:214-235).
 * Individual-major records of bytes_per_record bytes, 4 codes per byte, first SNP in
 * the two most significant bits, code = reference-allele count, 3 = missing.
 * Individuals of population p are [pop_offsets[p], pop_offsets[p+1]) (host array).
 * Ploidy: sample_ploidy (host, n_individuals entries) if non-NULL, else on-device
 * detection if detect_ploidy_on_device, else the scalar ploidy for every sample.
 * packed is a device pointer when packed_on_device is non-zero, else a host pointer. */
typedef struct stp_tile {
    const uint8_t* packed;
    size_t bytes_per_record;
    size_t n_snp;
    size_t n_individuals;
    const uint64_t* pop_offsets;
    int32_t n_pop;
    const int32_t* sample_ploidy;
    int32_t ploidy;
    int32_t detect_ploidy_on_device;
    int32_t packed_on_device;
} stp_tile;

/* ---- context ------------------------------------------------------------------ */

/* Open the engine on one HIP device (reference: build_resources + make_cuda_backend,
 * src/device/resources.hpp:70, src/device/backend_factory.hpp:18-24). */
int stp_open(int device, stp_context** out);
int stp_close(stp_context* ctx);
const char* stp_last_error(void);
/* Name and compute-unit count of the context's device (reference: BackendCapabilities,
 * src/device/backend.hpp:1649). */
int stp_device_info(stp_context* ctx, char* name, int name_len, int32_t* n_cu);

/* ---- decode ------------------------------------------------------------------- */

/* Per-sample ploidy detection: 2 if a heterozygote occurs among the first
 * min(1000, n_snp) SNPs, else 1 (reference: ComputeBackend::detect_sample_ploidy_device,
 * src/device/backend.hpp:765; src/device/cpu/cpu_backend.cpp:73-95). out: host int32[n_individuals]. */
int stp_detect_ploidy(stp_context* ctx, const stp_tile* tile, int32_t* out_ploidy);

/* Integer decode (reference: ComputeBackend::decode_af, src/device/backend.hpp:736,
 * in its exact integer form): d_a2[p*ld + s] = 2*ac, d_n[p*ld + s] = n as uint16,
 * ld >= n_snp and a multiple of 16.  q = a2 / (2 n) equals the reference's ac / n. */
int stp_decode_counts(stp_context* ctx, const stp_tile* tile, uint16_t* d_a2, uint16_t* d_n,
                      int64_t ld);

/* stp_decode_counts that allocates no device memory.  No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The tile must be device-resident (packed_on_device);
 * the population offsets and the per-sample ploidy go through a fixed device scratch array
 * of about 5.5 MiB that each GPU holds until the process ends, so one call takes at most
 * 65536 populations and 2^20 individuals (larger inputs: STP_ERR_INVALID).  Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU race.  Outputs are the
 * same as stp_decode_counts on the same inputs. */
int stp_decode_counts_noalloc(stp_context* ctx, const stp_tile* tile, uint16_t* d_a2, uint16_t* d_n,
                              int64_t ld);

/* Floating decode, bit-identical to the reference's DecodeResult{q, v, n}
 * (src/device/backend.hpp:250-256): host [P x M] arrays at i + P*s. */
int stp_decode_af(stp_context* ctx, const stp_tile* tile, double* q, double* v, double* n);

/* ---- f2 ----------------------------------------------------------------------- */

/* Per-block f2 from device counts (reference: ComputeBackend::compute_f2_blocks_device,
 * src/device/backend.hpp:694, fed by the decode).  block_id: host int32[M], contiguous
 * non-decreasing in [0, n_block).  Writes device d_f2, d_vpair [P x P x n_block]. */
int stp_f2_blocks_from_counts(stp_context* ctx, const uint16_t* d_a2, const uint16_t* d_n,
                              int64_t ld, int32_t P, int64_t M, const int32_t* block_id,
                              int32_t n_block, int32_t precision, double* d_f2, double* d_vpair);

/* Autotuner of the f2 contraction's launch configuration (replaces the reference's fixed
 * size buckets and VRAM chunk budget, src/device/cuda/cuda_backend_f2_blocks.cu:55-74,
 * src/device/vram_budget.hpp:63): runs stp_f2_blocks_from_counts on these device inputs
 * under every candidate configuration (work-item schedule over the dies, FP64 wavefront
 * tile 64x64 or 32x64, FP64 K-step prefetch, FP64 and pair-count grids), keeps the fastest
 * for the workload's key (device, precision, P, block-size histogram) in the context and
 * in the tune cache file (STEPPE_TUNE_CACHE, else ~/.cache/steppe_amd/f2_tune.tsv), which
 * later f2 calls and graphs of the same key use.  Every candidate writes the same tensor;
 * d_f2 / d_vpair hold it on return.  report (NUL-terminated JSON: key, chosen index,
 * candidates with their milliseconds) is truncated to report_len; report_needed receives
 * the full size.  Either may be NULL. */
int stp_f2_autotune(stp_context* ctx, const uint16_t* d_a2, const uint16_t* d_n, int64_t ld, int32_t P, int64_t M,
                    const int32_t* block_id, int32_t n_block, int32_t precision, double* d_f2, double* d_vpair,
                    char* report, int32_t report_len, int32_t* report_needed);

/* Per-block f2 from host FP64 Q/V/N [P x M] (reference: ComputeBackend::compute_f2_blocks,
 * src/device/backend.hpp:687).  Host outputs f2, vpair [P x P x n_block], block_sizes[n_block]. */
int stp_compute_f2_blocks(stp_context* ctx, const double* Q, const double* V, const double* N,
                          int32_t P, int64_t M, const int32_t* block_id, int32_t n_block,
                          int32_t precision, double* f2, double* vpair, int32_t* block_sizes);

/* Genotypes -> resident f2 handle in one call: decode, block assembly and the f2
 * contraction on the device (reference: run_extract_f2's compute core,
 * src/app/extract_f2_core.cpp:93-407, with the filter keeping every SNP). */
int stp_extract_f2(stp_context* ctx, const stp_tile* tile, const int32_t* block_id,
                   int32_t n_block, int32_t precision, stp_f2_handle** out);

/* Where an f2 result lives (reference: DeviceConfig::ForceTier and ExtractTier,
 * include/steppe/config.hpp:121-124, include/steppe/extract.hpp:33; the automatic choice of
 * src/device/tier_select.hpp:57-80).  AUTO keeps it in HBM when the result plus the
 * extraction's working set (operand planes, int8 V plane, fix-up list) take at most 0.70 of
 * the free device memory, else in host memory when it takes at most 0.60 of the available
 * host memory, else in an STPF2BK1 f2.bin file at disk_path.  The host and disk tiers contract
 * chunk_blocks blocks at a time into one device slab (0: 256 MiB of f2 + vpair, at most a
 * quarter of free device memory; the environment variable STEPPE_F2_CHUNK_BLOCKS sets it
 * when chunk_blocks is 0) and copy
 * each chunk out while the next is computed.  Every tier holds the same values.  Fits
 * (qpAdm, qpGraph, f-statistics) need a resident handle and return STP_ERR_INVALID for
 * the others; stp_f2_handle_download reads every tier. */
/* A resident extraction of one fixed shape captured once as a HIP graph and replayed
 * (replaces the host-driven chunk loop of the reference's f2 engine, which synchronises
 * after every size-bucket chunk: src/device/cuda/cuda_backend_f2_blocks.cu:301-347).
 * The tile must be device-resident (packed_on_device); its records, d_f2 and d_vpair
 * [P x P x n_block] must stay allocated while the graph exists.  d_a2 / d_n (uint16
 * [P x ld], ld >= n_snp, ld % 16 == 0) receive the counts, or are both NULL for counts
 * held by the graph.  Create runs the extraction once eagerly, then captures ploidy
 * detection (when the tile asks for it), decode, operand prep, pair counts, contraction
 * and the fix-up of cancellation-prone entries; each launch replays them with one graph
 * launch and returns when the tensors are complete.  Every launch rewrites the same
 * values as stp_decode_counts + stp_f2_blocks_from_counts on the same inputs. */
typedef struct stp_f2_graph stp_f2_graph;
int stp_f2_graph_create(stp_context* ctx, const stp_tile* tile, const int32_t* block_id, int32_t n_block,
                        int32_t precision, uint16_t* d_a2, uint16_t* d_n, int64_t ld, double* d_f2,
                        double* d_vpair, stp_f2_graph** out);
int stp_f2_graph_launch(stp_f2_graph* g);
int stp_f2_graph_free(stp_f2_graph* g);

#define STP_TIER_AUTO (-1)
#define STP_TIER_RESIDENT 0
#define STP_TIER_HOST 1
#define STP_TIER_DISK 2
/* stp_extract_f2 with a tier choice (reference: the tiered engines of
 * src/device/cuda/cuda_backend_f2_blocks.cu:349-603 and the disk sink of
 * src/device/cuda/block_sink.cu).  disk_path may be NULL unless the disk tier is used. */
int stp_extract_f2_tiered(stp_context* ctx, const stp_tile* tile, const int32_t* block_id,
                          int32_t n_block, int32_t precision, int32_t tier, const char* disk_path,
                          int64_t chunk_blocks, stp_f2_handle** out);
/* The tier a handle's tensor lives in (reference: F2ExtractResult::tier). */
int stp_f2_handle_tier(const stp_f2_handle* h, int32_t* tier);

/* ---- genotype files ---------------------------------------------------------------- */

/* Population selection of stp_extract_f2_files (reference: PopSelection,
 * src/io/ind_reader.hpp:20-31): the populations are sorted by label (the f2 P axis). */
#define STP_POPS_EXPLICIT 0  /* the n_pops labels in pops (every one must be in the .ind)   */
#define STP_POPS_TOP_K 1     /* the top_k populations with most individuals (ties: first seen) */
#define STP_POPS_MIN_N 2     /* every population with at least min_n individuals             */
/* SNP filters and blocking (reference: FilterConfig, include/steppe/config.hpp:130-160, as
 * bindings/bind_fstats.cpp:131-186 fills it): pooled minor allele frequency >= maf_min, at
 * most a fraction maxmiss of populations without data (1 disables), autosomes 1..22 only,
 * monomorphic dropped, transversions only; multiallelic or non-ACGT allele pairs are always
 * dropped, strand-ambiguous pairs (A/T, C/G) when strand_mode is 0 (1 keep, 2 flip = keep).
 * blgsize: jackknife block size in Morgans (0.05); the 2 Mb physical window is used when
 * every genetic position is 0.  ploidy: 0 detect per sample, 1 pseudo-haploid, 2 diploid. */
typedef struct stp_extract_request {
    const char* geno;            /* TGENO, GENO, EIGENSTRAT text or ANCESTRYMAP .geno, or a PLINK .bed */
    const char* snp;             /* .snp, or the PLINK .bim */
    const char* ind;             /* .ind, or the PLINK .fam (phenotype column = population) */
    int32_t pop_mode;
    const char* const* pops;
    int32_t n_pops;
    int64_t top_k;
    int64_t min_n;
    double maf_min;
    double maxmiss;
    int32_t autosomes_only;
    int32_t drop_monomorphic;
    int32_t transversions_only;
    int32_t strand_mode;
    double blgsize;
    int32_t ploidy;
    int32_t precision;           /* STP_PREC_* */
    int32_t tier;                /* STP_TIER_*; 0 is resident, as in stp_extract_f2 */
    const char* disk_path;       /* f2.bin path of the disk tier (may be NULL otherwise) */
    int64_t chunk_blocks;        /* blocks per streamed chunk of the host / disk tiers (0: auto) */
} stp_extract_request;
typedef struct stp_extract_info {
    int64_t n_snp_total;         /* min(header SNPs, .snp rows) */
    int64_t n_snp_kept;
    int64_t n_pseudo_haploid;
    int64_t n_diploid;
    int32_t P;
    int32_t n_block;
} stp_extract_info;

/* Genotype files -> resident f2 (reference: steppe::run_extract_f2,
 * src/app/extract_f2_core.cpp:96-407, with the genotype front end of
 * src/core/stats/genotype_front_end.cpp:9-37): reads the triple, builds the
 * individual-major tile in HBM (GENO is transposed on the device), detects ploidy,
 * decodes, keeps / compacts SNPs on the device, assigns blocks from the kept SNPs' .snp
 * positions and contracts f2.  The handle carries the population labels
 * (stp_f2_handle_label).  Unknown explicit populations, a filter that removes every SNP
 * or unreadable files return STP_ERR_INVALID. */
int stp_extract_f2_files(stp_context* ctx, const stp_extract_request* req, stp_f2_handle** out,
                         stp_extract_info* info);
/* Label of population i of a handle built from files ("" for other handles); writes at
 * most len bytes including the terminating NUL; returns the label's length in *n. */
int stp_f2_handle_label(const stp_f2_handle* h, int32_t i, char* buf, int32_t len, int32_t* n);

/* ---- f2 handle (reference: DeviceF2Blocks, src/device/device_f2_blocks.hpp:18-53) ---- */

/* Upload a host tensor (reference: upload_f2_blocks_to_device, device_f2_blocks.hpp:53).
 * vpair may be NULL (treated as all-present). */
int stp_f2_handle_upload(stp_context* ctx, const double* f2, const double* vpair,
                         const int32_t* block_sizes, int32_t P, int32_t n_block,
                         stp_f2_handle** out);
int stp_f2_handle_shape(const stp_f2_handle* h, int32_t* P, int32_t* n_block);
int stp_f2_handle_block_sizes(const stp_f2_handle* h, int32_t* block_sizes);
/* Device-resident tensors (reference: DeviceF2Blocks::f2_device / vpair_device); NULL for a
 * handle whose tensor is in the host or disk tier. */
const double* stp_f2_handle_f2_device(const stp_f2_handle* h);
const double* stp_f2_handle_vpair_device(const stp_f2_handle* h);
/* Copy back to host (reference: DeviceF2Blocks::to_host). */
int stp_f2_handle_download(const stp_f2_handle* h, double* f2, double* vpair);
int stp_f2_handle_free(stp_f2_handle* h);

/* ---- qpAdm on a resident f2 handle ------------------------------------------------ */

/* Fit options (reference: steppe::QpAdmOptions, include/steppe/qpadm.hpp:35-44). */
#define STP_JACKKNIFE_NONE 0
#define STP_JACKKNIFE_FEASIBLE 1  /* SEs only when every weight is in [0, 1] */
#define STP_JACKKNIFE_ALL 2
typedef struct stp_qpadm_options {
    double fudge;            /* ridge factor on tr(Q) and on the GLS normal matrices (1e-4) */
    int32_t als_iterations;  /* fixed ALS iterations per fit (20) */
    int32_t rank;            /* fitted rank; -1 = number of sources - 1 */
    double rank_alpha;       /* rank-test significance level (0.05) */
    int32_t jackknife;       /* STP_JACKKNIFE_*: when leave-one-block-out SEs are computed */
    int32_t se_require_p;    /* with STP_JACKKNIFE_FEASIBLE: also require p >= p_se_threshold */
    double p_se_threshold;   /* (0.05) */
} stp_qpadm_options;

/* Batched qpAdm on the device-resident tensor (reference: run_qpadm_search and
 * ComputeBackend::fit_models_batched, src/core/qpadm/model_search.cpp:30-218,
 * src/device/backend.hpp:1636-1647; a single model is run_qpadm,
 * src/core/qpadm/qpadm_fit.cpp:205-226).  Model i has target[i], sources
 * left[left_offsets[i] .. left_offsets[i+1]) and right set
 * right[right_offsets[i] .. right_offsets[i+1]) (right[...][0] is the base outgroup),
 * all population indices of the handle.  Results go to out[i] (host array of
 * n_models records), in input order, with out[i].model_index = i.  Models with
 * (sources x (right - 1)) <= 64 and right - 1 <= 32 run on the batched kernel, larger
 * ones (the reference's large path, model_search.cpp:85-106) on the large-model kernels
 * with the same arithmetic.  Limits: at most STP_QPADM_MAX_NL sources, sources <= right
 * populations beyond the base + 1 (the reference reads past its SVD's right vectors for
 * more: seed_AB, src/device/cpu/cpu_backend.cpp:2738-2744), sources x (right - 1) <= 1024
 * and right - 1 <= 128; violations return STP_ERR_INVALID.  Per-model domain outcomes
 * (rank deficient, non-SPD covariance, undefined chi-square) are reported in
 * out[i].status, as in the reference. */
int stp_qpadm_batch(stp_context* ctx, const stp_f2_handle* f2, int32_t n_models, const int32_t* target,
                    const int32_t* left_offsets, const int32_t* left, const int32_t* right_offsets,
                    const int32_t* right, const stp_qpadm_options* options, stp_qpadm_result* out);

/* ---- standalone f-statistics on a resident f2 handle ----------------------------------- */

/* f4 statistics (reference: steppe::run_f4 over the device tensor, src/core/qpadm/f4.cpp:38-111,
 * assemble_f4_quartets + jackknife_diag, src/device/backend.hpp; qpDstat on the f2 path is
 * the same computation, steppe::run_qpdstat).  quartets holds n rows (p1, p2, p3, p4) of
 * population indices; f4 = (f2(p2,p3) + f2(p1,p4) - f2(p1,p3) - f2(p2,p4)) / 2 per surviving
 * block, the weighted block-jackknife estimate est, its standard error se (NaN when the
 * variance is not positive), z = est / se and the two-sided normal p.  Out arrays hold n
 * values each; all NaN when no block survives.  Out-of-range indices: STP_ERR_INVALID. */
int stp_f4(stp_context* ctx, const stp_f2_handle* f2, const int32_t* quartets, int64_t n, double* est, double* se,
           double* z, double* p);

/* f3 statistics (reference: steppe::run_f3, src/core/qpadm/f3.cpp:37-105): rows (C, A, B),
 * f3 = (f2(C,A) + f2(C,B) - f2(A,B)) / 2 per block, read out as stp_f4. */
int stp_f3(stp_context* ctx, const stp_f2_handle* f2, const int32_t* triples, int64_t n, double* est, double* se,
           double* z, double* p);

/* f4-ratio (reference: steppe::run_f4ratio, src/core/qpadm/f4ratio.cpp:34-100): rows
 * (p1, p2, p3, p4, p5), alpha = f4(p1,p2;p3,p4) / f4(p1,p2;p5,p4) by the ratio block
 * jackknife; blocks whose denominator block value is below 1e-6 in magnitude are left
 * out for that row.  alpha / se / z are NaN when fewer than two blocks remain. */
int stp_f4ratio(stp_context* ctx, const stp_f2_handle* f2, const int32_t* tuples, int64_t n, double* alpha,
                double* se, double* z);

/* All-combination sweep request (reference: SweepRequest, include/steppe/fstat_sweep.hpp:25-38). */
#define STP_SWEEP_MIN_Z 0   /* keep |z| >= min_z */
#define STP_SWEEP_TOP_K 1   /* keep the top_k largest |z| among those with |z| >= min_z */
#define STP_SWEEP_AUTO 0    /* method: library's choice */
#define STP_SWEEP_DIRECT 1  /* method: each statistic from its own block series */
#define STP_SWEEP_GRAM 2    /* method: f4 only, block-covariance Gram matrices on FP64 MFMA */
typedef struct stp_sweep_request {
    int32_t k;              /* 4: f4 of (c0, c1, c2, c3); 3: f3 of (c0; c1, c2), c0 < c1 < ... */
    int32_t filter;         /* STP_SWEEP_MIN_Z or STP_SWEEP_TOP_K */
    double min_z;           /* 3.0 */
    int64_t top_k;          /* 100; <= 0 takes the default reservoir of 1e6 */
    int32_t sure;           /* non-zero lifts the 1e8-combination cap */
    int32_t method;         /* STP_SWEEP_AUTO / DIRECT / GRAM */
    const int32_t* subset;  /* n_subset population indices to combine (NULL: all P) */
    int32_t n_subset;
} stp_sweep_request;

/* All-combination f4 / f3 sweep (reference: steppe::run_f4_sweep / run_f3_sweep,
 * src/core/qpadm/fstat_sweep.cpp:31-110).  Enumerates the C(range, k) combinations of the
 * population range, keeps those passing the filter, at most K of them (top_k in top-K mode,
 * 1e6 otherwise), ordered by |z| descending then combination order.  keys receives 4 ints
 * per row (population indices; the fourth is 0 for f3), est / se / z / p one value per
 * row, up to capacity rows; *n_out is the number of survivors (may exceed capacity: then
 * only the first capacity rows are written).  *enumerated = C(range, k); when it exceeds
 * 1e8 and sure is 0 nothing is computed, *capped = 1 and *status = 5 (invalid_config);
 * otherwise *status = 0. */
int stp_fstat_sweep(stp_context* ctx, const stp_f2_handle* f2, const stp_sweep_request* req, int64_t capacity,
                    int32_t* keys, double* est, double* se, double* z, double* p, int64_t* n_out,
                    uint64_t* enumerated, int32_t* capped, int32_t* status);

/* ---- qpGraph ------------------------------------------------------------------------ */

typedef struct stp_qpgraph_topology stp_qpgraph_topology;
typedef struct stp_qpgraph_set stp_qpgraph_set;

/* Parse a parent->child edge list into an admixture graph whose leaves are named among
 * names (the f2 handle's population labels) (reference: parse_qpgraph,
 * src/core/qpadm/qpgraph_model.cpp:41-243).  Invalid graphs return STP_ERR_INVALID
 * with the reference's message. */
int stp_qpgraph_parse(int32_t n_edges, const char* const* from, const char* const* to, int32_t n_names,
                      const char* const* names, stp_qpgraph_topology** out);
int stp_qpgraph_topology_free(stp_qpgraph_topology* t);
/* dims = [npop, nedge_norm, nadmix, npair, npath, base_leaf] (QpGraphModel fields). */
int stp_qpgraph_topology_dims(const stp_qpgraph_topology* t, int32_t* dims);
/* Labels: kind 0 leaf, 1 admixture-edge parent, 2 admixture child, 3 normal-edge parent,
 * 4 normal-edge child (QpGraphResult's leaves / admix_from / admix_to / edge_from /
 * edge_to); index within that list. */
int stp_qpgraph_topology_label(const stp_qpgraph_topology* t, int32_t kind, int32_t index, char* buf,
                               int32_t len);

/* A resident set of graphs, graph g scored against f2[g]: the f3 basis
 * (base; a, b), its block-jackknife covariance with the diag_f3 ridge and its inverse
 * run on the device once (reference: assemble_f3_triples + jackknife_cov,
 * src/core/qpadm/qpgraph_fit.cpp:57-80).  The same topology or handle may appear several
 * times. */
int stp_qpgraph_set_create(stp_context* ctx, int32_t n_graphs, const stp_qpgraph_topology* const* topologies,
                           const stp_f2_handle* const* f2, double diag_f3, stp_qpgraph_set** out);
int stp_qpgraph_set_free(stp_qpgraph_set* s);
/* Per graph: 0 ok, 3 no surviving block or singular covariance. */
int stp_qpgraph_set_status(const stp_qpgraph_set* s, int32_t* status);
/* Row strides of the per-instance arrays of stp_qpgraph_eval: dmax (theta, gradient),
 * emax (edge lengths) and pmax (fitted f3) = the largest nadmix / nedge_norm / npair of
 * the set (dmax >= 1); pmax may be NULL. */
int stp_qpgraph_set_strides(const stp_qpgraph_set* s, int32_t* dmax, int32_t* emax, int32_t* pmax);
/* Graph g's f_obs (npair), unridged Q diagonal (npair) and Qinv (npair x npair, row-major). */
int stp_qpgraph_set_basis(stp_context* ctx, const stp_qpgraph_set* s, int32_t graph, double* f_obs, double* qdiag,
                          double* qinv);

/* Batched score and adjoint (reference: qpgraph_score, src/core/qpadm/qpgraph_objective.hpp:163,
 * evaluated by qpgraph_fit_fleet, src/device/cpu/cpu_backend.cpp:2936; the adjoint has no
 * counterpart there: the reference differentiates by finite differences).  Instance i
 * scores graph inst_graph[i] at theta[i*dmax .. +nadmix).
 *   mode 0: the score with edge lengths from the ridge NNLS (constrained) or solve;
 *   mode 1: mode 0 + d score / d theta through the edge-length solve;
 *   mode 2: the score for given edge lengths bl[i*emax ..], with d / d theta and d / d bl.
 * Outputs (host, may be NULL except score): score[n_inst], grad_theta[n_inst*dmax],
 * grad_bl[n_inst*emax], bl_out[n_inst*emax] (the edge lengths used), fit[n_inst*pmax]
 * (the fitted f3 basis).  A failed solve gives score = +inf. */
int stp_qpgraph_eval(stp_context* ctx, stp_qpgraph_set* s, int32_t n_inst, const int32_t* inst_graph,
                     const double* theta, const double* bl, int32_t mode, double fudge, int32_t constrained,
                     double* score, double* grad_theta, double* grad_bl, double* bl_out, double* fit);

/* A set from a given f3 basis and inverse covariance per graph, the inputs of the
 * reference's fleet fit (ComputeBackend::qpgraph_fit_fleet, src/device/backend.hpp:1176;
 * qpgraph_fit_fleet_batch, :1188): f_obs holds npair values per graph and qinv
 * npair x npair row-major per graph, concatenated in graph order.  The set's Q diagonal
 * (stp_qpgraph_set_basis) is reported as zeros. */
int stp_qpgraph_set_create_basis(stp_context* ctx, int32_t n_graphs, const stp_qpgraph_topology* const* topologies,
                                 const double* f_obs, const double* qinv, stp_qpgraph_set** out);

/* Options of the device fit.  The reference's coordinate search with finite
 * differences (cpu_backend.cpp:2936-3045) is replaced by a projected L-BFGS on [0,1]^D
 * driven by the device adjoint, one instance per workgroup to convergence: it stops when
 * the projected gradient is <= tol * 1e-2, the step <= tol * 1e-3, the relative decrease
 * <= 1e-14, the line search fails, or after maxit iterations.  Defaults (the
 * reference's QpGraphOptions): numstart 10, maxit 200, tol 1e-9, fudge 1e-4,
 * constrained 1. */
typedef struct stp_qpgraph_fit_options {
    int32_t numstart;
    int32_t maxit;
    double tol;
    double fudge;
    int32_t constrained;
} stp_qpgraph_fit_options;

/* Fit instances from given start points: instance i fits graph inst_graph[i] from
 * theta0[i*dmax ..]; outputs theta_out[n_inst*dmax] (may be NULL), score_out[n_inst]
 * (+inf when the start point's edge-length system is singular) and iterations[n_inst]
 * (-1 then; may be NULL).  numstart is not used. */
int stp_qpgraph_fit(stp_context* ctx, stp_qpgraph_set* s, int32_t n_inst, const int32_t* inst_graph,
                    const double* theta0, const stp_qpgraph_fit_options* options, double* theta_out,
                    double* score_out, int32_t* iterations);

/* Multistart fit of every graph of the set (reference: qpgraph_fit_fleet per topology,
 * src/device/cpu/cpu_backend.cpp:2936-3045, with its splitmix start points): numstart
 * restarts per graph (one when it has no admixture), all restarts of all graphs in one
 * launch; per graph the best score, the restart spread (max - min score), the best
 * theta and each coordinate's range over the restarts (rows of dmax), the edge lengths
 * (rows of emax) and fitted f3 basis (rows of pmax) at the best theta, the largest
 * iteration count and a status (0 ok, 3 singular covariance or edge-length system).
 * Any output except score and status may be NULL. */
int stp_qpgraph_fit_fleet(stp_context* ctx, stp_qpgraph_set* s, const stp_qpgraph_fit_options* options, double* score,
                          double* restart_spread, double* theta, double* theta_lo, double* theta_hi,
                          double* edge_length, double* f3_fit, int32_t* iterations, int32_t* status);

/* ---- qpGraph topology enumeration and search ------------------------------------- */

/* A list of candidate graphs (reference: EnumeratedTopology,
 * src/core/qpadm/qpgraph_enumerate.hpp:19-25): edges (parent, child names; leaves keep
 * their labels, other nodes are "N<id>"), admixture count, id and shape hash. */
typedef struct stp_qpgraph_list stp_qpgraph_list;
/* Every tree on the leaves, then (max_nadmix >= 1) every distinct one-admixture graph
 * (reference: enumerate_bounded_space, src/core/qpadm/qpgraph_enumerate.cpp:229-236; 105
 * trees and 1485 admixture graphs on 5 leaves).  Same order, names and hashes. */
int stp_qpgraph_enumerate(int32_t n_leaves, const char* const* leaves, int32_t max_nadmix,
                          stp_qpgraph_list** out);
/* Hill-climb moves from a graph (reference: topology_neighbors,
 * src/core/qpadm/qpgraph_enumerate.cpp:383-407). */
int stp_qpgraph_neighbors(int32_t n_edges, const char* const* from, const char* const* to, int32_t nadmix,
                          uint64_t hash, int32_t n_leaves, const char* const* leaves, int32_t max_nadmix,
                          stp_qpgraph_list** out);
/* Shape hash of an edge list, leaves ranked by label (reference: graph_hash,
 * src/core/qpadm/qpgraph_enumerate.cpp:183-202). */
int stp_qpgraph_graph_hash(int32_t n_edges, const char* const* from, const char* const* to, uint64_t* hash);
int stp_qpgraph_list_size(const stp_qpgraph_list* l, int32_t* n);
int stp_qpgraph_list_get(const stp_qpgraph_list* l, int32_t i, int32_t* nadmix, int32_t* id, uint64_t* hash,
                         int32_t* n_edges);
/* Edge k of graph i; names are cut to len - 1 bytes. */
int stp_qpgraph_list_edge(const stp_qpgraph_list* l, int32_t i, int32_t k, char* from, char* to, int32_t len);
int stp_qpgraph_list_free(stp_qpgraph_list* l);

/* Options of stp_qpgraph_search_run (reference: QpGraphSearchOptions,
 * include/steppe/qpgraph_search.hpp:30-36, and QpGraphOptions for the fits). */
typedef struct stp_qpgraph_search_options {
    int32_t max_nadmix;          /* 0: trees only; 1: trees and one-admixture graphs (larger acts as 1) */
    int32_t numstart;            /* restarts per admixture graph (10) */
    int32_t maxit;               /* L-BFGS iterations per restart (200) */
    double tol;                  /* 1e-9 */
    double fudge;                /* 1e-4 */
    double diag_f3;              /* 1e-5 */
    int32_t constrained;         /* 1 */
    int32_t run_heuristic;       /* 1: also run the hill-climb from heuristic_seeds starts (8) */
    int32_t heuristic_seeds;
} stp_qpgraph_search_options;
typedef struct stp_qpgraph_search_summary {
    int32_t status;              /* 0 ok, 3 non_spd_covariance */
    int32_t n_trees;
    int32_t n_admix1;
    int32_t n_candidates;        /* enumerated */
    int32_t n_fitted;            /* candidates that parse on this pop-set */
    int32_t best;                /* index into the fitted candidates */
    int32_t best_candidate;      /* index into the enumeration */
    int32_t best_nadmix;
    uint64_t best_hash;
    double best_score;
    double second_best_score;
    int32_t heuristic_recovered;
    int32_t n_seeds;
    double fit_all_wall_ms;      /* the fleet fit of every candidate */
    double topologies_per_s;
} stp_qpgraph_search_summary;
typedef struct stp_qpgraph_search stp_qpgraph_search;
/* The topology search (reference: run_qpgraph_search, src/core/qpadm/qpgraph_search.cpp:
 * 110-297): every candidate on pops (labels among names, the f2 handle's population
 * labels; at least 3, distinct) is parsed with pops[0] as f3 base, every candidate shares
 * one f3 basis f3(pops[0]; a, b), a <= b over pops[1..], with its covariance and inverse
 * computed once on the device, and all restarts of all candidates are fitted in one
 * device fleet; the best is the first lowest score.  A pop outside names returns
 * STP_ERR_INVALID. */
int stp_qpgraph_search_run(stp_context* ctx, const stp_f2_handle* f2, int32_t n_names, const char* const* names,
                           int32_t n_pops, const char* const* pops, const stp_qpgraph_search_options* options,
                           stp_qpgraph_search** out);
int stp_qpgraph_search_get_summary(const stp_qpgraph_search* s, stp_qpgraph_search_summary* out);
/* The enumeration the search scored (borrowed; valid until stp_qpgraph_search_free). */
const stp_qpgraph_list* stp_qpgraph_search_candidates(const stp_qpgraph_search* s);
/* Per fitted candidate (n_fitted rows, enumeration order): its index in the enumeration,
 * best score and restart spread (reference: QpGraphCandidate).  Any output may be NULL. */
int stp_qpgraph_search_scores(const stp_qpgraph_search* s, int32_t* candidate, double* score,
                              double* restart_spread);
/* Final graph hash of each hill-climb seed (n_seeds). */
int stp_qpgraph_search_seeds(const stp_qpgraph_search* s, uint64_t* hashes);
/* The best graph's fit (reference: QpGraphSearchResult::best_fit): dims as
 * stp_qpgraph_topology_dims; weights, their restart range (nadmix), edge lengths (nedge),
 * fitted and observed f3 and the unridged Q diagonal (npair).  Outputs may be NULL. */
int stp_qpgraph_search_best_fit(const stp_qpgraph_search* s, int32_t* dims, double* weight, double* weight_lo,
                                double* weight_hi, double* edge_length, double* f3_fit, double* f_obs,
                                double* qdiag);
/* Labels of the best graph, kinds as stp_qpgraph_topology_label. */
int stp_qpgraph_search_best_label(const stp_qpgraph_search* s, int32_t kind, int32_t index, char* buf, int32_t len);
int stp_qpgraph_search_free(stp_qpgraph_search* s);

/* ---- STPF2BK1 f2.bin (reference: src/device/f2_disk_format.hpp:17-50,
 *      src/app/f2_dir_writer.cpp:306-440, src/app/f2_dir_io.cpp:52-158) ------------- */

int stp_f2bin_write(const char* path, const double* f2, const double* vpair,
                    const int32_t* block_sizes, int32_t P, int32_t n_block);
int stp_f2bin_read_header(const char* path, int32_t* P, int32_t* n_block);
int stp_f2bin_read(const char* path, double* f2, double* vpair, int32_t* block_sizes);

/* ---- instrumentation (no counterpart in the reference) ---------------------------- */

/* The automatic tier choice of stp_extract_f2_tiered (host only, no device): the
 * contraction's device working set for this partition and precision, and the tier it
 * picks given free device and host memory in bytes. */
int stpx_f2_tier_choice(int32_t P, int64_t M, const int32_t* block_id, int32_t n_block, int32_t precision,
                        uint64_t free_vram, uint64_t free_host, uint64_t* working_bytes, int32_t* tier);

/* Enable per-stage device timing: HIP events recorded on the context's own stream
 * around the decode, prep and f2 contraction launches. */
int stpx_timing(stp_context* ctx, int32_t enable);
/* Accumulated milliseconds and launch counts per stage [decode, prep, f2, count, qpadm,
 * qpgraph, fstat, frontend, fixup] (arrays of 9; f2 = the contraction kernel, count = the
 * pair-count kernel, qpadm = the batched fit kernels of one stp_qpadm_batch call,
 * qpgraph = one stp_qpgraph_eval or stp_qpgraph_fit launch, fstat = the kernels of one
 * f4 / f3 / f4-ratio / sweep call, frontend = each GENO transpose block, the keep mask and
 * the compaction of stp_extract_f2_files, fixup = the FP64 recomputation of the f2
 * entries the contraction flagged as cancellation-prone); reset clears them. */
int stpx_timing_read(stp_context* ctx, double* ms, int64_t* count, int32_t reset);

/* ---- synthetic panels (bench/test input; no counterpart in the reference) ------------ */

/* Generate a packed panel directly in device memory.  Per population: F_ST fst[p],
 * missing rate pop_missing[p]; per individual: population ind_pop[g] and
 * ind_haploid[g] (1 = pseudo-haploid).  d_packed must hold n_ind * bytes_per_record
 * bytes; bytes_per_record is a multiple of 4 and >= ceil(M/4). */
int stpx_synth_panel(stp_context* ctx, uint64_t seed, int32_t P, int64_t M, int64_t n_ind,
                     int64_t bytes_per_record, const float* fst, const float* pop_missing,
                     const int32_t* ind_pop, const int8_t* ind_haploid, uint8_t* d_packed);

#ifdef __cplusplus
}
#endif

#endif /* STEPPE_HIP_H */
