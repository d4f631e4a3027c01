// steppe_amd/csrc/frontend_host.cpp — genotype files to a resident f2 tensor
// (HipBackend::extract_f2_files).
//
// Follows the reference's run_extract_f2 (src/app/extract_f2_core.cpp:96-407) on its
// device path: the genotype front end (src/core/stats/genotype_front_end.cpp:9-37: open
// the .geno, read the .ind selection against the records present, read the .snp,
// M0 = min(header SNPs, .snp rows), build the canonical individual-major tile — TGENO
// records gathered on the host, GENO and PLINK .bed records streamed in blocks of about
// 1 GiB and transposed (and, for .bed, recoded) on the device,
// read_canonical_tile.cpp:60-180), per-sample ploidy, the
// decode, the regime-B keep mask and compaction (decode_compact_kernel.cu,
// cuda_backend_decode.cu:490-620), assign_blocks on the kept SNPs' genetic positions
// (block_partition_rule.cpp:14-87, with the 2 Mb physical fallback when the map is all
// zero), then the f2 contraction over the kept columns.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <unordered_set>

#include "engine.hpp"
#include "frontend.hpp"
#include "geno_io.hpp"

namespace stp {

namespace {

constexpr std::size_t kStageBytes = std::size_t{1} << 30;   // GENO source bytes per transpose block
constexpr double kBpWindow = 2.0e6;                          // include/steppe/config.hpp:55
constexpr std::size_t kMaxPopIndividuals = 16383;            // uint16 allele counts

// Jackknife blocks of the kept SNPs (src/core/domain/block_partition_rule.cpp:14-87).
int block_walk(const std::vector<std::int32_t>& chrom, const std::vector<double>& genpos,
               const std::vector<double>& physpos, double blgsize, std::vector<std::int32_t>& out) {
    const std::size_t M = chrom.size();
    out.assign(M, 0);
    if (!(blgsize > 0.0) || M == 0) return 0;
    const std::vector<double>* pos = &genpos;
    double window = blgsize;
    const bool gen_zero = std::all_of(genpos.begin(), genpos.end(), [](double v) { return v == 0.0; });
    const bool phys_zero = std::all_of(physpos.begin(), physpos.end(), [](double v) { return v == 0.0; });
    if (gen_zero && !phys_zero) {
        pos = &physpos;
        window = kBpWindow;
    }
    long blk = -1;
    int cur = -1;
    double start = -1.0e20;
    for (std::size_t s = 0; s < M; ++s) {
        if (chrom[s] != cur || (*pos)[s] - start >= window) {
            ++blk;
            cur = chrom[s];
            start = (*pos)[s];
        }
        out[s] = static_cast<std::int32_t>(blk);
    }
    return static_cast<int>(blk + 1);
}

}  // namespace

F2Blocks* HipBackend::extract_f2_files(const ExtractRequest& req, ExtractInfo& info) {
    activate();
    info = ExtractInfo{};
    STP_REQUIRE(!req.geno.empty() && !req.snp.empty() && !req.ind.empty(),
                "extract_f2: the genotype triple (.geno/.snp/.ind) is required");
    STP_REQUIRE(req.ploidy == 0 || req.ploidy == 1 || req.ploidy == 2, "extract_f2: ploidy must be 0 (auto), 1 or 2");
    STP_REQUIRE(req.strand_mode >= 0 && req.strand_mode <= 2, "extract_f2: strand_mode must be 0, 1 or 2");
    STP_REQUIRE(req.precision == STP_PREC_FP64 || req.precision == STP_PREC_SPLIT_BF16,
                "extract_f2: unknown precision form");

    // ---- front end: files ----
    const GenoFile f = open_geno(req.geno);
    const bool plink = f.header.format == GenoFormat::Plink;
    const std::vector<PopGroup> groups =
        plink ? read_fam(req.ind, req.pops, f.records_present) : read_ind(req.ind, req.pops, f.records_present);
    if (req.pops.mode == PopSelect::kExplicit)
        for (const auto& want : req.pops.labels) {
            const bool found = std::any_of(groups.begin(), groups.end(), [&](const PopGroup& g) { return g.label == want; });
            STP_REQUIRE(found, "extract_f2: pops contains an unknown population '" + want + "' (not present in the .ind)");
        }
    const SnpTable snp = plink ? read_bim(req.snp) : read_snp(req.snp);
    const std::size_t M0 = std::min(f.header.n_snp, snp.size());
    const int P = static_cast<int>(groups.size());
    STP_REQUIRE(P > 0 && M0 > 0, "extract_f2: empty selection or no SNPs");
    std::vector<std::int32_t> sel;
    std::vector<std::uint64_t> offs{0};
    for (const auto& g : groups) {
        STP_REQUIRE(g.rows.size() <= kMaxPopIndividuals,
                    "extract_f2: population '" + g.label + "' exceeds 16383 individuals");
        for (std::size_t r : g.rows) sel.push_back(static_cast<std::int32_t>(r));
        offs.push_back(sel.size());
    }
    const std::int64_t n_sel = static_cast<std::int64_t>(sel.size());

    // ---- canonical individual-major tile in HBM ----
    const std::int64_t out_bpr = round_up(ceil_div(static_cast<std::int64_t>(M0), 4), 16);
    fe_tile_.ensure(static_cast<std::size_t>(n_sel * out_bpr));
    STP_HIP(hipMemsetAsync(fe_tile_.get(), 0, static_cast<std::size_t>(n_sel * out_bpr), stream_));
    if (f.header.format == GenoFormat::Tgeno) {
        std::vector<std::uint8_t> host;
        read_tgeno_selected(f, groups, M0, host);
        const std::size_t w = (M0 + 3) / 4;
        STP_HIP(hipMemcpy2DAsync(fe_tile_.get(), static_cast<std::size_t>(out_bpr), host.data(), w, w,
                                 static_cast<std::size_t>(n_sel), hipMemcpyHostToDevice, stream_));
        sync();
    } else {
        for (std::int32_t r : sel)
            STP_REQUIRE(static_cast<std::size_t>(r) < f.header.n_ind,
                        "extract_f2: individual row " + std::to_string(r) + " beyond the header's individuals");
        STP_REQUIRE(M0 <= f.records_present, "extract_f2: the .geno holds fewer SNP records than its header");
        fe_i32_.ensure(static_cast<std::size_t>(n_sel));
        STP_HIP(hipMemcpyAsync(fe_i32_.get(), sel.data(), sel.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                               stream_));
        const std::size_t bpr = f.header.bytes_per_record;
        // Source bytes per transpose block (about 1 GiB; STEPPE_GENO_BLOCK_BYTES overrides it).
        std::size_t stage = kStageBytes;
        if (const char* env = std::getenv("STEPPE_GENO_BLOCK_BYTES")) {
            const long long v = std::strtoll(env, nullptr, 10);
            if (v > 0) stage = static_cast<std::size_t>(v);
        }
        std::size_t blk = std::max<std::size_t>(stage / bpr / 64 * 64, 64);
        blk = std::min(blk, (M0 + 63) / 64 * 64);
        fe_stage_.ensure(blk * bpr);
        std::vector<std::uint8_t> host;
        std::unique_ptr<AsciiGenoReader> text;
        if (f.header.format == GenoFormat::Eigenstrat || f.header.format == GenoFormat::Ancestrymap)
            text = std::make_unique<AsciiGenoReader>(f);
        for (std::size_t s0 = 0; s0 < M0; s0 += blk) {
            const std::size_t s1 = std::min(M0, s0 + blk);
            if (text) text->next(s1 - s0, host);
            else read_geno_records(f, s0, s1, host);
            STP_HIP(hipMemcpyAsync(fe_stage_.get(), host.data(), host.size(), hipMemcpyHostToDevice, stream_));
            mark_begin(kFront);
            launch_transpose_gather(fe_stage_.get(), static_cast<std::int64_t>(bpr),
                                    static_cast<std::int64_t>(s1 - s0), fe_i32_.get(), n_sel, fe_tile_.get(), out_bpr,
                                    static_cast<std::int64_t>(s0 / 4), plink, stream_);
            mark_end(kFront);
            sync();   // host is reused by the next block
        }
    }

    // ---- ploidy and decode ----
    stp_tile tile{};
    tile.packed = fe_tile_.get();
    tile.bytes_per_record = static_cast<std::size_t>(out_bpr);
    tile.n_snp = M0;
    tile.n_individuals = static_cast<std::size_t>(n_sel);
    tile.pop_offsets = offs.data();
    tile.n_pop = P;
    tile.sample_ploidy = nullptr;
    tile.ploidy = req.ploidy == 0 ? 2 : req.ploidy;
    tile.detect_ploidy_on_device = req.ploidy == 0 ? 1 : 0;
    tile.packed_on_device = 1;
    prepare_tile(tile);
    {
        std::vector<std::int8_t> pl(static_cast<std::size_t>(n_sel));
        STP_HIP(hipMemcpyAsync(pl.data(), ploidy_.get(), pl.size(), hipMemcpyDeviceToHost, stream_));
        sync();
        for (std::int8_t v : pl) (v == 1 ? info.n_pseudo_haploid : info.n_diploid)++;
    }
    const std::int64_t M = static_cast<std::int64_t>(M0);
    const std::int64_t ld = round_up(M, 64);
    a2_.ensure(static_cast<std::size_t>(P) * ld);
    n_.ensure(static_cast<std::size_t>(P) * ld);
    mark_begin(kDecode);
    launch_decode_counts(d_packed_, out_bpr, M, pop_offsets_.get(), ploidy_.get(), P, a2_.get(), n_.get(), ld, stream_);
    mark_end(kDecode);

    // ---- keep mask ----
    std::vector<char> alleles(2 * M0);
    std::copy(snp.ref.begin(), snp.ref.begin() + M0, alleles.begin());
    std::copy(snp.alt.begin(), snp.alt.begin() + M0, alleles.begin() + M0);
    fe_alleles_.ensure(alleles.size());
    STP_HIP(hipMemcpyAsync(fe_alleles_.get(), alleles.data(), alleles.size(), hipMemcpyHostToDevice, stream_));
    fe_i32_.ensure(std::max<std::size_t>(M0, static_cast<std::size_t>(n_sel)));
    STP_HIP(hipMemcpyAsync(fe_i32_.get(), snp.chrom.data(), M0 * sizeof(std::int32_t), hipMemcpyHostToDevice, stream_));
    fe_flags_.ensure(M0);
    KeepRule rule;
    rule.maf_min = req.maf_min;
    rule.snp_max_missing = 1.0;
    rule.maxmiss = req.maxmiss;
    rule.ploidy = 2.0;
    rule.total_individuals = static_cast<double>(n_sel);
    rule.autosomes_only = req.autosomes_only;
    rule.drop_monomorphic = req.drop_monomorphic;
    rule.transversions_only = req.transversions_only;
    rule.strand_drop = req.strand_mode == 0;
    mark_begin(kFront);
    launch_snp_keep(a2_.get(), n_.get(), ld, P, M, fe_alleles_.get(), fe_alleles_.get() + M0, fe_i32_.get(), rule,
                    fe_flags_.get(), stream_);
    mark_end(kFront);
    std::vector<std::uint8_t> flags(M0);
    STP_HIP(hipMemcpyAsync(flags.data(), fe_flags_.get(), M0, hipMemcpyDeviceToHost, stream_));
    sync();
    std::vector<std::int32_t> kept, chrom_k;
    std::vector<double> gen_k, phys_k;
    for (std::size_t s = 0; s < M0; ++s)
        if (flags[s]) {
            kept.push_back(static_cast<std::int32_t>(s));
            chrom_k.push_back(snp.chrom[s]);
            gen_k.push_back(snp.genpos[s]);
            phys_k.push_back(snp.physpos[s]);
        }
    const std::int64_t Mk = static_cast<std::int64_t>(kept.size());
    STP_REQUIRE(Mk > 0, "extract_f2: every SNP was filtered out (0 of " + std::to_string(M0) +
                            " kept) - relax the filters");
    std::vector<std::int32_t> block_id;
    const int nb = block_walk(chrom_k, gen_k, phys_k, req.blgsize, block_id);
    STP_REQUIRE(nb > 0, "extract_f2: assign_blocks produced 0 blocks (check blgsize and the .snp genetic positions)");

    // ---- compact the kept columns, then the contraction ----
    const std::uint16_t* a2 = a2_.get();
    const std::uint16_t* nn = n_.get();
    std::int64_t ldk = ld;
    if (Mk < M) {
        ldk = round_up(Mk, 64);
        fe_i32_.ensure(static_cast<std::size_t>(Mk));
        STP_HIP(hipMemcpyAsync(fe_i32_.get(), kept.data(), kept.size() * sizeof(std::int32_t), hipMemcpyHostToDevice,
                               stream_));
        fe_a2_.ensure(static_cast<std::size_t>(P) * ldk);
        fe_n_.ensure(static_cast<std::size_t>(P) * ldk);
        mark_begin(kFront);
        launch_compact_counts(a2_.get(), n_.get(), ld, fe_i32_.get(), Mk, P, fe_a2_.get(), fe_n_.get(), ldk, stream_);
        mark_end(kFront);
        a2 = fe_a2_.get();
        nn = fe_n_.get();
    }
    auto out = std::unique_ptr<F2Blocks>(f2_counts_tiered(a2, nn, ldk, P, Mk, block_id.data(), nb, req.precision,
                                                          req.tier));
    for (const auto& g : groups) out->labels.push_back(g.label);
    info.n_snp_total = M;
    info.n_snp_kept = Mk;
    return out.release();
}

}  // namespace stp
