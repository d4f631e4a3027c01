// steppe_amd/csrc/frontend.hpp — launch interface of the genotype-file front-end kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace stp {

// The per-SNP keep rule of the extract (the reference's FilterConfig as its extract
// applies it, include/steppe/config.hpp:130-160 with src/app/extract_f2_core.cpp:184-262).
struct KeepRule {
    double maf_min = 0.0;             // pooled minor allele frequency floor
    double snp_max_missing = 1.0;     // pooled individual-missingness cap (1: inactive)
    double maxmiss = 1.0;             // cap on the fraction of populations with no data
    double ploidy = 2.0;              // divisor turning allele counts into individuals
    double total_individuals = 0.0;   // selected individuals
    int autosomes_only = 0, drop_monomorphic = 0, transversions_only = 0, strand_drop = 1;
};

// SNP-major records (src, n_snp records of src_bpr bytes) -> individual-major canonical
// records: individual g of the selection (source column sel[g]) gets ceil(n_snp/64)*16
// bytes at out + g*out_bpr + col_off.  out_bpr and col_off are multiples of 16.
// bed: the source is a PLINK .bed (LSB-first codes, recoded to copies of A1).
void launch_transpose_gather(const std::uint8_t* src, std::int64_t src_bpr, std::int64_t n_snp, const std::int32_t* sel,
                             std::int64_t n_sel, std::uint8_t* out, std::int64_t out_bpr, std::int64_t col_off,
                             bool bed, hipStream_t st);

// flags[s] = 1 when SNP s (counts a2 / n at p*ld + s) passes the rule.
void launch_snp_keep(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, std::int64_t M,
                     const char* ref, const char* alt, const std::int32_t* chrom, const KeepRule& rule,
                     std::uint8_t* flags, hipStream_t st);

// Columns kept[0 .. M_kept) of the counts into [P x ld_out] (zero past M_kept).
void launch_compact_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, const std::int32_t* kept,
                           std::int64_t M_kept, int P, std::uint16_t* a2_out, std::uint16_t* n_out, std::int64_t ld_out,
                           hipStream_t st);

}  // namespace stp
