// steppe_amd/csrc/frontend.hip — device half of the genotype-file front end: the
// SNP-major (GENO) to individual-major transpose with the population gather, the per-SNP
// keep decision over the decoded counts, and the compaction of the kept SNP columns.
//
//   transpose + gather   src/device/cuda/transpose_canonical_kernel.cu:34-68 (one output
//                        byte per thread there; here one thread packs 64 SNPs of one
//                        individual into a 16-byte store)
//   pooled summary       src/io/filter/snp_summary_reduce.hpp:36-62 (sums in population
//                        order, each product rounded before the add: this file is built
//                        with FP contraction off)
//   keep decision        src/io/filter/snp_summary_reduce.hpp:65-77, filter_decision.hpp:17-91,
//                        with the population-axis missingness test of
//                        src/device/cuda/decode_compact_kernel.cu:39-67 and the per-SNP
//                        individual-missingness threshold lifted to 1 as the reference's
//                        extract does (cuda_backend_decode.cu:512-513)
//   compaction           decode_compact_kernel.cu:94-108 (kept columns of a2 / n instead of
//                        Q / V / N: q, v, n follow from the counts exactly)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "frontend.hpp"

namespace stp {
namespace {

// ---- transpose + gather ---------------------------------------------------------------
// Thread (g, w): selected individual g, SNP group w (SNPs 64w .. 64w+63 of the block).
// src holds the block's SNP records (record s at src + s*src_bpr); individual sel[g]'s
// code of SNP s sits in byte sel[g]/4 at position sel[g]%4 (MSB first).  The 64 codes go
// to bytes 16w .. 16w+15 of the individual's canonical record (offset col_off), first SNP
// in the two most significant bits.  Codes past n_snp are 0.
__global__ __launch_bounds__(256) void transpose_gather_kernel(const std::uint8_t* __restrict__ src,
                                                               std::int64_t src_bpr, std::int64_t n_snp,
                                                               const std::int32_t* __restrict__ sel, std::int64_t n_sel,
                                                               std::uint8_t* __restrict__ out, std::int64_t out_bpr,
                                                               std::int64_t col_off, std::int64_t n_groups, int bed) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t g = t % n_sel;
    const std::int64_t w = t / n_sel;
    if (w >= n_groups) return;
    const std::int32_t row = sel[g];
    const std::uint8_t* col = src + (row >> 2);
    // GENO: individual r in bits 6-2(r%4) (first in the top bits).  PLINK .bed: bits
    // 2(r%4), codes 00 / 01 / 10 / 11 = 2 / missing / 1 / 0 copies, mapped through the
    // 2-bit table 0x1E (entry c at bits 2c).
    const int shift = bed ? 2 * (row & 3) : 6 - 2 * (row & 3);
    const std::uint32_t table = bed ? 0x1Eu : 0xE4u;   // 0xE4: the identity
    std::uint32_t word[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int j = 0; j < 64; ++j) {
        const std::int64_t s = 64 * w + j;
        const std::uint32_t raw = s < n_snp ? (col[s * src_bpr] >> shift) & 3u : 3u;
        const std::uint32_t code = s < n_snp ? (table >> (2 * raw)) & 3u : 0u;
        // output byte j/4 (within the 16), position j%4 from the top; bytes little-endian in a word
        word[j >> 4] |= code << (8 * ((j >> 2) & 3) + 6 - 2 * (j & 3));
    }
    uint4* dst = reinterpret_cast<uint4*>(out + g * out_bpr + col_off + 16 * w);
    *dst = make_uint4(word[0], word[1], word[2], word[3]);
}

// ---- keep decision ----------------------------------------------------------------------
__device__ inline char norm_allele(char a) {
    const char u = (a >= 'a' && a <= 'z') ? static_cast<char>(a - 0x20) : a;
    return (u == 'A' || u == 'C' || u == 'G' || u == 'T') ? u : '\0';
}
__device__ inline char complement_of(char a) {
    switch (norm_allele(a)) {
        case 'A': return 'T';
        case 'T': return 'A';
        case 'C': return 'G';
        case 'G': return 'C';
        default: return '\0';
    }
}
__device__ inline bool multiallelic(char a, char b) {
    const char x = norm_allele(a), y = norm_allele(b);
    return x == '\0' || y == '\0' || x == y;
}
__device__ inline bool strand_ambiguous(char a, char b) {
    const char x = norm_allele(a), y = norm_allele(b);
    if (x == '\0' || y == '\0') return false;
    return x == complement_of(y);
}
__device__ inline bool transversion(char a, char b) {
    if (multiallelic(a, b)) return false;
    const char x = norm_allele(a), y = norm_allele(b);
    const bool px = x == 'A' || x == 'G', py = y == 'A' || y == 'G';
    return px != py;
}

// One thread per SNP; populations summed in order.  q = a2 / (2n) is the decoded
// allele frequency (a2 = 2 x allele count), n the allele count.
__global__ __launch_bounds__(256) void snp_keep_kernel(const std::uint16_t* __restrict__ a2,
                                                       const std::uint16_t* __restrict__ nn, std::int64_t ld, int P,
                                                       std::int64_t M, const char* __restrict__ ref,
                                                       const char* __restrict__ alt, const std::int32_t* __restrict__ chrom,
                                                       KeepRule rule, std::uint8_t* __restrict__ flags) {
    const std::int64_t s = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= M) return;
    double ref_count = 0.0, allele_count = 0.0, nonmissing = 0.0;
    int missing_pops = 0;
    for (int p = 0; p < P; ++p) {
        const std::int64_t o = static_cast<std::int64_t>(p) * ld + s;
        const unsigned c = nn[o];
        const double n = static_cast<double>(c);
        const double q = c ? static_cast<double>(a2[o]) / (2.0 * n) : 0.0;
        ref_count = ref_count + q * n;
        allele_count += n;
        nonmissing += n / rule.ploidy;
        missing_pops += c == 0;
    }
    const double ref_af = allele_count > 0.0 ? ref_count / allele_count : 0.0;
    const double minor_af = ref_af < 1.0 - ref_af ? ref_af : 1.0 - ref_af;
    const double missing = rule.total_individuals > 0.0 ? 1.0 - nonmissing / rule.total_individuals : 1.0;
    const char r = ref[s], a = alt[s];
    bool keep = !multiallelic(r, a);
    keep = keep && !(rule.strand_drop && strand_ambiguous(r, a));
    keep = keep && minor_af >= rule.maf_min;
    keep = keep && missing <= rule.snp_max_missing;
    keep = keep && !(rule.drop_monomorphic && minor_af == 0.0);
    keep = keep && !(rule.transversions_only && !transversion(r, a));
    keep = keep && !(rule.autosomes_only && !(chrom[s] >= 1 && chrom[s] <= 22));
    if (keep && rule.maxmiss < 1.0) {
        const double frac = static_cast<double>(missing_pops) / static_cast<double>(P);
        if (frac > rule.maxmiss) keep = false;
    }
    flags[s] = keep ? 1 : 0;
}

// ---- compaction -----------------------------------------------------------------------
// Thread (k, p): output column k of population p takes source column kept[k].
__global__ __launch_bounds__(256) void compact_counts_kernel(const std::uint16_t* __restrict__ a2,
                                                             const std::uint16_t* __restrict__ nn, std::int64_t ld,
                                                             const std::int32_t* __restrict__ kept, std::int64_t M_kept,
                                                             int P, std::uint16_t* __restrict__ a2_out,
                                                             std::uint16_t* __restrict__ n_out, std::int64_t ld_out) {
    const std::int64_t k = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (k >= ld_out) return;
    const std::int64_t o = static_cast<std::int64_t>(p) * ld_out + k;
    if (k < M_kept) {
        const std::int64_t i = static_cast<std::int64_t>(p) * ld + kept[k];
        a2_out[o] = a2[i];
        n_out[o] = nn[i];
    } else {
        a2_out[o] = 0;
        n_out[o] = 0;
    }
}

}  // namespace

void launch_transpose_gather(const std::uint8_t* src, std::int64_t src_bpr, std::int64_t n_snp, const std::int32_t* sel,
                             std::int64_t n_sel, std::uint8_t* out, std::int64_t out_bpr, std::int64_t col_off,
                             bool bed, hipStream_t st) {
    if (n_snp <= 0 || n_sel <= 0) return;
    const std::int64_t groups = ceil_div(n_snp, 64);
    STP_REQUIRE(out_bpr % 16 == 0 && col_off % 16 == 0, "transpose: output stride and offset must be 16-byte aligned");
    STP_REQUIRE(col_off + 16 * groups <= out_bpr, "transpose: block exceeds the output record");
    const std::int64_t threads = groups * n_sel;
    STP_REQUIRE(ceil_div(threads, 256) <= 0x7fffffffLL, "transpose: block too large for one launch");
    transpose_gather_kernel<<<static_cast<unsigned>(ceil_div(threads, 256)), 256, 0, st>>>(
        src, src_bpr, n_snp, sel, n_sel, out, out_bpr, col_off, groups, bed ? 1 : 0);
    STP_LAUNCH_CHECK();
}

void launch_snp_keep(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, std::int64_t M,
                     const char* ref, const char* alt, const std::int32_t* chrom, const KeepRule& rule,
                     std::uint8_t* flags, hipStream_t st) {
    if (M <= 0) return;
    snp_keep_kernel<<<static_cast<unsigned>(ceil_div(M, 256)), 256, 0, st>>>(a2, n, ld, P, M, ref, alt, chrom, rule,
                                                                             flags);
    STP_LAUNCH_CHECK();
}

void launch_compact_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, const std::int32_t* kept,
                           std::int64_t M_kept, int P, std::uint16_t* a2_out, std::uint16_t* n_out, std::int64_t ld_out,
                           hipStream_t st) {
    if (P <= 0 || ld_out <= 0) return;
    STP_REQUIRE(P <= 65535, "compact: too many populations for one launch");
    const dim3 grid(static_cast<unsigned>(ceil_div(ld_out, 256)), static_cast<unsigned>(P));
    compact_counts_kernel<<<grid, 256, 0, st>>>(a2, n, ld, kept, M_kept, P, a2_out, n_out, ld_out);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
