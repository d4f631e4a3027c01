// oracle/cpu_ref.cpp — CPU restatement of the Steppe CpuBackend arithmetic for the
// f-statistics hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP product
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
// it, and only to compute expected values or to time the CPU baseline.  Nothing in
// steppe_amd/ links or calls it; the product path fails loudly without its own HIP
// library instead of falling back here.
//
// Every function restates one reference function, written fresh from its
// semantics (file:line cites are relative to the reference repository root):
//
//   cref_detect_ploidy     src/device/cpu/cpu_backend.cpp:73-95   (detect_ploidy_host)
//   cref_decode_counts     src/device/cpu/cpu_backend.cpp:267-322 (decode_af) with
//                          src/core/internal/decode_af.hpp:38-97, returning the exact
//                          integers 2*ac and n instead of q/v/n
//   cref_decode_af         the same loop, finishing with finalize_af_counts
//   cref_assign_blocks     src/core/domain/block_partition_rule.cpp:14-87
//   cref_block_ranges      src/core/domain/block_partition_rule.hpp:54-96
//   cref_f2_blocks         src/device/cpu/cpu_backend.cpp:155-174,217-265 and
//                          src/core/internal/f2_estimator.hpp:22-51, with the
//                          long-double pairwise cascade of cpu_backend.cpp:60-70
//   qpAdm functions        see the second half of this file
//
// The long-double pairwise cascade (base case 128, split at n/2) and the per-SNP
// term order are kept identical to the reference because the FP64 bits of the
// oracle depend on them; everything else is free-standing code.
//
// Parity of this restatement is pinned three ways (tests/test_oracle_*.py):
// bit-identity against the real reference CpuBackend compiled into oracle/_ref
// (where /root/reference is present), exact-rational brute force on tiny inputs,
// and the reference's own unit KATs (tests/unit/test_decode.cpp, test_f2.cpp).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no -march, so no FMA).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

constexpr int kMissing = 3;
constexpr int kPloidyWindow = 1000;  // SNPs scanned by the ploidy pre-pass
constexpr std::size_t kCascadeLeaf = 128;

inline int code_at(const std::uint8_t* rec, std::size_t s) {
    // Four codes per byte, first SNP in the two most significant bits.
    const unsigned byte = rec[s >> 2];
    const unsigned shift = 6u - 2u * static_cast<unsigned>(s & 3u);
    return static_cast<int>((byte >> shift) & 3u);
}

long double cascade_sum(const long double* a, std::size_t n) {
    if (n == 0) return 0.0L;
    if (n <= kCascadeLeaf) {
        long double s = 0.0L;
        for (std::size_t k = 0; k < n; ++k) s += a[k];
        return s;
    }
    const std::size_t half = n / 2;
    return cascade_sum(a, half) + cascade_sum(a + half, n - half);
}

double het_bias(double q, double n) {
    // q(1-q)/max(n-1, 1)
    const double d = n - 1.0;
    return q * (1.0 - q) / (d > 1.0 ? d : 1.0);
}

struct Tile {
    const std::uint8_t* packed;
    std::size_t bpr, n_snp, n_ind;
};

std::vector<int> ploidy_scan(const Tile& t) {
    std::vector<int> out(t.n_ind, 1);
    const std::size_t w = std::min<std::size_t>(kPloidyWindow, t.n_snp);
    for (std::size_t g = 0; g < t.n_ind; ++g) {
        const std::uint8_t* rec = t.packed + g * t.bpr;
        for (std::size_t s = 0; s < w; ++s)
            if (code_at(rec, s) == 1) { out[g] = 2; break; }
    }
    return out;
}

// Effective ploidy of every sample: explicit array, else detection, else the
// tile-wide scalar (matches DecodeTileView's three-way choice).
std::vector<int> effective_ploidy(const Tile& t, const std::int32_t* sample_ploidy,
                                  int ploidy, int detect) {
    if (sample_ploidy) return std::vector<int>(sample_ploidy, sample_ploidy + t.n_ind);
    if (detect) return ploidy_scan(t);
    return std::vector<int>(t.n_ind, ploidy);
}

}  // namespace

extern "C" {

int cref_detect_ploidy(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                       std::size_t n_ind, std::int32_t* out) {
    const Tile t{packed, bpr, n_snp, n_ind};
    const std::vector<int> p = ploidy_scan(t);
    for (std::size_t g = 0; g < n_ind; ++g) out[g] = p[g];
    return 0;
}

// Integer decode: out_a2 = 2*ac, out_n = n, both [P x M] at i + P*s.
// ac accumulates code/(3-ploidy) for ploidy in {1,2} and valid codes; so 2*ac is
// the integer sum of 2*code (diploid) or code (pseudo-haploid).
int cref_decode_counts(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                       std::size_t n_ind, const std::uint64_t* pop_offsets, int n_pop,
                       const std::int32_t* sample_ploidy, int ploidy, int detect,
                       std::int32_t* out_a2, std::int32_t* out_n) {
    const Tile t{packed, bpr, n_snp, n_ind};
    const std::vector<int> pl = effective_ploidy(t, sample_ploidy, ploidy, detect);
    const std::size_t P = static_cast<std::size_t>(n_pop);
    for (std::size_t i = 0; i < P; ++i) {
        for (std::size_t s = 0; s < n_snp; ++s) {
            long a2 = 0, n = 0;
            for (std::uint64_t g = pop_offsets[i]; g < pop_offsets[i + 1]; ++g) {
                const int c = code_at(packed + g * bpr, s);
                const int p = pl[g];
                if (c == kMissing || (p != 1 && p != 2)) continue;
                a2 += (p == 2) ? 2 * c : c;
                n += p;
            }
            out_a2[i + P * s] = static_cast<std::int32_t>(a2);
            out_n[i + P * s] = static_cast<std::int32_t>(n);
        }
    }
    return 0;
}

// Floating decode, bit-for-bit what CpuBackend::decode_af produces: ac summed in
// double as code/(3-ploidy), q = ac/n, v = 1, N = n; all zero when n == 0.
int cref_decode_af(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                   std::size_t n_ind, const std::uint64_t* pop_offsets, int n_pop,
                   const std::int32_t* sample_ploidy, int ploidy, int detect,
                   double* out_q, double* out_v, double* out_n) {
    const Tile t{packed, bpr, n_snp, n_ind};
    const std::vector<int> pl = effective_ploidy(t, sample_ploidy, ploidy, detect);
    const std::size_t P = static_cast<std::size_t>(n_pop);
    for (std::size_t i = 0; i < P; ++i) {
        for (std::size_t s = 0; s < n_snp; ++s) {
            double ac = 0.0;
            std::int64_t n = 0;
            for (std::uint64_t g = pop_offsets[i]; g < pop_offsets[i + 1]; ++g) {
                const int c = code_at(packed + g * bpr, s);
                const int p = pl[g];
                if (c == kMissing || (p != 1 && p != 2)) continue;
                ac += static_cast<double>(c) / (3.0 - static_cast<double>(p));
                n += p;
            }
            const std::size_t o = i + P * s;
            if (n > 0) {
                out_n[o] = static_cast<double>(n);
                out_q[o] = ac / static_cast<double>(n);
                out_v[o] = 1.0;
            } else {
                out_q[o] = out_v[o] = out_n[o] = 0.0;
            }
        }
    }
    return 0;
}

// Jackknife block walk: a new block starts on a chromosome change or once the
// position has moved >= window past the block's first SNP.  Falls back to the
// physical axis (window bp_window) when every genetic position is zero.
// Returns n_block (0 for an invalid block size or M <= 0).
int cref_assign_blocks(const std::int32_t* chrom, const double* genpos, std::int64_t M,
                       double blgsize, const double* physpos, double bp_window,
                       std::int32_t* out_block_id) {
    if (!(blgsize > 0.0) || M <= 0) return 0;
    const double* pos = genpos;
    double window = blgsize;
    if (physpos && bp_window > 0.0) {
        bool gen_zero = true, phys_zero = true;
        for (std::int64_t s = 0; s < M; ++s) {
            gen_zero = gen_zero && genpos[s] == 0.0;
            phys_zero = phys_zero && physpos[s] == 0.0;
        }
        if (gen_zero && !phys_zero) { pos = physpos; window = bp_window; }
    }
    long blk = -1;
    int cur_chrom = -1;
    double start = -1.0e20;
    for (std::int64_t s = 0; s < M; ++s) {
        if (chrom[s] != cur_chrom || pos[s] - start >= window) {
            ++blk;
            cur_chrom = chrom[s];
            start = pos[s];
        }
        out_block_id[s] = static_cast<std::int32_t>(blk);
    }
    return static_cast<int>(blk + 1);
}

// Contiguous [begin, end) per block; -1 when block_id is out of range or not
// non-decreasing (the reference throws in both cases).
int cref_block_ranges(const std::int32_t* block_id, std::int64_t M, int nb,
                      std::int64_t* begin, std::int64_t* end) {
    for (int b = 0; b < nb; ++b) begin[b] = end[b] = 0;
    int prev = -1;
    std::int64_t s = 0;
    while (s < M) {
        const int b = block_id[s];
        if (b < 0 || b >= nb || b < prev) return -1;
        std::int64_t e = s;
        while (e < M && block_id[e] == b) ++e;
        begin[b] = s;
        end[b] = e;
        prev = b;
        s = e;
    }
    return 0;
}

// Per-block f2 over Q/V/N [P x M] (i + P*s).  For every block and pair i <= j:
// long-double terms (q_i-q_j)^2 - hc_i - hc_j over the jointly valid SNPs, in SNP
// order, summed by the pairwise cascade; f2 = num/count (0 when count == 0),
// vpair = count; mirrored to (j,i).  n_threads > 1 splits the (block, i) rows
// across threads; the arithmetic per entry is unchanged, so results are
// identical for any thread count.
int cref_f2_blocks(const double* Q, const double* V, const double* N, int P, std::int64_t M,
                   const std::int32_t* block_id, int nb, double* f2, double* vpair,
                   std::int32_t* block_sizes, int n_threads) {
    if (P <= 0 || M <= 0 || nb <= 0) return 0;
    std::vector<std::int64_t> b0(nb), b1(nb);
    if (cref_block_ranges(block_id, M, nb, b0.data(), b1.data()) != 0) return -1;
    for (int b = 0; b < nb; ++b) block_sizes[b] = static_cast<std::int32_t>(b1[b] - b0[b]);
    const std::size_t Pz = static_cast<std::size_t>(P);
    const std::size_t slab = Pz * Pz;

    std::int64_t max_len = 0;
    for (int b = 0; b < nb; ++b) max_len = std::max(max_len, b1[b] - b0[b]);

    const std::int64_t rows = static_cast<std::int64_t>(nb) * P;
    std::atomic<std::int64_t> next{0};
    auto worker = [&]() {
        std::vector<long double> terms(static_cast<std::size_t>(max_len) + 1);
        for (;;) {
            const std::int64_t r = next.fetch_add(1);
            if (r >= rows) break;
            const int b = static_cast<int>(r / P);
            const int i = static_cast<int>(r % P);
            const std::size_t base = slab * static_cast<std::size_t>(b);
            for (int j = i; j < P; ++j) {
                std::size_t cnt = 0;
                for (std::int64_t s = b0[b]; s < b1[b]; ++s) {
                    const std::size_t oi = static_cast<std::size_t>(i) + Pz * s;
                    const std::size_t oj = static_cast<std::size_t>(j) + Pz * s;
                    if (V[oi] == 0.0 || V[oj] == 0.0) continue;
                    const double pi = Q[oi], pj = Q[oj];
                    const double d = pi - pj;
                    const double t = d * d - het_bias(pi, N[oi]) - het_bias(pj, N[oj]);
                    terms[cnt++] = static_cast<long double>(t);
                }
                const double num = static_cast<double>(cascade_sum(terms.data(), cnt));
                const double c = static_cast<double>(cnt);
                const double val = c > 0.0 ? num / c : 0.0;
                const std::size_t ij = base + static_cast<std::size_t>(i) + Pz * j;
                const std::size_t ji = base + static_cast<std::size_t>(j) + Pz * i;
                f2[ij] = f2[ji] = val;
                vpair[ij] = vpair[ji] = c;
            }
        }
    };
    const int nt = n_threads > 1 ? n_threads : 1;
    if (nt == 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t) pool.emplace_back(worker);
        for (auto& th : pool) th.join();
    }
    return 0;
}

}  // extern "C"
