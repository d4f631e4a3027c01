// steppe_amd/csrc/synth.hip — on-device synthetic genotype panels.
//
// Benchmark/test input generator (no counterpart in the reference, whose tests read
// real AADR data).  Produces the packed 2-bit individual-major tile directly in HBM
// with the same distributions as steppe_amd/synth.py:
//   p0_s ~ U(0.05, 0.95);  p_{pop,s} ~ Beta(p0 (1-F)/F, (1-p0)(1-F)/F)  (Balding-Nichols)
//   diploid code ~ Binomial(2, p); pseudo-haploid code = 2 * Bernoulli(p); missing at the
//   population's rate.
// Randomness is counter-based (splitmix64 of (seed, stream, indices)), so a panel is
// a pure function of its parameters and independent of the launch shape.  Beta
// draws use two Marsaglia-Tsang gamma variates (boosted by U^(1/a) for a < 1).

#include "common.hpp"

namespace stp {
namespace {

__device__ inline std::uint64_t mix64(std::uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ inline std::uint64_t key(std::uint64_t seed, std::uint64_t stream, std::uint64_t a,
                                    std::uint64_t b) {
    return mix64(mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ a) ^ (b * 0x9E3779B97F4A7C15ull));
}

// Uniform in (0, 1): 53 random bits, never exactly 0.
__device__ inline double unif(std::uint64_t h) {
    return (static_cast<double>(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

__device__ double gamma_draw(double a, std::uint64_t k) {
    double boost = 1.0;
    if (a < 1.0) {
        boost = pow(unif(mix64(k ^ 0x5151ull)), 1.0 / a);
        a += 1.0;
    }
    const double d = a - 1.0 / 3.0, c = 1.0 / sqrt(9.0 * d);
    for (std::uint64_t it = 1;; ++it) {
        const std::uint64_t h = mix64(k + it * 0x2545F4914F6CDD1Dull);
        const double u1 = unif(h), u2 = unif(mix64(h));
        const double x = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        double v = 1.0 + c * x;
        if (v <= 0.0) continue;
        v = v * v * v;
        const double u = unif(mix64(h ^ 0xABCDull));
        if (log(u) < 0.5 * x * x + d - d * v + d * log(v)) return boost * d * v;
        if (it > 1000) return boost * d;  // unreachable in practice; keeps the loop bounded
    }
}

// One thread per (population, SNP): Balding-Nichols frequency as float.
__global__ void freq_kernel(std::uint64_t seed, int P, std::int64_t M, const float* __restrict__ fst,
                            float* __restrict__ freq) {
    const std::int64_t idx = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<std::int64_t>(P) * M) return;
    const int p = static_cast<int>(idx / M);
    const std::int64_t s = idx - static_cast<std::int64_t>(p) * M;
    const double p0 = 0.05 + 0.9 * unif(key(seed, 1, static_cast<std::uint64_t>(s), 0));
    const double F = fst[p];
    const double a = p0 * (1.0 - F) / F, b = (1.0 - p0) * (1.0 - F) / F;
    const double x = gamma_draw(a, key(seed, 2, static_cast<std::uint64_t>(p), static_cast<std::uint64_t>(s)));
    const double y = gamma_draw(b, key(seed, 3, static_cast<std::uint64_t>(p), static_cast<std::uint64_t>(s)));
    freq[idx] = static_cast<float>(x / (x + y));
}

// One thread per (individual, 16-SNP dword) of the packed record.
__global__ void genotype_kernel(std::uint64_t seed, std::int64_t n_ind, std::int64_t M, std::int64_t bpr,
                                const std::int32_t* __restrict__ ind_pop,
                                const std::int8_t* __restrict__ ind_haploid,
                                const float* __restrict__ pop_missing, const float* __restrict__ freq,
                                std::uint8_t* __restrict__ packed) {
    const std::int64_t dwords = bpr / 4;
    const std::int64_t idx = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= n_ind * dwords) return;
    const std::int64_t g = idx / dwords;
    const std::int64_t d = idx - g * dwords;
    const int p = ind_pop[g];
    const bool hap = ind_haploid[g] != 0;
    const double miss = pop_missing[p];
    std::uint32_t word = 0;
    for (int j = 0; j < 16; ++j) {
        const std::int64_t s = 16 * d + j;
        if (s >= M) break;
        const double f = freq[static_cast<std::int64_t>(p) * M + s];
        const std::uint64_t h = key(seed, 4, static_cast<std::uint64_t>(g), static_cast<std::uint64_t>(s));
        const double um = unif(h);
        unsigned code;
        if (um < miss) {
            code = 3;
        } else if (hap) {
            code = unif(mix64(h ^ 0x11ull)) < f ? 2u : 0u;
        } else {
            code = (unif(mix64(h ^ 0x22ull)) < f ? 1u : 0u) + (unif(mix64(h ^ 0x33ull)) < f ? 1u : 0u);
        }
        // SNP j of the dword: byte j/4 (little-endian), shift 6 - 2*(j%4).
        word |= code << (8 * (j >> 2) + 6 - 2 * (j & 3));
    }
    reinterpret_cast<std::uint32_t*>(packed + g * bpr)[d] = word;
}

}  // namespace

void launch_synth_panel(std::uint64_t seed, int P, std::int64_t M, std::int64_t n_ind, std::int64_t bpr,
                        const float* fst, const float* pop_missing, const std::int32_t* ind_pop,
                        const std::int8_t* ind_haploid, float* freq_scratch, std::uint8_t* packed,
                        hipStream_t st) {
    STP_REQUIRE(bpr % 4 == 0 && bpr >= (M + 3) / 4, "synth: bytes_per_record must be a multiple of 4 and >= ceil(M/4)");
    const std::int64_t nf = static_cast<std::int64_t>(P) * M;
    freq_kernel<<<static_cast<unsigned>(ceil_div(nf, 256)), 256, 0, st>>>(seed, P, M, fst, freq_scratch);
    STP_LAUNCH_CHECK();
    const std::int64_t ng = n_ind * (bpr / 4);
    genotype_kernel<<<static_cast<unsigned>(ceil_div(ng, 256)), 256, 0, st>>>(
        seed, n_ind, M, bpr, ind_pop, ind_haploid, pop_missing, freq_scratch, packed);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
