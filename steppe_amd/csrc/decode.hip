// steppe_amd/csrc/decode.hip — packed 2-bit genotypes -> per-population counts.
//
// Replaces the reference's decode_af_kernel (src/device/cuda/decode_af_kernel.cu:32-87)
// and detect_ploidy_kernel (src/device/cuda/detect_ploidy_kernel.cu:32-58), with the
// semantics of src/core/internal/decode_af.hpp:38-97.  Instead of FP64 q/v/n (24 B
// per population-SNP) the kernel emits the two exact integers the frequency is
// built from:
//     a2 = 2*ac = sum over valid genotypes of 2*code (diploid) or code (pseudo-haploid)
//     n  = sum of ploidy over valid genotypes
// so q = a2 / (2n) reproduces the reference's q = ac/n bit for bit (one correctly
// rounded division of the same rational), v = (n > 0), N = n.
//
// Input layout (the reference's DecodeTileView): individual-major records of
// bytes_per_record bytes, four 2-bit codes per byte, first SNP in bits 7:6,
// code 3 = missing; individuals of population p are [pop_offsets[p], pop_offsets[p+1]).
//
// Output layout: a2, n as uint16 [P][ld] (population-major, SNP-contiguous rows,
// ld a multiple of 16, entries for SNPs >= n_snp are zero).
//
// Decode scheme: one wavefront owns a 1024-SNP column tile (256 contiguous bytes
// of every record: one coalesced 4-byte load per lane per individual) and walks the
// individuals of one population at a time.  Each lane keeps 16 SNPs in SWAR form:
// a dword's sixteen codes are split into four byte-lane vectors, missing codes are
// masked with two ANDs, and counts accumulate in byte lanes (flushed to 16-bit
// lanes every 60 individuals), so decoding costs a few integer ops per 16 genotypes
// and the kernel runs at HBM speed.

#include "common.hpp"
#include "engine.hpp"

namespace stp {
namespace {

constexpr int kSnpPerLane = 16;
constexpr int kLanes = 64;
constexpr int kTileSnps = kSnpPerLane * kLanes;  // 1024
constexpr int kTileBytes = kTileSnps / 4;        // 256
constexpr int kFlushEvery = 60;                  // 60 * 4 <= 255 per byte lane
constexpr int kBatch = 8;                        // records loaded ahead per wavefront
constexpr int kPloidyWindow = 1000;

template <bool kAligned>
__device__ inline std::uint32_t load_codes(const std::uint8_t* __restrict__ rec, std::int64_t boff,
                                           std::int64_t bpr) {
    if constexpr (kAligned) {
        // bpr % 4 == 0 and boff % 4 == 0: the whole dword lies in the record.
        return (boff < bpr) ? *reinterpret_cast<const std::uint32_t*>(rec + boff) : 0u;
    } else {
        std::uint32_t w = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (boff + b < bpr) w |= static_cast<std::uint32_t>(rec[boff + b]) << (8 * b);
        return w;
    }
}

template <bool kAligned>
__global__ __launch_bounds__(256) void decode_counts_kernel(
    const std::uint8_t* __restrict__ packed, std::int64_t bpr, std::int64_t n_snp,
    const std::uint64_t* __restrict__ pop_offsets, const std::int8_t* __restrict__ ploidy,
    int n_pop, std::uint16_t* __restrict__ out_a2, std::uint16_t* __restrict__ out_n,
    std::int64_t ld) {
    const int lane = threadIdx.x & 63;
    // wavefront-uniform: population offsets and ploidies then come through scalar loads
    const int slot = __builtin_amdgcn_readfirstlane(static_cast<int>(blockIdx.y * (blockDim.x / 64) + (threadIdx.x >> 6)));
    const int n_slots = gridDim.y * (blockDim.x / 64);
    const std::int64_t tile = blockIdx.x;
    const std::int64_t boff = tile * kTileBytes + 4 * lane;
    const std::int64_t s0 = tile * kTileSnps + kSnpPerLane * lane;

    for (int p = slot; p < n_pop; p += n_slots) {
        const std::uint64_t g0 = pop_offsets[p], g1 = pop_offsets[p + 1];
        std::uint32_t a8[4] = {0, 0, 0, 0}, n8[4] = {0, 0, 0, 0};
        std::uint32_t alo[4] = {0, 0, 0, 0}, ahi[4] = {0, 0, 0, 0};
        std::uint32_t nlo[4] = {0, 0, 0, 0}, nhi[4] = {0, 0, 0, 0};
        int pending = 0;
        // Individuals in batches of kBatch: all of a batch's loads are issued before any
        // is decoded, so each wavefront keeps kBatch records' words in flight.
        for (std::uint64_t gb = g0; gb < g1; gb += kBatch) {
            // lane u < kBatch fetches the ploidy of individual gb + u (one load for the batch)
            const int myp = (lane < kBatch && gb + lane < g1) ? ploidy[gb + lane] : 0;
            std::uint32_t wb[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const std::uint64_t g = gb + u;
                wb[u] = g < g1 ? load_codes<kAligned>(packed + g * bpr, boff, bpr) : 0u;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int pl = __builtin_amdgcn_readlane(myp, u);
                if (pl != 1 && pl != 2) continue;  // the reference skips any other ploidy
                const std::uint32_t w = wb[u];
                const int sh = (pl == 2) ? 1 : 0;  // a2 += 2*code, n += 2 for diploids
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const std::uint32_t x = (w >> (2 * k)) & 0x03030303u;
                    const std::uint32_t miss = x & (x >> 1) & 0x01010101u;  // code == 3
                    const std::uint32_t ok = miss ^ 0x01010101u;
                    const std::uint32_t code = x & ~(miss * 3u);
                    a8[k] += code << sh;
                    n8[k] += ok << sh;
                }
                if (++pending == kFlushEvery) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        alo[k] += a8[k] & 0x00FF00FFu;
                        ahi[k] += (a8[k] >> 8) & 0x00FF00FFu;
                        nlo[k] += n8[k] & 0x00FF00FFu;
                        nhi[k] += (n8[k] >> 8) & 0x00FF00FFu;
                        a8[k] = n8[k] = 0;
                    }
                    pending = 0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            alo[k] += a8[k] & 0x00FF00FFu;
            ahi[k] += (a8[k] >> 8) & 0x00FF00FFu;
            nlo[k] += n8[k] & 0x00FF00FFu;
            nhi[k] += (n8[k] >> 8) & 0x00FF00FFu;
        }
        if (s0 >= ld) continue;
        // Lane SNP j sits in byte j/4 at shift 6 - 2*(j%4), i.e. in vector
        // k = 3 - j%4, byte lane j/4: even byte lanes in *lo, odd in *hi.
        std::uint32_t pa[8], pn[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            std::uint32_t va[2], vn[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int j = 2 * m + h;
                const int bl = j >> 2, k = 3 - (j & 3);
                const std::uint32_t ra = (bl & 1) ? ahi[k] : alo[k];
                const std::uint32_t rn = (bl & 1) ? nhi[k] : nlo[k];
                const int sh16 = 16 * (bl >> 1);
                const bool live = (s0 + j) < n_snp;
                va[h] = live ? ((ra >> sh16) & 0xFFFFu) : 0u;
                vn[h] = live ? ((rn >> sh16) & 0xFFFFu) : 0u;
            }
            pa[m] = va[0] | (va[1] << 16);
            pn[m] = vn[0] | (vn[1] << 16);
        }
        uint4* da = reinterpret_cast<uint4*>(out_a2 + static_cast<std::int64_t>(p) * ld + s0);
        uint4* dn = reinterpret_cast<uint4*>(out_n + static_cast<std::int64_t>(p) * ld + s0);
        da[0] = make_uint4(pa[0], pa[1], pa[2], pa[3]);
        da[1] = make_uint4(pa[4], pa[5], pa[6], pa[7]);
        dn[0] = make_uint4(pn[0], pn[1], pn[2], pn[3]);
        dn[1] = make_uint4(pn[4], pn[5], pn[6], pn[7]);
    }
}

// Per-sample ploidy: 2 if any heterozygote (code 1) occurs among the first
// min(1000, n_snp) SNPs, else 1 (src/device/cpu/cpu_backend.cpp:73-95).
// One wavefront per individual; lanes scan disjoint bytes.
__global__ __launch_bounds__(256) void detect_ploidy_kernel(const std::uint8_t* __restrict__ packed,
                                                            std::int64_t bpr, std::int64_t n_snp,
                                                            std::int64_t n_ind,
                                                            std::int8_t* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const std::int64_t g = static_cast<std::int64_t>(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (g >= n_ind) return;
    const std::int64_t window = n_snp < kPloidyWindow ? n_snp : kPloidyWindow;
    const std::int64_t nbytes = (window + 3) / 4;
    const std::uint8_t* rec = packed + g * bpr;
    bool het = false;
    for (std::int64_t b = lane; b < nbytes; b += 64) {
        const unsigned byte = rec[b];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const std::int64_t s = 4 * b + j;
            if (s < window && ((byte >> (6 - 2 * j)) & 3u) == 1u) het = true;
        }
    }
    const bool any = __any(het);
    if (lane == 0) out[g] = any ? 2 : 1;
}

__global__ void fill_ploidy_kernel(std::int8_t* out, std::int64_t n, int value) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<std::int8_t>(value);
}

// Counts -> the reference's floating DecodeResult: q = a2/(2n) (== ac/n), v, N,
// written column-major [P x M] at i + P*s.
__global__ void counts_to_qvn_kernel(const std::uint16_t* __restrict__ a2, const std::uint16_t* __restrict__ nn,
                                     std::int64_t ld, int P, std::int64_t M, double* __restrict__ q,
                                     double* __restrict__ v, double* __restrict__ n) {
    const std::int64_t idx = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= static_cast<std::int64_t>(P) * M) return;
    const std::int64_t s = idx / P;
    const int p = static_cast<int>(idx - s * P);
    const unsigned c = nn[static_cast<std::int64_t>(p) * ld + s];
    if (c) {
        const double nd = static_cast<double>(c);
        q[idx] = static_cast<double>(a2[static_cast<std::int64_t>(p) * ld + s]) / (2.0 * nd);
        v[idx] = 1.0;
        n[idx] = nd;
    } else {
        q[idx] = v[idx] = n[idx] = 0.0;
    }
}

// Per-sample ploidy from the caller's int32 values: 1 or 2 as given, anything else 0.
__global__ void ploidy_from_i32_kernel(const std::int32_t* __restrict__ in, std::int64_t n,
                                       std::int8_t* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const int v = in[i];
        out[i] = static_cast<std::int8_t>((v == 1 || v == 2) ? v : 0);
    }
}

}  // namespace

// ---- scratch of the allocation-free decode (stp_decode_counts_noalloc) -------------------
// Device copies of the population offsets and the per-sample ploidy, in a fixed array that
// every GPU the library runs on holds until the process ends.  Layout: offsets
// (uint64, kNoallocMaxPops + 1), then ploidy (int8, kNoallocMaxIndividuals), then the
// caller's int32 ploidy values staged for conversion (kNoallocMaxIndividuals int32).
constexpr std::size_t kNoallocOffsBytes = (8 * (static_cast<std::size_t>(kNoallocMaxPops) + 1) + 255) / 256 * 256;
constexpr std::size_t kNoallocBytes =
    kNoallocOffsBytes + static_cast<std::size_t>(kNoallocMaxIndividuals) * (1 + 4);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

NoallocScratch resolve_noalloc_scratch() {
    void* ptr = nullptr;
    STP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    NoallocScratch s;
    auto* base = static_cast<unsigned char*>(ptr);
    s.offs = reinterpret_cast<std::uint64_t*>(base);
    s.ploidy = reinterpret_cast<std::int8_t*>(base + kNoallocOffsBytes);
    s.ploidy_i32 = reinterpret_cast<std::int32_t*>(base + kNoallocOffsBytes + kNoallocMaxIndividuals);
    s.bytes = kNoallocBytes;
    return s;
}

void launch_ploidy_from_i32(const std::int32_t* in, std::int64_t n, std::int8_t* out, hipStream_t st) {
    if (n <= 0) return;
    ploidy_from_i32_kernel<<<static_cast<unsigned>(ceil_div(n, 256)), 256, 0, st>>>(in, n, out);
    STP_LAUNCH_CHECK();
}

void launch_counts_to_qvn(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                          std::int64_t M, double* q, double* v, double* nout, hipStream_t st) {
    const std::int64_t tot = static_cast<std::int64_t>(P) * M;
    if (tot <= 0) return;
    counts_to_qvn_kernel<<<static_cast<unsigned>(ceil_div(tot, 256)), 256, 0, st>>>(a2, n, ld, P, M, q, v, nout);
    STP_LAUNCH_CHECK();
}

// Launchers (host).  All pointers are device pointers; the caller owns them.
void launch_detect_ploidy(const std::uint8_t* packed, std::int64_t bpr, std::int64_t n_snp,
                          std::int64_t n_ind, std::int8_t* out, hipStream_t st) {
    if (n_ind <= 0) return;
    if (n_snp <= 0) {
        fill_ploidy_kernel<<<static_cast<unsigned>(ceil_div(n_ind, 256)), 256, 0, st>>>(out, n_ind, 1);
        STP_LAUNCH_CHECK();
        return;
    }
    detect_ploidy_kernel<<<static_cast<unsigned>(ceil_div(n_ind, 4)), 256, 0, st>>>(packed, bpr, n_snp,
                                                                                   n_ind, out);
    STP_LAUNCH_CHECK();
}

void launch_fill_ploidy(std::int8_t* out, std::int64_t n, int value, hipStream_t st) {
    if (n <= 0) return;
    fill_ploidy_kernel<<<static_cast<unsigned>(ceil_div(n, 256)), 256, 0, st>>>(out, n, value);
    STP_LAUNCH_CHECK();
}

void launch_decode_counts(const std::uint8_t* packed, std::int64_t bpr, std::int64_t n_snp,
                          const std::uint64_t* pop_offsets, const std::int8_t* ploidy, int n_pop,
                          std::uint16_t* a2, std::uint16_t* n, std::int64_t ld, hipStream_t st) {
    STP_REQUIRE(ld % kSnpPerLane == 0 && ld >= n_snp, "decode: ld must be >= n_snp and a multiple of 16");
    STP_REQUIRE(bpr >= (n_snp + 3) / 4, "decode: bytes_per_record < ceil(n_snp/4)");
    if (n_pop <= 0 || n_snp <= 0) return;
    const std::int64_t tiles = ceil_div(ld, kTileSnps);
    STP_REQUIRE(tiles <= 0x7fffffff, "decode: too many SNP tiles");
    // About four populations per wavefront keeps >= 4 waves per SIMD on wide panels.
    const int waves_y = static_cast<int>(ceil_div(n_pop, 4));
    const unsigned gy = static_cast<unsigned>(std::min<std::int64_t>(ceil_div(waves_y, 4), 65535));
    dim3 grid(static_cast<unsigned>(tiles), gy);
    const bool aligned = (bpr % 4 == 0) && (reinterpret_cast<std::uintptr_t>(packed) % 4 == 0);
    if (aligned)
        decode_counts_kernel<true><<<grid, 256, 0, st>>>(packed, bpr, n_snp, pop_offsets, ploidy, n_pop,
                                                          a2, n, ld);
    else
        decode_counts_kernel<false><<<grid, 256, 0, st>>>(packed, bpr, n_snp, pop_offsets, ploidy, n_pop,
                                                           a2, n, ld);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
