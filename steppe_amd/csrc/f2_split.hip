// steppe_amd/csrc/f2_split.hip — per-block f2 from fixed-point digit slices on BF16 and
// FP16 MFMA.
//
// The reference's default precision is cuBLAS's closed FP64 emulation on integer
// units (src/device/cuda/f2_block_kernel.cu:121-144).  This is an independent
// construction from the published Ozaki-scheme idea (Ozaki, Ogita, Oishi & Rump,
// Numer. Algorithms 59:95-118, 2012; Ootomo, Ozaki & Yokota, IJHPCA 2024): split
// every operand into fixed-point digits whose pairwise products and partial sums
// are exact in the MFMA's FP32 accumulator, and recombine the exact partial sums
// in FP64.  It computes the same estimator as the FP64 path (f2_fp64.hip):
//
//   num_ij = sum_s v_i v_j [(q_i - q_j)^2 - h_i - h_j],  f2 = num / vpair
//
// Operands.  The estimator only sees differences q_i - q_j, so any per-SNP shift of q
// leaves it unchanged; with q' = q - 1/2 (|q'| <= 1/2) and w = -(q'^2 - h)/2
// (|w| <= 1/8):
//   q' is written as six balanced 8-bit digits (|digit| <= 128, exact in BF16),
//       q~ = 2^-7 sum_k d_k 2^-8k,  k = 0..5;
//   w (computed from q~, not q') as four FP16 digits of up to 11 significant bits,
//       w~ = A 2^-14 + B 2^-30 + C 2^-38 + D 2^-46,
//       A = rint(w 2^14) (|A| <= 2048), B = 32 b, C = 4 c (|b|, |c| <= 1024), |D| <= 512,
//   so |w - w~| <= 2^-47, and each digit is an integer in the units of the level it
//   enters.  The device evaluates the estimator for the perturbed frequencies q~
//   (|q~ - q'| <= 2^-48).  Then, with level units U_t = 2^(-14-8t),
//     -num_ij / 2 = sum_s [ q~_i q~_j + w~_i v_j + v_i w~_j ] = sum_t U_t * L_t,
//     L_t = sum_s [ sum_{k+l=t} d_k(i) d_l(j) + W_t(i) v(j) + v(i) W_t(j) ],
//   W_0 = A, W_1 = 0, W_2 = B, W_3 = C, W_4 = D.
// The digit products d_k d_l run on BF16 MFMA and the W_t v products on FP16 MFMA
// (v in {0, 1}); both accumulate into the same FP32 level accumulators.  Levels
// t = 0..4 each have one.  Per SNP, |L_2| grows by at most 3 * 2^15 and the others by
// less, so 128 SNPs keep every L_t below 2^24 and the MFMA sums are exact.
// Level 5 (d_k d_{5-k} for k = 0..5) is folded into L_4 through operands scaled by
// 2^-8 (exact in BF16): those sums are rounded at 2^-47 per MFMA, far below the
// result's precision.  Products with k + l >= 6 are dropped (<= 2^-45 per SNP).
// Every 128 SNPs each level carries its multiple of 2^8 into the level above (exact
// FP32 operations) and L_0 carries multiples of 2^12 into an FP32 high word, so a
// block of any length accumulates in six FP32 words per output and FP64 appears only
// in the per-block epilogue.
//
// MFMA work per (pair, SNP): 29 products (21 BF16 digit-by-digit, 8 FP16
// digit-by-V), plus the pair count V V^T, computed exactly by a separate int8 MFMA
// kernel.  Operand bytes per (population, SNP): 22 (six q digits, four w digits, V).
//
// Kernel shape: persistent workgroups of 8 wavefronts (2 per SIMD) own a 128x128
// output tile of one jackknife block; each wavefront owns 64x32 (two 32x32 MFMA
// tiles, six FP32 accumulators per output).  Each 16-SNP K-step is staged through
// LDS in two phases by direct global->LDS loads: the q digits (two slots, alternating
// between K-steps), then the w digits and V (A, V and, in the second ring layout, B in
// two alternating slots; the rest in one).

#include <utility>

#include "common.hpp"
#include "f2_layout.hpp"

namespace stp {

void launch_zero_u32(unsigned* p, int n, hipStream_t st);

namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

constexpr int kPlanes = kSplitPlanes;  // d0..d5 (0-5), w digit A (6), V (7), w digits B, C, D (8-10)
constexpr int kPlaneA = 6, kVPlane = 7;
constexpr int kChunk = 1024;           // bytes per (32-row group, 16-SNP step, plane)
constexpr int kGroups = 8;             // 32-row groups per workgroup tile: 4 left + 4 right
constexpr int kDPlanes = 6;                                 // q digits d0..d5 per slot
constexpr int kDSlotBytes = kGroups * kDPlanes * kChunk;    // 48 KB
// The w-digit and V planes in two slot kinds, X (double-buffered) and Y (single), for two
// ring layouts: RING 0 holds A, V in X and B, C, D in Y (152 KB of LDS in all, the item
// index broadcast through LDS); RING 1 holds A, V, B in X and C, D in Y (160 KB, all of
// the LDS, so the item index goes through a per-workgroup word in device memory).
template <int RING>
constexpr int x_planes() { return RING ? 3 : 2; }
template <int RING>
constexpr int y_planes() { return RING ? 2 : 3; }
constexpr int kWaves = 8;

// f(std::integral_constant<int, I>) for I = 0 .. N-1, as straight-line code.
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    [&]<int... I>(std::integer_sequence<int, I...>) {
        (f(std::integral_constant<int, I>{}), ...);
    }(std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------------
// prep: frequencies -> 11 digit planes (32x32x16 MFMA fragment order: six BF16 q
// digits, then FP16 A, V, B, C, D) + the int8 V plane (16x16x64 fragment order, for
// the count kernel), in one pass.
//   plane chunk (g32, t16, p) holds rows 32 g32 + lane%32, SNPs 16 t16 + 8 (lane/32) + j
//   at byte offset ((g32 * KS16 + t16) * 11 + p) * 1024 + lane * 16 + 2 j.
//   Plane 5 (d5) is stored scaled by 2^-8 (it only enters level 5).
// A workgroup converts one (32-row group, 64-SNP chunk) unit at a time: it stages the
// unit's frequencies and heterozygosity terms in LDS with coalesced row reads, then
// every thread expands one 8-SNP fragment into its digits.
// ---------------------------------------------------------------------------------
constexpr int kStageLd = 65;  // LDS row stride (doubles) of the staged unit

// Staging reads of the two frequency sources.  Each thread stages 8 elements of the
// 32 x 64 unit: fetch() issues the global loads into registers (so a workgroup can
// fetch its next unit while it converts the current one), put() turns them into
// (q - 1/2, h) in sq/sh, with h = -1 marking an invalid entry.
// h = q(1-q)/max(n-1, 1) as in the reference's het_correction
// (src/core/internal/f2_estimator.hpp:22-29).
struct CountsSrc {
    const std::uint16_t* a2;
    const std::uint16_t* n;
    std::int64_t ld;
    struct Raw {
        unsigned a[8], c[8];
        bool ok[8];
    };
    // Lanes run along the SNP axis: each wavefront reads 128 contiguous bytes per row.
    // The loads are unconditional (entries outside the unit re-read a valid one and are
    // masked), so the compiler knows how many are in flight and waits for them alone,
    // not for the previous unit's plane stores as well.
    __device__ Raw fetch(int r0, int P, std::int64_t snp0, int live, int tid) const {
        Raw raw;
        const int kk = tid & 63;
        const bool col_ok = kk < live;
        const std::int64_t col = snp0 + (col_ok ? kk : 0);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = (tid >> 6) + 4 * i;
            const bool ok = col_ok && r0 + row < P;
            const std::int64_t o = static_cast<std::int64_t>(ok ? r0 + row : r0) * ld + col;
            raw.a[i] = a2[o];
            raw.c[i] = n[o];
            raw.ok[i] = ok;
        }
        return raw;
    }
    __device__ void put(const Raw& raw, int tid, double* sq, double* sh) const {
        const int kk = tid & 63;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int row = (tid >> 6) + 4 * i;
            double qp = 0.0, h = -1.0;
            if (raw.ok[i] && raw.c[i]) {
                const double nd = static_cast<double>(raw.c[i]);
                const double q = static_cast<double>(raw.a[i]) / (2.0 * nd);
                const double nm1 = nd - 1.0;
                h = q * (1.0 - q) / (nm1 > 1.0 ? nm1 : 1.0);
                qp = q - 0.5;
            }
            sq[row * kStageLd + kk] = qp;
            sh[row * kStageLd + kk] = h;
        }
    }
};

struct QvnSrc {
    const double* q;
    const double* v;
    const double* n;
    struct Raw {
        double q[8], v[8], n[8];
        bool ok[8];
    };
    // Q/V/N are population-fastest (i + P*s): lanes run along the population axis.
    // Unconditional loads, as in CountsSrc::fetch.
    __device__ Raw fetch(int r0, int P, std::int64_t snp0, int live, int tid) const {
        Raw raw;
        const int row = tid & 31;
        const bool row_ok = r0 + row < P;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kk = (tid >> 5) + 8 * i;
            const bool ok = row_ok && kk < live;
            const std::int64_t o = (row_ok ? r0 + row : r0) + static_cast<std::int64_t>(P) * (snp0 + (ok ? kk : 0));
            raw.v[i] = v[o];
            raw.q[i] = q[o];
            raw.n[i] = n[o];
            raw.ok[i] = ok;
        }
        return raw;
    }
    __device__ void put(const Raw& raw, int tid, double* sq, double* sh) const {
        const int row = tid & 31;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kk = (tid >> 5) + 8 * i;
            double qp = 0.0, h = -1.0;
            if (raw.ok[i] && raw.v[i] != 0.0) {
                const double qq = raw.q[i], nm1 = raw.n[i] - 1.0;
                h = qq * (1.0 - qq) / (nm1 > 1.0 ? nm1 : 1.0);
                qp = qq - 0.5;
            }
            sq[row * kStageLd + kk] = qp;
            sh[row * kStageLd + kk] = h;
        }
    }
};

// Rows, SNP range and block of a prep unit (32-row group, 64-SNP chunk).
struct PrepUnit {
    int g, live;
    std::int64_t c, snp0;
};

// The block-table arrays arrive as separate read-only, non-aliased kernel arguments, so
// that these uniform lookups compile to scalar loads: a vector load here would have to
// wait (in order) for the previous unit's plane stores before its value could be used.
__device__ inline PrepUnit prep_unit(std::int64_t u, std::int64_t n_chunks, const std::int64_t* __restrict__ s0,
                                     const std::int32_t* __restrict__ len, const std::int64_t* __restrict__ k0,
                                     const std::int32_t* __restrict__ chunk_block) {
    PrepUnit pu;
    pu.g = static_cast<int>(u / n_chunks);
    pu.c = u - static_cast<std::int64_t>(pu.g) * n_chunks;
    const int b = chunk_block[pu.c];
    const std::int64_t local0 = 64 * pu.c - k0[b];
    pu.snp0 = s0[b] + local0;
    pu.live = static_cast<int>(max<std::int64_t>(0, min<std::int64_t>(64, len[b] - local0)));
    return pu;
}

// Workgroup barrier for the LDS stage only: waits for this wavefront's LDS operations,
// not for its outstanding global stores (__syncthreads would also wait for those).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ inline unsigned bf16_bits(double integer_digit) {
    // Integer digits |d| <= 2^14 and their 2^-8 multiples have <= 8 significant
    // bits (a power of two times an integer below 2^8 in magnitude, or a power
    // of two): the float and bf16 conversions are exact.
    return __float_as_uint(static_cast<float>(integer_digit)) >> 16;
}

__device__ inline unsigned f16_bits(double digit) {
    // The w digits are integers of at most 11 significant bits times a power of two
    // (|A| <= 2048, 32 b, 4 c with |b|, |c| <= 1024, |D| <= 512) and V is 0 or 1: the
    // float and FP16 conversions are exact.
    return __builtin_bit_cast(unsigned short, static_cast<_Float16>(static_cast<float>(digit)));
}

template <class Src>
__global__ __launch_bounds__(256) void prep_split_kernel(Src src, int P, int G32, int G16,
                                                         const std::int64_t* __restrict__ bt_s0,
                                                         const std::int32_t* __restrict__ bt_len,
                                                         const std::int64_t* __restrict__ bt_k0,
                                                         const std::int32_t* __restrict__ chunk_block,
                                                         std::int64_t n_chunks, std::int64_t KS16,
                                                         std::uint8_t* __restrict__ planes,
                                                         i4* __restrict__ pV8) {
    __shared__ double sq[32 * kStageLd];
    __shared__ double sh[32 * kStageLd];
    const int tid = threadIdx.x;
    const int t4 = tid >> 6, lane = tid & 63;
    const std::int64_t n_units = static_cast<std::int64_t>(G32) * n_chunks;
    // Units run chunk-fastest so consecutive workgroups read adjacent row segments and
    // write adjacent plane chunks.  The next unit's inputs are fetched into registers
    // before the current unit is converted, so their latency overlaps the conversion,
    // and the stage is refilled only after the current unit's stores are issued: each
    // refill then waits for its own loads, never for the plane stores issued after them.
    // The first unit's stage is filled before the loop, and the last iteration re-reads
    // a valid unit, so that every path through the loop issues the same operations.
    std::int64_t u = blockIdx.x;
    if (u >= n_units) return;
    PrepUnit cur = prep_unit(u, n_chunks, bt_s0, bt_len, bt_k0, chunk_block);
    src.put(src.fetch(32 * cur.g, P, cur.snp0, cur.live, tid), tid, sq, sh);
    lds_barrier();
    for (;;) {
        const bool more = u + gridDim.x < n_units;
        const PrepUnit nxt = prep_unit(more ? u + gridDim.x : u, n_chunks, bt_s0, bt_len, bt_k0, chunk_block);
        const typename Src::Raw raw = src.fetch(32 * nxt.g, P, nxt.snp0, nxt.live, tid);
        const int g = cur.g;
        const std::int64_t c = cur.c;

        // BF16 digit planes: thread -> (16-SNP step t4 of the chunk, fragment lane).
        const int r = lane & 31;
        unsigned w[kPlanes][4];
#pragma unroll
        for (int p = 0; p < kPlanes; ++p)
#pragma unroll
            for (int hh = 0; hh < 4; ++hh) w[p][hh] = 0u;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int kk = 16 * t4 + 8 * (lane >> 5) + j;
            const double h = sh[r * kStageLd + kk];
            if (h >= 0.0) {
                double x = sq[r * kStageLd + kk] * 128.0;
                double d[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) {
                    d[k] = rint(x);
                    x = (x - d[k]) * 256.0;
                }
                const double qt =
                    (d[0] + (d[1] + (d[2] + (d[3] + (d[4] + d[5] / 256.0) / 256.0) / 256.0) / 256.0) / 256.0) / 128.0;
                // w in units of 2^-14, then the remainders in units of 2^-30 / 32, 2^-38 / 4
                // and 2^-46 (every step is exact in FP64).
                const double y = -0.5 * (qt * qt - h) * 16384.0;
                const double wA = rint(y);
                const double r1 = (y - wA) * 2048.0;
                const double wb = rint(r1);
                const double r2 = (r1 - wb) * 2048.0;
                const double wc = rint(r2);
                const double wD = rint((r2 - wc) * 1024.0);
                const int sft = 16 * (j & 1);
#pragma unroll
                for (int k = 0; k < 5; ++k) w[k][j >> 1] |= bf16_bits(d[k]) << sft;
                w[5][j >> 1] |= bf16_bits(d[5] / 256.0) << sft;
                w[kPlaneA][j >> 1] |= f16_bits(wA) << sft;
                w[kVPlane][j >> 1] |= f16_bits(1.0) << sft;
                w[8][j >> 1] |= f16_bits(32.0 * wb) << sft;
                w[9][j >> 1] |= f16_bits(4.0 * wc) << sft;
                w[10][j >> 1] |= f16_bits(wD) << sft;
            }
        }
        // int8 V plane for the pair-count kernel (16x16x64 layout): the unit's two
        // 16-row groups, one 16-byte fragment per thread of the first two wavefronts.
        // Stored before the digit planes, so that the next unit's loads are followed by
        // exactly the 11 plane stores of every wavefront.
        if (tid < 128) {
            const int g16 = 2 * g + (tid >> 6);
            if (g16 < G16) {
                const int rr = 16 * (tid >> 6) + (lane & 15);
                const int kb = 16 * (lane >> 4);
                unsigned v4[4] = {0, 0, 0, 0};
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (sh[rr * kStageLd + kb + j] >= 0.0) v4[j >> 2] |= 1u << (8 * (j & 3));
                i4 v;
                v[0] = static_cast<int>(v4[0]);
                v[1] = static_cast<int>(v4[1]);
                v[2] = static_cast<int>(v4[2]);
                v[3] = static_cast<int>(v4[3]);
                pV8[(static_cast<std::int64_t>(g16) * (KS16 / 4) + c) * 64 + lane] = v;
            }
        }
        const std::int64_t t16 = 4 * c + t4;
        std::uint8_t* base = planes + ((static_cast<std::int64_t>(g) * KS16 + t16) * kPlanes) * kChunk + lane * 16;
        // Streaming stores: the planes are far larger than the caches and are read back
        // by a later kernel in another order.
        typedef unsigned u4v __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int p = 0; p < kPlanes; ++p)
            __builtin_nontemporal_store(u4v{w[p][0], w[p][1], w[p][2], w[p][3]}, reinterpret_cast<u4v*>(base + p * kChunk));
        if (!more) break;
        lds_barrier();  // every reader of this unit's stage is done
        src.put(raw, tid, sq, sh);
        lds_barrier();
        cur = nxt;
        u += gridDim.x;
    }
}

// ---------------------------------------------------------------------------------
// pair counts: vpair_ij = sum_s v_i v_j by int8 MFMA (exact int32), written for both
// (i,j) and (j,i).  Persistent 4-wavefront workgroups over (block, 128x128 tile pair)
// items; wavefront w owns the 64x64 quadrant (rows 64 (w & 1), columns 64 (w >> 1)),
// so every 16-row fragment a wavefront loads is loaded by one other wavefront of the
// workgroup at the same time (the second load hits the CU's L1).  Quadrants below the
// diagonal of a diagonal tile pair are skipped.
// ---------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void count_kernel(const i4* __restrict__ pV8, BlockTable bt,
                                                    const F2Item* __restrict__ items, ItemQueues queues,
                                                    unsigned* __restrict__ counters, int P, int G16,
                                                    std::int64_t KS64, double* __restrict__ vpair) {
    __shared__ int s_item;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int rh = wave & 1, ch = wave >> 1;
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;
    const int home = static_cast<int>(blockIdx.x % kXcds);
    int xq = 0;
    for (;;) {
        if (threadIdx.x == 0) s_item = claim_item(counters, queues, home, xq);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it < 0) break;
        const F2Item wi = items[it];
        const int b = __builtin_amdgcn_readfirstlane(wi.b);
        const int ti = __builtin_amdgcn_readfirstlane(static_cast<int>(wi.ti));
        const int tj = __builtin_amdgcn_readfirstlane(static_cast<int>(wi.tj));
        if (ti == tj && rh > ch) continue;
        const int gI0 = 8 * ti + 4 * rh;
        const int gJ0 = 8 * tj + 4 * ch;
        if (gI0 >= G16 || gJ0 >= G16) continue;
        const std::int64_t u0 = bt.k0[b] / 64;
        const int nu = bt.kpad[b] / 64;
        bool rowI[4], rowJ[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) { rowI[x] = gI0 + x < G16; rowJ[x] = gJ0 + x < G16; }
        i4 cnt[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) cnt[x][y] = i4{0, 0, 0, 0};
        for (int u = 0; u < nu; ++u) {
            i4 cI[4], cJ[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                cI[x] = rowI[x] ? pV8[((gI0 + x) * KS64 + u0 + u) * 64 + lane] : i4{0, 0, 0, 0};
                cJ[x] = rowJ[x] ? pV8[((gJ0 + x) * KS64 + u0 + u) * 64 + lane] : i4{0, 0, 0, 0};
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y)
                    if (rowI[x] && rowJ[y] && gI0 + x <= gJ0 + y)
                        cnt[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cI[x], cJ[y], cnt[x][y], 0, 0, 0);
        }
        double* vpb = vpair + slab * b;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int gI = gI0 + x, gJ = gJ0 + y;
                if (!(rowI[x] && rowJ[y] && gI <= gJ)) continue;
                const int col = 16 * gJ + (lane & 15);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = 16 * gI + 4 * (lane >> 4) + e;
                    if (row >= P || col >= P || (gI == gJ && row > col)) continue;
                    const double vp = static_cast<double>(cnt[x][y][e]);
                    vpb[row + static_cast<std::int64_t>(P) * col] = vp;
                    vpb[col + static_cast<std::int64_t>(P) * row] = vp;
                }
            }
    }
}

// ---------------------------------------------------------------------------------
// The split contraction.
// ---------------------------------------------------------------------------------
__device__ inline bf8 lds_frag(const std::uint8_t* slot, int group, int plane, int nplanes, int lane) {
    return *reinterpret_cast<const bf8*>(slot + (group * nplanes + plane) * kChunk + lane * 16);
}

// Multiply 8 BF16 values by 2^-8 (exact: the values are small integers or zero).
__device__ inline bf8 scale_2m8(bf8 a) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 x = __builtin_bit_cast(u4, a);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float lo = __uint_as_float(x[i] << 16) * 0.00390625f;
        const float hi = __uint_as_float(x[i] & 0xFFFF0000u) * 0.00390625f;
        x[i] = (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xFFFF0000u);
    }
    return __builtin_bit_cast(bf8, x);
}

// s_waitcnt on the vector-memory count only (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14; expcnt 6:4 and lgkmcnt 11:8 left at their maxima).  As a builtin, unlike
// inline assembly, the compiler's wait-count tracking takes it into account.
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt is six bits");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Slot kinds of a K-step: 0 = q digits d0..d5 (planes 0-5), 1 = X (A, V and, in RING 1,
// B: planes 6-7 or 6-8), 2 = Y (the rest of B, C, D: planes 8-10 or 9-10).
template <int kind, int RING>
constexpr int slot_planes() {
    return kind == 0 ? kDPlanes : (kind == 1 ? x_planes<RING>() : y_planes<RING>());
}
template <int kind, int RING>
constexpr int slot_first_plane() {
    return kind == 0 ? 0 : (kind == 1 ? kPlaneA : kPlaneA + x_planes<RING>());
}

// Issue this wavefront's share of one slot's 1 KB chunks (direct global->LDS).
// Every wavefront issues exactly (8 groups * planes) / 8 = planes loads per slot, and
// chunk ci of a slot is always loaded by wavefront ci % 8; the compile-time count
// lets the compiler's wait-count tracking see how many loads follow each slot's, so
// that it adds no wait of its own beyond the explicit ones.
template <int kind, int RING>
__device__ inline void stage_slot(const std::uint8_t* __restrict__ planes, std::uint8_t* slot, std::int64_t t16,
                                  std::int64_t KS16, int gI0, int gJ0, int G32, int wave, int lane) {
    constexpr int np = slot_planes<kind, RING>();
#pragma unroll
    for (int r = 0; r < np; ++r) {
        const int ci = wave + kWaves * r;
        const int lg = ci / np, pp = ci - lg * np;
        int gg = lg < 4 ? gI0 + lg : gJ0 + lg - 4;
        if (gg >= G32) gg = 0;  // rows beyond P: any valid address; never read back
        const std::uint8_t* src =
            planes + ((static_cast<std::int64_t>(gg) * KS16 + t16) * kPlanes + slot_first_plane<kind, RING>() + pp) * kChunk +
            lane * 16;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src),
                                         (__attribute__((address_space(3))) void*)(slot + ci * kChunk), 16, 0, 0);
    }
}

__device__ inline h8 lds_frag_h(const std::uint8_t* slot, int group, int plane, int nplanes, int lane) {
    return *reinterpret_cast<const h8*>(slot + (group * nplanes + plane) * kChunk + lane * 16);
}

template <int RING>
__global__ __launch_bounds__(512) void f2_split_kernel(const std::uint8_t* __restrict__ planes, BlockTable bt,
                                                       const F2Item* __restrict__ items, ItemQueues queues,
                                                       unsigned* __restrict__ counters, int P, int G32,
                                                       std::int64_t KS16, const double* __restrict__ vpair,
                                                       double* __restrict__ f2, FixupList fl,
                                                       int* __restrict__ item_slot) {
    constexpr int kXPlanes = x_planes<RING>(), kYPlanes = y_planes<RING>();
    constexpr int kXSlotBytes = kGroups * kXPlanes * kChunk, kYSlotBytes = kGroups * kYPlanes * kChunk;
    // One LDS object per ring slot: the compiler then knows that reads of one slot do
    // not depend on the direct loads still landing in the others.  The q-digit slots
    // and the A + V slots alternate between K-steps; B..D have one slot.
    __shared__ __attribute__((aligned(16))) std::uint8_t lds_d0[kDSlotBytes];
    __shared__ __attribute__((aligned(16))) std::uint8_t lds_d1[kDSlotBytes];
    __shared__ __attribute__((aligned(16))) std::uint8_t lds_x0[kXSlotBytes];
    __shared__ __attribute__((aligned(16))) std::uint8_t lds_x1[kXSlotBytes];
    __shared__ __attribute__((aligned(16))) std::uint8_t lds_y[kYSlotBytes];
    const int lane = threadIdx.x & 63;
    // Wavefront-uniform values are moved to scalar registers so that every branch
    // on them is uniform and the MFMA accumulators never sit in divergent flow.
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int wa = wave & 1;   // left sub-tile: groups 2wa, 2wa+1 of the left tile
    const int wc = wave >> 1;  // right sub-tile: group wc of the right tile
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;
    const std::int64_t t_last = KS16 - 1;

    const int home = static_cast<int>(blockIdx.x % kXcds);
    int xq = 0;
    for (;;) {
        int it;
        if constexpr (RING == 0) {
            __shared__ int s_item;
            if (threadIdx.x == 0) s_item = claim_item(counters, queues, home, xq);
            __syncthreads();
            it = __builtin_amdgcn_readfirstlane(s_item);
            __syncthreads();
        } else {
            if (threadIdx.x == 0)
                __hip_atomic_store(&item_slot[blockIdx.x], claim_item(counters, queues, home, xq), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __syncthreads();
            it = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&item_slot[blockIdx.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            __syncthreads();
        }
        if (it < 0) break;
        const F2Item wi = items[it];
        const int b = __builtin_amdgcn_readfirstlane(wi.b);
        const int gI0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(wi.ti));
        const int gJ0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(wi.tj));
        const std::int64_t t0 = bt.k0[b] / 16;
        const int nsteps = __builtin_amdgcn_readfirstlane(bt.kpad[b] / 16);
        // K-step index of step s, clamped: look-ahead loads past the block end (and the
        // buffer end) re-read a valid K-step and are never read back.
        auto kstep = [&](int s) { return t0 + s < t_last ? t0 + s : t_last; };

        bool act[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int li = 2 * wa + x;
            act[x] = (gI0 + li < G32) && (gJ0 + wc < G32) && (gI0 < gJ0 || li <= wc);
        }

        f16v L[2][5], H[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
#pragma unroll
            for (int t = 0; t < 5; ++t) L[x][t] = f16v{};
            H[x] = f16v{};
        }

        // Load schedule (per wavefront, loads per slot D 6, X 2 or 3, Y 3 or 2; all loads
        // stage unconditionally, so every path issues the same counts and the explicit
        // waits below are exact):
        //   prologue:       D(0) -> d0, X(0) -> x0, D(1) -> d1
        //   step s, phase d: wait until <= 6 + |X| loads are outstanding (D(s) done: only
        //                    binding at s = 0), barrier (Y free), issue Y(s); q-digit
        //                    products from D(s)
        //   phase w:         wait for all (Y(s), and X(s), D(s+1) issued before it),
        //                    barrier (D(s) and X(s-1) free), issue D(s+2) into D(s)'s
        //                    slot and X(s+1) into X(s-1)'s; w-digit products from X(s), Y(s)
        stage_slot<0, RING>(planes, lds_d0, kstep(0), KS16, gI0, gJ0, G32, wave, lane);
        stage_slot<1, RING>(planes, lds_x0, kstep(0), KS16, gI0, gJ0, G32, wave, lane);
        stage_slot<0, RING>(planes, lds_d1, kstep(1), KS16, gI0, gJ0, G32, wave, lane);

        auto step_body = [&]<int par>(int step) {
            std::uint8_t* const slot_d = par ? lds_d1 : lds_d0;
            std::uint8_t* const slot_x = par ? lds_x1 : lds_x0;
            std::uint8_t* const slot_xn = par ? lds_x0 : lds_x1;
            // ---- q digits (BF16)
            wait_vmcnt<kDPlanes + kXPlanes>();
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            stage_slot<2, RING>(planes, lds_y, kstep(step), KS16, gI0, gJ0, G32, wave, lane);
            {
                bf8 dj[6];
#pragma unroll
                for (int k = 0; k < 6; ++k) dj[k] = lds_frag(slot_d, 4 + wc, k, kDPlanes, lane);
                static_for<6>([&](auto kc) {
                    constexpr int k = decltype(kc)::value;
                    // Level 5 into L4: d5 is stored pre-scaled; for k = 1..4 the right
                    // operand d_(5-k) is scaled here, once for both left sub-tiles (the
                    // products are exact either way).
                    const bf8 d5r = (k == 0 || k == 5) ? dj[5 - k] : scale_2m8(dj[5 - k]);
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        if (!act[x]) continue;
                        const bf8 di = lds_frag(slot_d, 2 * wa + x, k, kDPlanes, lane);
#pragma unroll
                        for (int t = k; t < 5; ++t)
                            L[x][t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(di, dj[t - k], L[x][t], 0, 0, 0);
                        L[x][4] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(di, d5r, L[x][4], 0, 0, 0);
                    }
                });
            }
            // ---- w digits x V (FP16): A into L0, B, C, D into L2, L3, L4
            wait_vmcnt<0>();
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            stage_slot<0, RING>(planes, slot_d, kstep(step + 2), KS16, gI0, gJ0, G32, wave, lane);
            stage_slot<1, RING>(planes, slot_xn, kstep(step + 1), KS16, gI0, gJ0, G32, wave, lane);
            {
                const h8 aj = lds_frag_h(slot_x, 4 + wc, 0, kXPlanes, lane);
                const h8 vj = lds_frag_h(slot_x, 4 + wc, 1, kXPlanes, lane);
                // w digits B, C, D: plane m of the three in X (after A, V) or Y
                auto wfrag = [&](int group, int m) {
                    return m < kXPlanes - 2 ? lds_frag_h(slot_x, group, 2 + m, kXPlanes, lane)
                                            : lds_frag_h(lds_y, group, m - (kXPlanes - 2), kYPlanes, lane);
                };
                h8 wj[3];
#pragma unroll
                for (int m = 0; m < 3; ++m) wj[m] = wfrag(4 + wc, m);
#pragma unroll
                for (int x = 0; x < 2; ++x) {
                    if (!act[x]) continue;
                    const h8 ai = lds_frag_h(slot_x, 2 * wa + x, 0, kXPlanes, lane);
                    const h8 vi = lds_frag_h(slot_x, 2 * wa + x, 1, kXPlanes, lane);
                    L[x][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ai, vj, L[x][0], 0, 0, 0);
                    L[x][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(vi, aj, L[x][0], 0, 0, 0);
#pragma unroll
                    for (int m = 0; m < 3; ++m) {
                        const h8 wi = wfrag(2 * wa + x, m);
                        L[x][2 + m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(wi, vj, L[x][2 + m], 0, 0, 0);
                        L[x][2 + m] = __builtin_amdgcn_mfma_f32_32x32x16_f16(vi, wj[m], L[x][2 + m], 0, 0, 0);
                    }
                }
                // Every 8 K-steps (128 SNPs) and at the block end: exact carries.
                if ((step & 7) == 7 || step + 1 == nsteps) {
#pragma unroll
                    for (int x = 0; x < 2; ++x) {
                        if (!act[x]) continue;
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
#pragma unroll
                            for (int t = 4; t >= 1; --t) {
                                const float cr = rintf(L[x][t][e] * (1.0f / 256.0f));
                                L[x][t][e] = fmaf(cr, -256.0f, L[x][t][e]);
                                L[x][t - 1][e] += cr;
                            }
                            const float c0 = rintf(L[x][0][e] * (1.0f / 4096.0f)) * 4096.0f;
                            L[x][0][e] -= c0;
                            H[x][e] += c0;
                        }
                    }
                }
            }
        };
        // Steps alternate between the two q-digit (and A + V) slots; unrolled by two so
        // that each slot is a distinct LDS object in every access, with an odd last step
        // after the loop (so that both bodies always run inside it).
        int step = 0;
        for (; step + 1 < nsteps; step += 2) {
            step_body.template operator()<0>(step);
            step_body.template operator()<1>(step + 1);
        }
        if (step < nsteps) step_body.template operator()<0>(step);

        // Epilogue: num = -2 sum_t U_t L_t in FP64; f2 = num / vpair.
        double* f2b = f2 + slab * b;
        const double* vpb = vpair + slab * b;
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            if (!act[x]) continue;
            const int gI = gI0 + 2 * wa + x, gJ = gJ0 + wc;
            const int col = 32 * gJ + (lane & 31);
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int row = 32 * gI + 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3);
                const bool ok = row < P && col < P && !(gI == gJ && row > col);
                double val = 0.0, vp = 0.0;
                if (ok) {
                    const double s = (static_cast<double>(H[x][e]) + static_cast<double>(L[x][0][e])) * 0x1p-14 +
                                     static_cast<double>(L[x][1][e]) * 0x1p-22 +
                                     static_cast<double>(L[x][2][e]) * 0x1p-30 +
                                     static_cast<double>(L[x][3][e]) * 0x1p-38 +
                                     static_cast<double>(L[x][4][e]) * 0x1p-46;
                    const double num = -2.0 * s;
                    const std::int64_t ij = row + static_cast<std::int64_t>(P) * col;
                    vp = vpb[ij];
                    val = vp > 0.0 ? num / vp : 0.0;
                    f2b[ij] = val;
                    f2b[col + static_cast<std::int64_t>(P) * row] = val;
                }
                fixup_append(fl, ok && vp > 0.0 && fabs(val) < fl.tau, b, row, col);
            }
        }
    }
    // The last item's look-ahead loads may still be landing in LDS: drain them before
    // the workgroup's LDS can be handed to another workgroup.
    wait_vmcnt<0>();
}

}  // namespace

// Persistent-style grid for the prep: enough workgroups to fill every CU several
// times over, each looping over (32-row group, chunk) units.
static unsigned prep_grid(std::int64_t n_units) {
    return static_cast<unsigned>(std::max<std::int64_t>(1, std::min<std::int64_t>(n_units, 8192)));
}

void launch_prep_split_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, int G32,
                              int G16, BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks,
                              std::int64_t KS16, std::uint8_t* planes, void* pV8, hipStream_t st) {
    if (n_chunks <= 0 || G32 <= 0) return;
    prep_split_kernel<CountsSrc><<<prep_grid(G32 * n_chunks), 256, 0, st>>>(
        CountsSrc{a2, n, ld}, P, G32, G16, bt.s0, bt.len, bt.k0, chunk_block, n_chunks, KS16, planes,
        static_cast<i4*>(pV8));
    STP_LAUNCH_CHECK();
}

void launch_prep_split_qvn(const double* q, const double* v, const double* n, int P, int G32, int G16,
                           BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks, std::int64_t KS16,
                           std::uint8_t* planes, void* pV8, hipStream_t st) {
    if (n_chunks <= 0 || G32 <= 0) return;
    prep_split_kernel<QvnSrc><<<prep_grid(G32 * n_chunks), 256, 0, st>>>(
        QvnSrc{q, v, n}, P, G32, G16, bt.s0, bt.len, bt.k0, chunk_block, n_chunks, KS16, planes,
        static_cast<i4*>(pV8));
    STP_LAUNCH_CHECK();
}

void launch_count(const void* pV8, BlockTable bt, const F2Item* items, const ItemQueues& queues,
                  unsigned* counters, int P, int G16, std::int64_t KS64, double* vpair, int n_workgroups,
                  hipStream_t st) {
    if (queues.off[kXcds] <= 0) return;
    launch_zero_u32(counters, kXcds, st);
    count_kernel<<<static_cast<unsigned>(n_workgroups), 256, 0, st>>>(static_cast<const i4*>(pV8), bt, items,
                                                                      queues, counters, P, G16, KS64, vpair);
    STP_LAUNCH_CHECK();
}

void launch_f2_split(const std::uint8_t* planes, BlockTable bt, const F2Item* items, const ItemQueues& queues,
                     unsigned* counters, int P, int G32, std::int64_t KS16, const double* vpair, double* f2,
                     const FixupList& fl, int n_workgroups, int ring, int* item_slot, hipStream_t st) {
    if (queues.off[kXcds] <= 0) return;
    launch_zero_u32(counters, kXcds, st);
    if (ring == 1)
        f2_split_kernel<1><<<static_cast<unsigned>(n_workgroups), 512, 0, st>>>(planes, bt, items, queues, counters, P,
                                                                               G32, KS16, vpair, f2, fl, item_slot);
    else
        f2_split_kernel<0><<<static_cast<unsigned>(n_workgroups), 512, 0, st>>>(planes, bt, items, queues, counters, P,
                                                                               G32, KS16, vpair, f2, fl, item_slot);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
