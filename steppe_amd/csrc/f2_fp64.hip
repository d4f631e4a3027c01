// steppe_amd/csrc/f2_fp64.hip — the plain FP64 per-block f2 contraction.
//
// Replaces the reference's FP64 feeder + size-bucketed gather + three strided-
// batched cuBLAS GEMMs + assemble kernel (src/device/cuda/f2_block_kernel.cu:37-61,
// src/device/cuda/f2_batched_kernel.cu:31-187, src/device/cuda/cuda_backend_f2_blocks.cu:241-346)
// with one prep kernel and one persistent MFMA kernel, computing for every block b
// and pair i <= j exactly the reference estimator (src/core/internal/f2_estimator.hpp:22-51):
//
//   num_ij = sum_s v_i v_j [ (q_i - q_j)^2 - h_i - h_j ],   h = q(1-q)/max(n-1, 1)
//   vpair_ij = sum_s v_i v_j,   f2_ij = num_ij / vpair_ij  (0 when vpair_ij == 0)
//
// Algebra used on the device.  With a per-SNP centre c_s (shift-invariance of
// (q_i - q_j)^2), q' = v (q - c), A = v (q'^2 - h):
//
//   num_ij = sum_s [ A_i v_j + v_i A_j - 2 q'_i q'_j ]
//
// i.e. three rank-K products accumulated into ONE FP64 accumulator (left operand
// rows [A, V, -2q'], right rows [V, A, q']), so the cancellation between the
// squared and cross terms happens inside the FP64 sum of centred values instead of
// between three separately rounded GEMM outputs.  The pair count is an exact int8
// MFMA product V V^T (int32 accumulation).

#include "common.hpp"
#include "f2_layout.hpp"

namespace stp {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

constexpr double kCentreScale = 256.0;  // centres are multiples of 2^-8

// ---------------------------------------------------------------------------
// prep: per-(population, SNP) frequencies -> FP64 operand planes A, V, q' and the
// int8 V plane.  One workgroup per 64 padded columns (one K-chunk of one block).
// The source is either the decode's integer counts or host-supplied FP64 Q/V/N.
// ---------------------------------------------------------------------------
struct CountsSrc {
    const std::uint16_t* a2;
    const std::uint16_t* n;
    std::int64_t ld;
    __device__ bool get(int r, std::int64_t s, double& q, double& nd) const {
        const std::int64_t o = static_cast<std::int64_t>(r) * ld + s;
        const unsigned c = n[o];
        if (!c) return false;
        nd = static_cast<double>(c);
        q = static_cast<double>(a2[o]) / (2.0 * nd);  // == ac / n, correctly rounded
        return true;
    }
};

struct QvnSrc {
    const double* q;
    const double* v;
    const double* n;
    int P;
    __device__ bool get(int r, std::int64_t s, double& qq, double& nd) const {
        const std::int64_t o = r + static_cast<std::int64_t>(P) * s;
        if (v[o] == 0.0) return false;
        qq = q[o];
        nd = n[o];
        return true;
    }
};

template <class Src>
__global__ __launch_bounds__(256) void prep_fp64_kernel(
    Src src, int P, int G, BlockTable bt, const std::int32_t* __restrict__ chunk_block, std::int64_t KS,
    double* __restrict__ pA, double* __restrict__ pV, double* __restrict__ pQ,
    i4* __restrict__ pV8) {
    __shared__ double red_sum[4][64];
    __shared__ int red_cnt[4][64];
    __shared__ double centre[64];

    const int tid = threadIdx.x;
    const std::int64_t c = blockIdx.x;             // 64-column chunk
    const int b = chunk_block[c];
    const std::int64_t kg0 = 64 * c;               // first padded column of the chunk
    const std::int64_t local0 = kg0 - bt.k0[b];    // offset inside the block
    const std::int64_t snp0 = bt.s0[b] + local0;   // SNP of column kk is snp0 + kk
    const int live = static_cast<int>(max<std::int64_t>(0, min<std::int64_t>(64, bt.len[b] - local0)));

    // Per-SNP centre: mean frequency over the populations observed at the SNP,
    // rounded to a multiple of 2^-8 (any centre is exact algebra; this one keeps
    // |q'| small and is deterministic).
    {
        const int kk = tid & 63, pg = tid >> 6;
        double s = 0.0;
        int cnt = 0;
        if (kk < live) {
            for (int p = pg; p < P; p += 4) {
                double q, nd;
                if (src.get(p, snp0 + kk, q, nd)) { s += q; ++cnt; }
            }
        }
        red_sum[pg][kk] = s;
        red_cnt[pg][kk] = cnt;
        __syncthreads();
        if (tid < 64) {
            const double ts = red_sum[0][tid] + red_sum[1][tid] + red_sum[2][tid] + red_sum[3][tid];
            const int tc = red_cnt[0][tid] + red_cnt[1][tid] + red_cnt[2][tid] + red_cnt[3][tid];
            centre[tid] = tc ? rint(ts / tc * kCentreScale) / kCentreScale : 0.0;
        }
        __syncthreads();
    }

    const std::int64_t t0 = kg0 / 4;  // first 4-column K-step of the chunk
    const std::int64_t u = c;         // 64-column step index of the int8 plane
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + 256 * i;
            const int t4 = e >> 6, ln = e & 63;
            const int r = kRowGroup * g + (ln & 15);
            const int kk = 4 * t4 + (ln >> 4);
            double A = 0.0, V = 0.0, Qc = 0.0;
            double q, nd;
            if (r < P && kk < live && src.get(r, snp0 + kk, q, nd)) {
                const double nm1 = nd - 1.0;
                const double h = q * (1.0 - q) / (nm1 > 1.0 ? nm1 : 1.0);
                Qc = q - centre[kk];
                A = Qc * Qc - h;
                V = 1.0;
            }
            const std::int64_t o = (static_cast<std::int64_t>(g) * KS + t0 + t4) * 64 + ln;
            pA[o] = A;
            pV[o] = V;
            pQ[o] = Qc;
        }
        if (tid < 64) {
            const int r = kRowGroup * g + (tid & 15);
            const int kb = 16 * (tid >> 4);
            unsigned w[4] = {0, 0, 0, 0};
            if (r < P) {
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    const int kk = kb + j;
                    double q, nd;
                    if (kk < live && src.get(r, snp0 + kk, q, nd)) w[j >> 2] |= 1u << (8 * (j & 3));
                }
            }
            i4 v;
            v[0] = static_cast<int>(w[0]);
            v[1] = static_cast<int>(w[1]);
            v[2] = static_cast<int>(w[2]);
            v[3] = static_cast<int>(w[3]);
            pV8[(static_cast<std::int64_t>(g) * (KS / 16) + u) * 64 + tid] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// contraction: persistent wavefronts pull (block, 64x64 tile pair) items from a
// device counter (items are ordered block-major, largest blocks first) and
// accumulate the 4x4 grid of 16x16 sub-tiles over the block's K range.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void f2_fp64_kernel(
    const double* __restrict__ pA, const double* __restrict__ pV, const double* __restrict__ pQ,
    const i4* __restrict__ pV8, BlockTable bt, const F2Item* __restrict__ items, ItemQueues queues,
    unsigned* __restrict__ counters, int P, int G, std::int64_t KS, double* __restrict__ f2,
    double* __restrict__ vpair, FixupList fl) {
    const int lane = threadIdx.x & 63;
    // v_mfma_f64_16x16x4f64 returns D[(l/16) + 4e][l%16] in lane l, item e, while the
    // int8 count MFMA returns D[4(l/16) + e][l%16].  Feeding the left operand's rows
    // in the order pi(i) = 4(i%4) + i/4 makes the FP64 accumulator land in the same
    // lanes/items as the count: lane l of the left fragment reads plane lane lane_a.
    const int lane_a = (lane & 48) + 4 * (lane & 3) + ((lane >> 2) & 3);
    const std::int64_t KS64 = KS / 16;
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;

    const int home = static_cast<int>(blockIdx.x % kXcds);
    int xq = 0;
    for (;;) {
        int it = 0;
        if (lane == 0) it = claim_item(counters, queues, home, xq);
        // Uniform values in scalar registers keep every branch below uniform.
        it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
        if (it < 0) break;
        const F2Item w = items[it];
        const int b = __builtin_amdgcn_readfirstlane(w.b);
        const int gI0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(w.ti));
        const int gJ0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(w.tj));
        const std::int64_t t_beg = bt.k0[b] / 4;
        const int nsteps = bt.kpad[b] / 4;

        bool rowI[4], rowJ[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) { rowI[x] = gI0 + x < G; rowJ[x] = gJ0 + x < G; }

        d4 acc[4][4];
        i4 cnt[4][4];
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};
                cnt[x][y] = i4{0, 0, 0, 0};
            }

        for (int t = 0; t < nsteps; ++t) {
            double aI[4], vI[4], qI[4], aJ[4], vJ[4], qJ[4];
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                aI[x] = vI[x] = qI[x] = aJ[x] = vJ[x] = qJ[x] = 0.0;
                if (rowI[x]) {
                    const std::int64_t o = ((gI0 + x) * KS + t_beg + t) * 64 + lane_a;
                    aI[x] = pA[o];
                    vI[x] = pV[o];
                    qI[x] = -2.0 * pQ[o];
                }
                if (rowJ[x]) {
                    const std::int64_t o = ((gJ0 + x) * KS + t_beg + t) * 64 + lane;
                    aJ[x] = pA[o];
                    vJ[x] = pV[o];
                    qJ[x] = pQ[o];
                }
            }
#pragma unroll
            for (int x = 0; x < 4; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    if (!(rowI[x] && rowJ[y] && gI0 + x <= gJ0 + y)) continue;
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(aI[x], vJ[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(vI[x], aJ[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(qI[x], qJ[y], acc[x][y], 0, 0, 0);
                }
            if ((t & 15) == 15) {
                const std::int64_t u = (t_beg + t) / 16;
                i4 cI[4], cJ[4];
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    cI[x] = rowI[x] ? pV8[((gI0 + x) * KS64 + u) * 64 + lane] : i4{0, 0, 0, 0};
                    cJ[x] = rowJ[x] ? pV8[((gJ0 + x) * KS64 + u) * 64 + lane] : i4{0, 0, 0, 0};
                }
#pragma unroll
                for (int x = 0; x < 4; ++x)
#pragma unroll
                    for (int y = 0; y < 4; ++y) {
                        if (!(rowI[x] && rowJ[y] && gI0 + x <= gJ0 + y)) continue;
                        cnt[x][y] = __builtin_amdgcn_mfma_i32_16x16x64_i8(cI[x], cJ[y], cnt[x][y], 0, 0, 0);
                    }
            }
        }

        // Epilogue: with the row permutation above, D[i][j] of both accumulators sits
        // in lane j + 16*(i/4), item i%4.
        double* f2b = f2 + slab * b;
        double* vpb = vpair + slab * b;
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int gI = gI0 + x, gJ = gJ0 + y;
                if (!(rowI[x] && rowJ[y] && gI <= gJ)) continue;
                const int col = kRowGroup * gJ + (lane & 15);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = kRowGroup * gI + 4 * (lane >> 4) + e;
                    const bool ok = row < P && col < P && !(gI == gJ && row > col);
                    const double vp = static_cast<double>(cnt[x][y][e]);
                    const double val = vp > 0.0 ? acc[x][y][e] / vp : 0.0;
                    if (ok) {
                        f2b[row + static_cast<std::int64_t>(P) * col] = val;
                        f2b[col + static_cast<std::int64_t>(P) * row] = val;
                        vpb[row + static_cast<std::int64_t>(P) * col] = vp;
                        vpb[col + static_cast<std::int64_t>(P) * row] = vp;
                    }
                    fixup_append(fl, ok && vp > 0.0 && fabs(val) < fl.tau, b, row, col);
                }
            }
    }
}

__global__ void zero_u32_kernel(unsigned* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

void launch_zero_u32(unsigned* p, int n, hipStream_t st) {
    zero_u32_kernel<<<static_cast<unsigned>(ceil_div(n, 256)), 256, 0, st>>>(p, n);
    STP_LAUNCH_CHECK();
}

void launch_prep_fp64_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P, int G,
                             BlockTable bt, const std::int32_t* chunk_block, std::int64_t n_chunks,
                             std::int64_t KS, double* pA, double* pV, double* pQ, void* pV8, hipStream_t st) {
    if (n_chunks <= 0) return;
    prep_fp64_kernel<CountsSrc><<<static_cast<unsigned>(n_chunks), 256, 0, st>>>(
        CountsSrc{a2, n, ld}, P, G, bt, chunk_block, KS, pA, pV, pQ, static_cast<i4*>(pV8));
    STP_LAUNCH_CHECK();
}

void launch_prep_fp64_qvn(const double* q, const double* v, const double* n, int P, int G, BlockTable bt,
                          const std::int32_t* chunk_block, std::int64_t n_chunks, std::int64_t KS,
                          double* pA, double* pV, double* pQ, void* pV8, hipStream_t st) {
    if (n_chunks <= 0) return;
    prep_fp64_kernel<QvnSrc><<<static_cast<unsigned>(n_chunks), 256, 0, st>>>(
        QvnSrc{q, v, n, P}, P, G, bt, chunk_block, KS, pA, pV, pQ, static_cast<i4*>(pV8));
    STP_LAUNCH_CHECK();
}

void launch_f2_fp64(const double* pA, const double* pV, const double* pQ, const void* pV8, BlockTable bt,
                    const F2Item* items, const ItemQueues& queues, unsigned* counters, int P, int G,
                    std::int64_t KS, double* f2, double* vpair, const FixupList& fl, int n_workgroups,
                    hipStream_t st) {
    if (queues.off[kXcds] <= 0) return;
    launch_zero_u32(counters, kXcds, st);
    f2_fp64_kernel<<<static_cast<unsigned>(n_workgroups), 256, 0, st>>>(
        pA, pV, pQ, static_cast<const i4*>(pV8), bt, items, queues, counters, P, G, KS, f2, vpair, fl);
    STP_LAUNCH_CHECK();
}

}  // namespace stp
