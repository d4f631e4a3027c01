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
// between three separately rounded GEMM outputs.  The pair count is the exact int8
// MFMA product V V^T of the count kernel (f2_split.hip), run before the contraction.

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
    static constexpr bool kSnpFastest = true;  // rows are SNP-contiguous
    const std::uint16_t* a2;
    const std::uint16_t* n;
    std::int64_t ld;
    struct Raw {
        unsigned a, c;
    };
    // Loads are unconditional (the caller passes a valid address), so that a batch of
    // them is in flight before any is used.
    __device__ Raw fetch(int r, std::int64_t s) const {
        const std::int64_t o = static_cast<std::int64_t>(r) * ld + s;
        return Raw{a2[o], n[o]};
    }
    __device__ static bool conv(const Raw& w, double& q, double& nd) {
        if (!w.c) return false;
        nd = static_cast<double>(w.c);
        q = static_cast<double>(w.a) / (2.0 * nd);  // == ac / n, correctly rounded
        return true;
    }
};

struct QvnSrc {
    static constexpr bool kSnpFastest = false;  // population-fastest (i + P*s)
    const double* q;
    const double* v;
    const double* n;
    int P;
    struct Raw {
        double q, v, n;
    };
    __device__ Raw fetch(int r, std::int64_t s) const {
        const std::int64_t o = r + static_cast<std::int64_t>(P) * s;
        return Raw{q[o], v[o], n[o]};
    }
    __device__ static bool conv(const Raw& w, double& qq, double& nd) {
        if (w.v == 0.0) return false;
        qq = w.q;
        nd = w.n;
        return true;
    }
};

// Workgroup barrier for the LDS stage only: waits for this wavefront's LDS operations,
// not for its outstanding global stores (__syncthreads would also wait for those).
__device__ inline void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class Src>
__global__ __launch_bounds__(256) void prep_fp64_kernel(
    Src src, int P, int G, BlockTable bt, const std::int32_t* __restrict__ chunk_block, std::int64_t KS,
    double* __restrict__ pA, double* __restrict__ pV, double* __restrict__ pQ,
    i4* __restrict__ pV8) {
    __shared__ double red_sum[4][64];
    __shared__ int red_cnt[4][64];
    __shared__ double centre[64];
    // One 16-row group of the chunk, operand values at [row][column].
    constexpr int kLd = 65;
    __shared__ double sA[kRowGroup * kLd], sV[kRowGroup * kLd], sQ[kRowGroup * kLd];

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
            // eight populations' loads in flight at a time; the sum keeps p ascending
            for (int p0 = pg; p0 < P; p0 += 32) {
                typename Src::Raw w[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) w[j] = src.fetch(p0 + 4 * j < P ? p0 + 4 * j : pg, snp0 + kk);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    double q, nd;
                    if (p0 + 4 * j < P && Src::conv(w[j], q, nd)) {
                        s += q;
                        ++cnt;
                    }
                }
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
    // Element i of this thread in a group's stage: lanes run along the source's
    // contiguous axis, so every row (or column) segment is read once, coalesced.
    auto elem = [&](int i, int& rr, int& kk) {
        const int e = tid + 256 * i;
        rr = Src::kSnpFastest ? (e >> 6) : (e & 15);
        kk = Src::kSnpFastest ? (e & 63) : (e >> 4);
    };
    // Raw loads of group g (elements outside the populations or the block's SNPs read
    // row 0 at the block's first SNP, a valid address, and are masked in the stage).
    auto fetch_group = [&](int g, typename Src::Raw (&w)[4]) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int rr, kk;
            elem(i, rr, kk);
            const int r = kRowGroup * g + rr;
            const bool ok = r < P && kk < live;
            w[i] = src.fetch(ok ? r : 0, ok ? snp0 + kk : bt.s0[b]);
        }
    };
    typename Src::Raw cur[4];
    if (G > 0) fetch_group(0, cur);
    for (int g = 0; g < G; ++g) {
        // The next group's loads are issued before this group is converted and written.
        typename Src::Raw nxt[4];
        fetch_group(g + 1 < G ? g + 1 : g, nxt);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            int rr, kk;
            elem(i, rr, kk);
            const int r = kRowGroup * g + rr;
            double A = 0.0, V = 0.0, Qc = 0.0;
            double q, nd;
            if (r < P && kk < live && Src::conv(cur[i], q, nd)) {
                const double nm1 = nd - 1.0;
                const double h = q * (1.0 - q) / (nm1 > 1.0 ? nm1 : 1.0);
                Qc = q - centre[kk];
                A = Qc * Qc - h;
                V = 1.0;
            }
            sA[rr * kLd + kk] = A;
            sV[rr * kLd + kk] = V;
            sQ[rr * kLd + kk] = Qc;
        }
        lds_barrier();
        // Operand planes in fragment order: thread -> (4-column step t4, fragment lane).
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int e = tid + 256 * i;
            const int t4 = e >> 6, ln = e & 63;
            const int rr = ln & 15, kk = 4 * t4 + (ln >> 4);
            const std::int64_t o = (static_cast<std::int64_t>(g) * KS + t0 + t4) * 64 + ln;
            pA[o] = sA[rr * kLd + kk];
            pV[o] = sV[rr * kLd + kk];
            pQ[o] = sQ[rr * kLd + kk];
        }
        if (tid < 64) {
            const int rr = tid & 15;
            const int kb = 16 * (tid >> 4);
            unsigned w[4] = {0, 0, 0, 0};
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (sV[rr * kLd + kb + j] != 0.0) w[j >> 2] |= 1u << (8 * (j & 3));
            i4 v;
            v[0] = static_cast<int>(w[0]);
            v[1] = static_cast<int>(w[1]);
            v[2] = static_cast<int>(w[2]);
            v[3] = static_cast<int>(w[3]);
            pV8[(static_cast<std::int64_t>(g) * (KS / 16) + u) * 64 + tid] = v;
        }
        lds_barrier();  // every reader of this group's stage is done
#pragma unroll
        for (int i = 0; i < 4; ++i) cur[i] = nxt[i];
    }
}

// ---------------------------------------------------------------------------
// contraction: persistent wavefronts pull work units from per-die queues (items are
// (block, 64x64 tile pair), ordered largest blocks first; a unit is an item, or half
// of one for 32x64 wavefront tiles) and accumulate TI x 4 sub-tiles of 16 x 16 over the
// block's K range: per 4-SNP step three v_mfma_f64_16x16x4 per sub-tile,
// A_i V_j + V_i A_j - 2 q'_i q'_j into one accumulator.  With PF the next step's
// operands are loaded before this step's MFMAs (one step of software pipelining).
// vpair comes from the pair-count kernel, which runs first.
// ---------------------------------------------------------------------------
template <int TI, bool PF>
__global__ __launch_bounds__(256, TI == 4 ? 2 : 3) void f2_fp64_kernel(
    const double* __restrict__ pA, const double* __restrict__ pV, const double* __restrict__ pQ, BlockTable bt,
    const F2Item* __restrict__ items, ItemQueues queues, unsigned* __restrict__ counters, int P, int G,
    std::int64_t KS, double* __restrict__ f2, const double* __restrict__ vpair, FixupList fl) {
    constexpr int kHalves = 4 / TI;
    const int lane = threadIdx.x & 63;
    // v_mfma_f64_16x16x4f64 returns D[(l/16) + 4e][l%16] in lane l, item e.  Feeding the
    // left operand's rows in the order pi(i) = 4(i%4) + i/4 puts D[i][j] in lane
    // j + 16*(i/4), item i%4: lane l of the left fragment reads plane lane lane_a.
    const int lane_a = (lane & 48) + 4 * (lane & 3) + ((lane >> 2) & 3);
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;

    const int home = static_cast<int>(blockIdx.x % kXcds);
    int xq = 0;
    for (;;) {
        int un = 0;
        if (lane == 0) un = claim_item(counters, queues, home, xq);
        // Uniform values in scalar registers keep every branch below uniform.
        un = __builtin_amdgcn_readfirstlane(__shfl(un, 0));
        if (un < 0) break;
        const F2Item w = items[un / kHalves];
        const int half = un - (un / kHalves) * kHalves;
        const int b = __builtin_amdgcn_readfirstlane(w.b);
        const int gI0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(w.ti)) + TI * half;
        const int gJ0 = 4 * __builtin_amdgcn_readfirstlane(static_cast<int>(w.tj));
        const std::int64_t t_beg = bt.k0[b] / 4;
        const int nsteps = bt.kpad[b] / 4;

        bool rowI[TI], rowJ[4];
#pragma unroll
        for (int x = 0; x < TI; ++x) rowI[x] = gI0 + x < G;
#pragma unroll
        for (int y = 0; y < 4; ++y) rowJ[y] = gJ0 + y < G;

        d4 acc[TI][4];
#pragma unroll
        for (int x = 0; x < TI; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};

        double aI[TI], vI[TI], qI[TI], aJ[4], vJ[4], qJ[4];
        auto load = [&](int t, double (&a_i)[TI], double (&v_i)[TI], double (&q_i)[TI], double (&a_j)[4],
                        double (&v_j)[4], double (&q_j)[4]) {
#pragma unroll
            for (int x = 0; x < TI; ++x) {
                a_i[x] = v_i[x] = q_i[x] = 0.0;
                if (rowI[x]) {
                    const std::int64_t o = ((gI0 + x) * KS + t_beg + t) * 64 + lane_a;
                    a_i[x] = pA[o];
                    v_i[x] = pV[o];
                    q_i[x] = -2.0 * pQ[o];
                }
            }
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                a_j[y] = v_j[y] = q_j[y] = 0.0;
                if (rowJ[y]) {
                    const std::int64_t o = ((gJ0 + y) * KS + t_beg + t) * 64 + lane;
                    a_j[y] = pA[o];
                    v_j[y] = pV[o];
                    q_j[y] = pQ[o];
                }
            }
        };
        auto mfma = [&](const double (&a_i)[TI], const double (&v_i)[TI], const double (&q_i)[TI],
                        const double (&a_j)[4], const double (&v_j)[4], const double (&q_j)[4]) {
#pragma unroll
            for (int x = 0; x < TI; ++x)
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    if (!(rowI[x] && rowJ[y] && gI0 + x <= gJ0 + y)) continue;
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_i[x], v_j[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(v_i[x], a_j[y], acc[x][y], 0, 0, 0);
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(q_i[x], q_j[y], acc[x][y], 0, 0, 0);
                }
        };
        if constexpr (PF) {
            load(0, aI, vI, qI, aJ, vJ, qJ);
            for (int t = 0; t < nsteps; ++t) {
                double aI2[TI], vI2[TI], qI2[TI], aJ2[4], vJ2[4], qJ2[4];
                // the last step re-reads its own operands (in bounds; never used)
                load(t + 1 < nsteps ? t + 1 : t, aI2, vI2, qI2, aJ2, vJ2, qJ2);
                mfma(aI, vI, qI, aJ, vJ, qJ);
#pragma unroll
                for (int x = 0; x < TI; ++x) {
                    aI[x] = aI2[x];
                    vI[x] = vI2[x];
                    qI[x] = qI2[x];
                }
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    aJ[y] = aJ2[y];
                    vJ[y] = vJ2[y];
                    qJ[y] = qJ2[y];
                }
            }
        } else {
            for (int t = 0; t < nsteps; ++t) {
                load(t, aI, vI, qI, aJ, vJ, qJ);
                mfma(aI, vI, qI, aJ, vJ, qJ);
            }
        }

        // Epilogue: with the row permutation above, D[i][j] sits in lane j + 16*(i/4),
        // item i%4.
        double* f2b = f2 + slab * b;
        const double* vpb = vpair + slab * b;
#pragma unroll
        for (int x = 0; x < TI; ++x)
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int gI = gI0 + x, gJ = gJ0 + y;
                if (!(rowI[x] && rowJ[y] && gI <= gJ)) continue;
                const int col = kRowGroup * gJ + (lane & 15);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int row = kRowGroup * gI + 4 * (lane >> 4) + e;
                    const bool ok = row < P && col < P && !(gI == gJ && row > col);
                    double vp = 0.0, val = 0.0;
                    if (ok) {
                        vp = vpb[row + static_cast<std::int64_t>(P) * col];
                        val = vp > 0.0 ? acc[x][y][e] / vp : 0.0;
                        f2b[row + static_cast<std::int64_t>(P) * col] = val;
                        f2b[col + static_cast<std::int64_t>(P) * row] = val;
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

void launch_f2_fp64(const double* pA, const double* pV, const double* pQ, BlockTable bt, const F2Item* items,
                    const ItemQueues& queues, unsigned* counters, int P, int G, std::int64_t KS, double* f2,
                    const double* vpair, const FixupList& fl, int n_workgroups, int wave_rows, bool prefetch,
                    hipStream_t st) {
    if (queues.off[kXcds] <= 0) return;
    launch_zero_u32(counters, kXcds, st);
    // work units: whole items (64-row wavefront tiles) or item halves (32-row tiles)
    ItemQueues qu = queues;
    const int halves = wave_rows == 32 ? 2 : 1;
    for (int x = 0; x <= kXcds; ++x) qu.off[x] = queues.off[x] * halves;
    const dim3 grid(static_cast<unsigned>(n_workgroups));
    auto go = [&](auto kernel) { kernel<<<grid, 256, 0, st>>>(pA, pV, pQ, bt, items, qu, counters, P, G, KS, f2, vpair, fl); };
    if (wave_rows == 32) {
        if (prefetch) go(f2_fp64_kernel<2, true>);
        else go(f2_fp64_kernel<2, false>);
    } else {
        if (prefetch) go(f2_fp64_kernel<4, true>);
        else go(f2_fp64_kernel<4, false>);
    }
    STP_LAUNCH_CHECK();
}

}  // namespace stp
