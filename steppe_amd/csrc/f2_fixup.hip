// steppe_amd/csrc/f2_fixup.hip — FP64 recomputation of cancellation-prone f2 entries.
//
// Where a block's f2 is close to zero the per-SNP terms (q_i - q_j)^2 - h_i - h_j
// cancel almost completely, and the contraction kernels' rounding (FP64 MFMA sums
// over centred operands, or the split form's 48-bit digits) is large relative to the
// result.  The contraction epilogues list every upper-triangle entry with
// 0 < vpair and |f2| < tau (FixupList, f2_layout.hpp); this file recomputes each
// listed entry the way the reference's CpuBackend does (src/device/cpu/
// cpu_backend.cpp:217-265): over the block's SNPs where both populations have data,
// the term t = (q_i - q_j)^2 - h_i - h_j is formed in double with each operation
// rounded (this file is built without FP contraction), h = q(1 - q)/max(n - 1, 1)
// (src/core/internal/f2_estimator.hpp:22-29); the terms are summed in double-double
// (the reference sums them in long double) and f2 = num / vpair.
//
// One wavefront per entry: lanes stride over the block's SNPs, then the 64
// double-double partial sums are merged by shuffles.  A second kernel covers the
// overflow case (more flagged entries than the list holds): it reads the flag count
// and, only if the list overflowed, scans the whole tensor for entries below tau.

#include "common.hpp"
#include "f2_layout.hpp"

namespace stp {

namespace {

struct dd {
    double hi, lo;
};

__device__ inline dd dd_add(dd a, double b) {
    // TwoSum(a.hi, b), the rounding error carried in lo.
    const double s = a.hi + b;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b - bb);
    return dd{s, a.lo + err};
}

__device__ inline dd dd_merge(dd a, dd b) {
    const double s = a.hi + b.hi;
    const double bb = s - a.hi;
    const double err = (a.hi - (s - bb)) + (b.hi - bb);
    const double lo = err + a.lo + b.lo;
    const double hi = s + lo;
    return dd{hi, lo - (hi - s)};
}

__device__ inline double het_bias(double q, double n) {
    const double d = n - 1.0;
    return q * (1.0 - q) / (d > 1.0 ? d : 1.0);
}

// Frequencies from the decode's counts: q = a2 / (2n) (the reference's ac / n).
struct FixCounts {
    const std::uint16_t* a2;
    const std::uint16_t* n;
    std::int64_t ld;
    __device__ bool get(int pop, std::int64_t s, double& q, double& nn) const {
        const std::int64_t o = static_cast<std::int64_t>(pop) * ld + s;
        const unsigned c = n[o];
        if (c == 0u) return false;
        nn = static_cast<double>(c);
        q = static_cast<double>(a2[o]) / (2.0 * nn);
        return true;
    }
};

// Frequencies from Q/V/N stacks (i + P*s), the reference's decode_af output.
struct FixQvn {
    const double* q;
    const double* v;
    const double* n;
    int P;
    __device__ bool get(int pop, std::int64_t s, double& qq, double& nn) const {
        const std::int64_t o = pop + static_cast<std::int64_t>(P) * s;
        if (v[o] == 0.0) return false;
        qq = q[o];
        nn = n[o];
        return true;
    }
};

// One wavefront recomputes entry (b, i, j) and writes it to (i, j) and (j, i).
template <class Src>
__device__ void fix_entry(const Src& src, const BlockTable& bt, int P, int b, int i, int j,
                          const double* __restrict__ vpair, double* __restrict__ f2) {
    const int lane = static_cast<int>(threadIdx.x & 63);
    const std::int64_t s0 = bt.s0[b];
    const int len = bt.len[b];
    dd acc{0.0, 0.0};
    for (int k = lane; k < len; k += 64) {
        double qi, ni, qj, nj;
        if (!src.get(i, s0 + k, qi, ni)) continue;
        if (!src.get(j, s0 + k, qj, nj)) continue;
        const double d = qi - qj;
        const double t = d * d - het_bias(qi, ni) - het_bias(qj, nj);
        acc = dd_add(acc, t);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const dd o{__shfl_xor(acc.hi, off), __shfl_xor(acc.lo, off)};
        acc = dd_merge(acc, o);
    }
    if (lane == 0) {
        const std::int64_t slab = static_cast<std::int64_t>(P) * P;
        const std::int64_t ij = slab * b + i + static_cast<std::int64_t>(P) * j;
        const double vp = vpair[ij];
        const double num = acc.hi + acc.lo;
        const double val = vp > 0.0 ? num / vp : 0.0;
        f2[ij] = val;
        f2[slab * b + j + static_cast<std::int64_t>(P) * i] = val;
    }
}

template <class Src>
__global__ __launch_bounds__(256) void f2_fixup_kernel(Src src, BlockTable bt, FixupList fl, int P,
                                                       const double* __restrict__ vpair, double* __restrict__ f2) {
    const unsigned flagged = *reinterpret_cast<volatile unsigned*>(fl.count);
    const unsigned n = flagged < fl.cap ? flagged : fl.cap;
    const unsigned wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const unsigned n_waves = (gridDim.x * blockDim.x) >> 6;
    for (unsigned k = wave; k < n; k += n_waves) {
        const std::int32_t* e = fl.entries + 3 * static_cast<std::int64_t>(k);
        const int b = __builtin_amdgcn_readfirstlane(e[0]);
        const int i = __builtin_amdgcn_readfirstlane(e[1]);
        const int j = __builtin_amdgcn_readfirstlane(e[2]);
        fix_entry(src, bt, P, b, i, j, vpair, f2);
    }
}

// Overflow path: every (block, column) pair is one work item; a wavefront tests 64 rows
// of the column at a time and recomputes the entries below tau one by one.
template <class Src>
__global__ __launch_bounds__(256) void f2_fixup_scan_kernel(Src src, BlockTable bt, FixupList fl, int P,
                                                            int n_block, const double* __restrict__ vpair,
                                                            double* __restrict__ f2) {
    const unsigned flagged = *reinterpret_cast<volatile unsigned*>(fl.count);
    if (flagged <= fl.cap) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const std::int64_t wave = (static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const std::int64_t n_waves = (static_cast<std::int64_t>(gridDim.x) * blockDim.x) >> 6;
    const std::int64_t slab = static_cast<std::int64_t>(P) * P;
    const std::int64_t n_items = static_cast<std::int64_t>(n_block) * P;
    for (std::int64_t it = wave; it < n_items; it += n_waves) {
        const int b = static_cast<int>(it / P);
        const int j = static_cast<int>(it - static_cast<std::int64_t>(b) * P);
        for (int i0 = 0; i0 <= j; i0 += 64) {
            const int i = i0 + lane;
            bool flag = false;
            if (i <= j) {
                const std::int64_t ij = slab * b + i + static_cast<std::int64_t>(P) * j;
                flag = vpair[ij] > 0.0 && fabs(f2[ij]) < fl.tau;
            }
            unsigned long long m = __ballot(flag);
            while (m) {
                const int r = __ffsll(static_cast<long long>(m)) - 1;
                m &= m - 1ull;
                fix_entry(src, bt, P, b, i0 + r, j, vpair, f2);
            }
        }
    }
}

template <class Src>
void launch_fixup_pair(const Src& src, BlockTable bt, const FixupList& fl, int P, int n_block, const double* vpair,
                       double* f2, int n_workgroups, hipStream_t st) {
    f2_fixup_kernel<Src><<<static_cast<unsigned>(n_workgroups), 256, 0, st>>>(src, bt, fl, P, vpair, f2);
    STP_LAUNCH_CHECK();
    f2_fixup_scan_kernel<Src><<<static_cast<unsigned>(n_workgroups), 256, 0, st>>>(src, bt, fl, P, n_block, vpair,
                                                                                  f2);
    STP_LAUNCH_CHECK();
}

}  // namespace

void launch_f2_fixup_counts(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, BlockTable bt,
                            const FixupList& fl, int P, int n_block, const double* vpair, double* f2,
                            int n_workgroups, hipStream_t st) {
    launch_fixup_pair(FixCounts{a2, n, ld}, bt, fl, P, n_block, vpair, f2, n_workgroups, st);
}

void launch_f2_fixup_qvn(const double* q, const double* v, const double* n, BlockTable bt, const FixupList& fl,
                         int P, int n_block, const double* vpair, double* f2, int n_workgroups, hipStream_t st) {
    launch_fixup_pair(FixQvn{q, v, n, P}, bt, fl, P, n_block, vpair, f2, n_workgroups, st);
}

}  // namespace stp
