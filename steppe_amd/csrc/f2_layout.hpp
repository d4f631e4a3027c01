// steppe_amd/csrc/f2_layout.hpp — the block table and operand-plane layout shared by
// the f2 prep and contraction kernels.
//
// Per jackknife block b (contiguous SNP range [s0_b, s0_b + len_b)), the SNP axis
// is padded to kpad_b = round_up(len_b, 64) and blocks are concatenated: block b
// occupies padded columns [k0_b, k0_b + kpad_b) of a K axis of length K_total.
// Padding columns carry V = 0 (and zero operands), so they contribute nothing.
//
// Populations are grouped in rows of 16 (G = ceil(P/16) groups; rows >= P are
// zero).  Every operand plane is stored in MFMA fragment order so that one
// wavefront load of 64 consecutive elements is exactly one MFMA operand:
//
//   FP64 planes (v_mfma_f64_16x16x4f64, 4-SNP K-steps, KS = K_total/4 steps):
//     plane[(g*KS + t)*64 + lane],  lane -> row 16g + lane%16, SNP 4t + lane/16
//   int8 V plane (v_mfma_i32_16x16x64_i8, 64-SNP K-steps, KS64 = K_total/64):
//     v8[(g*KS64 + u)*64 + lane] is 16 bytes: row 16g + lane%16,
//     SNPs 64u + 16*(lane/16) + 0..15
//
// The same fragment serves as the A operand (rows of the left population tile)
// and the B operand (rows of the right tile), because both MFMA operand layouts
// put the output index on lane%16 and the K index on lane/16.
#pragma once

#include <cstdint>

namespace stp {

constexpr int kRowGroup = 16;
constexpr int kKPad = 64;   // padded block granularity (SNPs)
constexpr int kTileRows = 64;  // rows per wavefront output tile (4 row groups)
// Split form: 11 digit planes of 2 bytes per (population, SNP) (f2_split.hip).
constexpr int kSplitPlanes = 11;

struct BlockTable {
    std::int64_t* s0;    // first SNP of block b
    std::int32_t* len;   // SNPs in block b
    std::int64_t* k0;    // first padded column of block b
    std::int32_t* kpad;  // padded columns of block b
};

struct F2Item {
    std::int32_t b;   // block
    std::int16_t ti;  // left 64-row tile
    std::int16_t tj;  // right 64-row tile (ti <= tj)
};

// MI355X has 8 accelerator dies (XCDs), each with its own L2, and the dispatcher
// hands workgroup i of a launch to die i % 8.  The persistent f2 kernels therefore
// keep one work queue per die: every block (all its tile pairs) is assigned to one
// die, so the tiles that share a block's operand panels run under the same L2.
constexpr int kXcds = 8;

// Items of die x are items[off[x] .. off[x+1]); counters[x] is that queue's cursor.
struct ItemQueues {
    std::int32_t off[kXcds + 1];
};

// Cancellation-prone entries.  Where a block's f2 is close to zero, the per-SNP terms
// (q_i - q_j)^2 - h_i - h_j cancel almost completely and the contraction's rounding
// (FP64 MFMA sums, or the split form's 48-bit digits) is large relative to the result.
// The contraction epilogues append every upper-triangle entry with 0 < vpair and
// |f2| < tau to this list; f2_fixup.hip then recomputes those entries from the
// frequencies with per-SNP FP64 terms and a double-double sum.  Entries past `cap`
// are not stored, and a scan over the whole tensor handles that case.
struct FixupList {
    unsigned* count;          // entries flagged by the launch (may exceed cap)
    std::int32_t* entries;    // 3 per entry: block, row, column (row <= column)
    unsigned cap;
    double tau;
};

// |f2| thresholds below which the contraction's result is recomputed.  Split path: its
// 48-bit digits leave up to ~1.4e-14 absolute (measured over 684M entries; ~6e-14 worst
// case), so entries below 1e-4 are recomputed to keep every entry within 1e-9 relative.
// FP64 path: its sums are off by ~1e-18 on entries that cancel to ~0 and by ~1e-16 at
// |f2| ~ 1e-4, so the same threshold keeps it within 1e-12 relative; a larger one would
// send whole tensors of closely related populations (f2 ~ 1e-3 everywhere) through the
// recomputation.
constexpr double kFixupTauFp64 = 1e-4;
constexpr double kFixupTauSplit = 1e-4;

#ifdef __HIPCC__
// Claims the next item for a worker homed on die `home`: its own queue first, then
// the other dies' queues in turn once it is drained (`xq` counts the queues already
// exhausted by this worker and starts at 0).  Returns -1 when every queue is empty.
// Called by one lane; at most kXcds atomics per call.
__device__ inline int claim_item(unsigned* counters, const ItemQueues& q, int home, int& xq) {
    while (xq < kXcds) {
        const int x = (home + xq) % kXcds;
        const int n = q.off[x + 1] - q.off[x];
        if (n > 0) {
            const unsigned v = atomicAdd(&counters[x], 1u);
            if (static_cast<int>(v) < n) return q.off[x] + static_cast<int>(v);
        }
        ++xq;
    }
    return -1;
}

// Appends (b, row, col) for every lane of the wavefront whose flag is set; all 64 lanes
// call it together.  One atomic per wavefront with a flagged lane.
__device__ inline void fixup_append(const FixupList& fl, bool flag, int b, int row, int col) {
    const unsigned long long m = __ballot(flag);
    if (m == 0ull) return;
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(fl.count, static_cast<unsigned>(__popcll(m)));
    base = __shfl(base, leader);
    if (flag) {
        const unsigned k = base + static_cast<unsigned>(__popcll(m & ((1ull << lane) - 1ull)));
        if (k < fl.cap) {
            fl.entries[3 * static_cast<std::int64_t>(k)] = b;
            fl.entries[3 * static_cast<std::int64_t>(k) + 1] = row;
            fl.entries[3 * static_cast<std::int64_t>(k) + 2] = col;
        }
    }
}
#endif

}  // namespace stp
