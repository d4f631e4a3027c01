// steppe_amd/csrc/f2bin.hpp — the STPF2BK1 f2.bin header (src/device/f2_disk_format.hpp:18-60):
// 64 bytes (magic, version 1, dtype 1 = FP64, P, n_block, section offsets), then the f2
// region, the vpair region (each P*P*n_block doubles, block-major, i + P*j + P*P*b) and
// n_block int32 block sizes.
#pragma once

#include <cstdint>

namespace stp {

struct F2BinHeader {
    char magic[8];
    std::uint32_t version;
    std::uint32_t dtype;
    std::int32_t P;
    std::int32_t n_block;
    std::uint64_t f2_offset;
    std::uint64_t vpair_offset;
    std::uint64_t block_sizes_offset;
    std::uint8_t reserved[16];
};
static_assert(sizeof(F2BinHeader) == 64, "STPF2BK1 header is 64 bytes");
constexpr char kF2BinMagic[8] = {'S', 'T', 'P', 'F', '2', 'B', 'K', '1'};

// The header of a P x P x n_block file with the sections back to back.
inline F2BinHeader f2bin_header(std::int32_t P, std::int32_t n_block) {
    F2BinHeader h{};
    for (int i = 0; i < 8; ++i) h.magic[i] = kF2BinMagic[i];
    h.version = 1;
    h.dtype = 1;
    h.P = P;
    h.n_block = n_block;
    const std::uint64_t slab = static_cast<std::uint64_t>(P) * static_cast<std::uint64_t>(P) *
                               static_cast<std::uint64_t>(n_block) * sizeof(double);
    h.f2_offset = 64;
    h.vpair_offset = 64 + slab;
    h.block_sizes_offset = 64 + 2 * slab;
    return h;
}

}  // namespace stp
