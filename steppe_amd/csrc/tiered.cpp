// steppe_amd/csrc/tiered.cpp — f2 results larger than HBM: the host-memory and disk
// tiers (HipBackend::f2_counts_tiered, HipBackend::download).
//
// The reference streams such results through its HostRam / Disk engines
// (src/device/cuda/cuda_backend_f2_blocks.cu:349-603, src/device/cuda/block_sink.cu,
// tier choice src/device/tier_select.hpp:57-80 with kResidentTierVramFraction 0.70 and
// kHostTierRamFraction 0.60, include/steppe/config.hpp:41-45), re-decoding genotype
// columns per chunk because it keeps its inputs as FP64 Q/V/N.  Here the inputs are the
// decode's 16-bit counts, which stay resident (4 bytes per population and SNP), so only
// the output is chunked: consecutive block ranges are contracted into one HBM slab by the
// same kernels and copied into a pinned two-slot ring; while the device computes and
// copies range k, the host moves range k-1 into the host tensor or writes it into the
// f2.bin file at its block offsets (the file is the STPF2BK1 f2.bin the reference's disk
// sink writes).
#include <fcntl.h>
#include <sys/mman.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <fstream>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "f2bin.hpp"

namespace stp {

namespace {

constexpr std::size_t kChunkBytes = std::size_t{256} << 20;   // f2 + vpair of one streamed chunk

// Pinned host memory (hipHostMalloc / hipHostFree), movable.
class PinnedBuf {
public:
    explicit PinnedBuf(std::size_t n) : n_(n) {
        if (n) STP_HIP(hipHostMalloc(reinterpret_cast<void**>(&p_), n * sizeof(double), hipHostMallocDefault));
    }
    ~PinnedBuf() {
        if (p_) (void)hipHostFree(p_);
    }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    double* get() const { return p_; }

private:
    double* p_ = nullptr;
    std::size_t n_ = 0;
};

std::size_t free_host_bytes() {
    std::ifstream f("/proc/meminfo");
    std::string key;
    std::size_t kb = 0;
    std::string unit;
    while (f >> key >> kb >> unit)
        if (key == "MemAvailable:") return kb * 1024;
    return 0;
}

// Split [0, bytes) into page-aligned pieces and run f(offset, length) on up to 16 threads
// (first-touch page faults and page-cache writes of a multi-GB result scale with cores).
template <class F>
void parallel_pieces(std::size_t bytes, F&& f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const std::size_t T = std::min<std::size_t>({16, hw, std::max<std::size_t>(1, bytes >> 22)});
    if (T <= 1) {
        f(std::size_t{0}, bytes);
        return;
    }
    const std::size_t piece = (bytes / T + 4095) / 4096 * 4096;
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(T);
    for (std::size_t t = 0; t < T; ++t) {
        const std::size_t lo = std::min(bytes, t * piece), hi = std::min(bytes, lo + piece);
        if (lo >= hi) break;
        th.emplace_back([&, t, lo, hi] {
            try {
                f(lo, hi - lo);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    }
    for (auto& x : th) x.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

void pwrite_all(int fd, const void* buf, std::size_t bytes, std::uint64_t off, const std::string& path) {
    const char* p = static_cast<const char*>(buf);
    while (bytes > 0) {
        const ssize_t w = ::pwrite(fd, p, bytes, static_cast<off_t>(off));
        if (w < 0) {
            if (errno == EINTR) continue;
            throw ArgError("f2 disk tier: write failed (disk full?): " + path);
        }
        p += w;
        off += static_cast<std::uint64_t>(w);
        bytes -= static_cast<std::size_t>(w);
    }
}

}  // namespace

void HostArray::Free::operator()(double* p) const { std::free(p); }

void HostArray::alloc(std::size_t n) {
    constexpr std::size_t kHuge = std::size_t{2} << 20;
    const std::size_t bytes = std::max<std::size_t>((n * sizeof(double) + kHuge - 1) / kHuge * kHuge, kHuge);
    void* p = nullptr;
    if (::posix_memalign(&p, kHuge, bytes) != 0 || p == nullptr)
        throw std::bad_alloc();
    (void)::madvise(p, bytes, MADV_HUGEPAGE);   // advice only; ignored where THP is off
    p_.reset(static_cast<double*>(p));
    n_ = n;
}

std::size_t f2_working_set_bytes(const F2Plan& plan, int precision) {
    const std::size_t K = static_cast<std::size_t>(plan.K_total);
    const std::size_t rows16 = static_cast<std::size_t>(plan.G) * kRowGroup;
    const std::size_t v8 = rows16 * K;   // int8 V plane
    const std::size_t planes = precision == STP_PREC_FP64 ? 3 * sizeof(double) * rows16 * K
                                                          : 2 * kSplitPlanes * static_cast<std::size_t>(plan.G32) * 32 * K;
    const std::size_t upper = static_cast<std::size_t>(plan.P) * (plan.P + 1) / 2 * plan.n_block;
    const std::size_t fix = 12 * std::min<std::size_t>(std::max<std::size_t>(upper / 16, 1 << 16), 1 << 24);
    return planes + v8 + fix;
}

int choose_f2_tier(std::size_t result_bytes, std::size_t working_bytes, std::size_t free_vram,
                   std::size_t free_host) {
    if (result_bytes + working_bytes <= static_cast<std::size_t>(0.70 * static_cast<double>(free_vram)))
        return kTierResident;
    if (result_bytes <= static_cast<std::size_t>(0.60 * static_cast<double>(free_host))) return kTierHost;
    return kTierDisk;
}

F2Blocks* HipBackend::f2_counts_tiered(const std::uint16_t* a2, const std::uint16_t* n, std::int64_t ld, int P,
                                       std::int64_t M, const std::int32_t* block_id, int n_block, int precision,
                                       const TierRequest& tr) {
    STP_REQUIRE(tr.tier >= -1 && tr.tier <= kTierDisk, "f2: unknown tier");
    const std::size_t slab = static_cast<std::size_t>(P) * static_cast<std::size_t>(P);
    const std::size_t result_bytes = 2 * slab * static_cast<std::size_t>(n_block) * sizeof(double);
    std::size_t free_vram = 0, total_vram = 0;
    STP_HIP(hipMemGetInfo(&free_vram, &total_vram));
    const F2Plan full = make_f2_plan(P, M, block_id, n_block);   // validates the partition
    int tier = tr.tier;
    if (tier < 0) {
        // resident only when the result and the contraction's working set (planes, int8 V
        // plane, fix-up list) fit together, as the reference's select_output_tier adds
        // its resident working set (src/device/tier_select.hpp:62-72)
        tier = choose_f2_tier(result_bytes, f2_working_set_bytes(full, precision), free_vram, free_host_bytes());
    }
    STP_REQUIRE(tier != kTierDisk || !tr.disk_path.empty(),
                "f2: the result exceeds device and host memory; a disk path for the f2.bin is required");

    auto out = std::make_unique<F2Blocks>();
    out->P = P;
    out->n_block = n_block;
    out->tier = tier;
    out->block_sizes = full.len;
    if (tier == kTierResident) {
        const F2Plan& plan = plan_for(P, M, block_id, n_block, precision);
        out->f2.alloc(slab * n_block);
        out->vpair.alloc(slab * n_block);
        F2Source src;
        src.a2 = a2;
        src.n = n;
        src.ld = ld;
        run_f2(plan, precision, src, out->f2.get(), out->vpair.get());
        sync();
        return out.release();
    }

    // ---- streamed tiers ----
    std::int64_t nbc = tr.chunk_blocks;
    if (const char* env = std::getenv("STEPPE_F2_CHUNK_BLOCKS")) {
        const long long v = std::strtoll(env, nullptr, 10);
        if (v > 0 && nbc <= 0) nbc = v;
    }
    if (nbc <= 0) {
        // a quarter of free device memory, at most kChunkBytes per chunk (each chunk also
        // takes twice its size in pinned host memory for the ring)
        const double budget = std::min(0.25 * static_cast<double>(free_vram), static_cast<double>(kChunkBytes));
        nbc = static_cast<std::int64_t>(budget / (2.0 * slab * sizeof(double)));
    }
    nbc = std::max<std::int64_t>(1, std::min<std::int64_t>(nbc, n_block));
    const std::size_t chunk_elems = slab * static_cast<std::size_t>(nbc);
    DevBuf<double> d_f2(chunk_elems), d_vp(chunk_elems);
    PinnedBuf ring0(2 * chunk_elems), ring1(2 * chunk_elems);
    double* ring[2] = {ring0.get(), ring1.get()};
    hipEvent_t done[2];
    STP_HIP(hipEventCreateWithFlags(&done[0], hipEventDisableTiming));
    STP_HIP(hipEventCreateWithFlags(&done[1], hipEventDisableTiming));
    struct EventGuard {
        hipEvent_t* e;
        ~EventGuard() {
            (void)hipEventDestroy(e[0]);
            (void)hipEventDestroy(e[1]);
        }
    } eg{done};

    int fd = -1;
    F2BinHeader hdr = f2bin_header(P, n_block);
    if (tier == kTierHost) {
        out->host_f2.alloc(slab * n_block);
        out->host_vpair.alloc(slab * n_block);
    } else {
        fd = ::open(tr.disk_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        STP_REQUIRE(fd >= 0, "f2 disk tier: cannot open for write: " + tr.disk_path);
        out->disk_path = tr.disk_path;
    }
    struct FdGuard {
        int fd;
        ~FdGuard() {
            if (fd >= 0) ::close(fd);
        }
    } fg{fd};
    if (fd >= 0) pwrite_all(fd, &hdr, sizeof(hdr), 0, tr.disk_path);

    // Hand a finished range [b0, b1) in ring slot k to the sink.
    auto consume = [&](int k, int b0, int b1) {
        STP_HIP(hipEventSynchronize(done[k]));
        const std::size_t cnt = slab * static_cast<std::size_t>(b1 - b0);
        const double* f = ring[k];
        const double* v = ring[k] + chunk_elems;
        const std::size_t bytes = cnt * sizeof(double), at = slab * b0 * sizeof(double);
        if (tier == kTierHost) {
            char* hf = reinterpret_cast<char*>(out->host_f2.get()) + at;
            char* hv = reinterpret_cast<char*>(out->host_vpair.get()) + at;
            parallel_pieces(bytes, [&](std::size_t o, std::size_t len) {
                std::memcpy(hf + o, reinterpret_cast<const char*>(f) + o, len);
                std::memcpy(hv + o, reinterpret_cast<const char*>(v) + o, len);
            });
        } else {
            parallel_pieces(bytes, [&](std::size_t o, std::size_t len) {
                pwrite_all(fd, reinterpret_cast<const char*>(f) + o, len, hdr.f2_offset + at + o, tr.disk_path);
                pwrite_all(fd, reinterpret_cast<const char*>(v) + o, len, hdr.vpair_offset + at + o, tr.disk_path);
            });
        }
    };

    std::vector<std::int32_t> sub;
    int prev_b0 = -1, prev_b1 = -1, prev_k = 0, k = 0;
    for (int b0 = 0; b0 < n_block; b0 += static_cast<int>(nbc)) {
        const int b1 = static_cast<int>(std::min<std::int64_t>(n_block, b0 + nbc));
        // SNPs of blocks [b0, b1) (block_id is non-decreasing; blocks may be empty)
        const std::int64_t s_lo = std::lower_bound(block_id, block_id + M, b0) - block_id;
        const std::int64_t s_hi = std::lower_bound(block_id, block_id + M, b1) - block_id;
        const std::size_t cnt = slab * static_cast<std::size_t>(b1 - b0);
        if (s_hi > s_lo) {
            sub.resize(static_cast<std::size_t>(s_hi - s_lo));
            for (std::int64_t s = s_lo; s < s_hi; ++s) sub[s - s_lo] = block_id[s] - b0;
            const F2Plan& plan = plan_for(P, s_hi - s_lo, sub.data(), b1 - b0, precision);
            F2Source src;
            src.a2 = a2 + s_lo;
            src.n = n + s_lo;
            src.ld = ld;
            run_f2(plan, precision, src, d_f2.get(), d_vp.get());
        } else {
            STP_HIP(hipMemsetAsync(d_f2.get(), 0, cnt * sizeof(double), stream_));
            STP_HIP(hipMemsetAsync(d_vp.get(), 0, cnt * sizeof(double), stream_));
        }
        STP_HIP(hipMemcpyAsync(ring[k], d_f2.get(), cnt * sizeof(double), hipMemcpyDeviceToHost, stream_));
        STP_HIP(hipMemcpyAsync(ring[k] + chunk_elems, d_vp.get(), cnt * sizeof(double), hipMemcpyDeviceToHost,
                               stream_));
        STP_HIP(hipEventRecord(done[k], stream_));
        if (prev_b0 >= 0) consume(prev_k, prev_b0, prev_b1);
        prev_b0 = b0;
        prev_b1 = b1;
        prev_k = k;
        k ^= 1;
    }
    if (prev_b0 >= 0) consume(prev_k, prev_b0, prev_b1);
    sync();
    if (fd >= 0)
        pwrite_all(fd, out->block_sizes.data(), out->block_sizes.size() * sizeof(std::int32_t), hdr.block_sizes_offset,
                   tr.disk_path);
    return out.release();
}

void HipBackend::download(const F2Blocks& h, double* f2, double* vpair) {
    activate();
    const std::size_t tot = static_cast<std::size_t>(h.P) * h.P * h.n_block;
    if (h.tier == kTierResident) {
        if (f2) STP_HIP(hipMemcpyAsync(f2, h.f2.get(), tot * sizeof(double), hipMemcpyDeviceToHost, stream_));
        if (vpair) STP_HIP(hipMemcpyAsync(vpair, h.vpair.get(), tot * sizeof(double), hipMemcpyDeviceToHost, stream_));
        sync();
    } else if (h.tier == kTierHost) {
        if (f2) std::memcpy(f2, h.host_f2.get(), tot * sizeof(double));
        if (vpair) std::memcpy(vpair, h.host_vpair.get(), tot * sizeof(double));
    } else {
        std::ifstream in(h.disk_path, std::ios::binary);
        STP_REQUIRE(static_cast<bool>(in), "f2 disk tier: cannot reopen " + h.disk_path);
        F2BinHeader hdr{};
        in.read(reinterpret_cast<char*>(&hdr), sizeof(hdr));
        STP_REQUIRE(std::memcmp(hdr.magic, kF2BinMagic, 8) == 0 && hdr.P == h.P && hdr.n_block == h.n_block,
                    "f2 disk tier: " + h.disk_path + " no longer holds this tensor");
        if (f2) {
            in.seekg(static_cast<std::streamoff>(hdr.f2_offset));
            in.read(reinterpret_cast<char*>(f2), static_cast<std::streamsize>(tot * sizeof(double)));
        }
        if (vpair) {
            in.seekg(static_cast<std::streamoff>(hdr.vpair_offset));
            in.read(reinterpret_cast<char*>(vpair), static_cast<std::streamsize>(tot * sizeof(double)));
        }
        STP_REQUIRE(static_cast<bool>(in), "f2 disk tier: short read of " + h.disk_path);
    }
}

}  // namespace stp
