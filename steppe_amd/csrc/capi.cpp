// steppe_amd/csrc/capi.cpp — the extern "C" boundary (include/steppe_hip.h).
//
// Every entry point runs inside guard(): exceptions never cross the ABI; they are
// mapped to status codes and the message is kept per thread for stp_last_error().
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <new>
#include <string>

#include "engine.hpp"
#include "f2bin.hpp"
#include "steppe_hip.h"

struct stp_context {
    std::unique_ptr<stp::HipBackend> be;
};

struct stp_f2_handle {
    stp_context* ctx = nullptr;
    std::unique_ptr<stp::F2Blocks> blocks;
};

struct stp_f2_graph {
    stp_context* ctx = nullptr;
    std::unique_ptr<stp::F2Graph> g;
};

struct stp_qpgraph_topology {
    stp::QgTopology t;
};

struct stp_qpgraph_set {
    stp_context* ctx = nullptr;
    std::unique_ptr<stp::QgSet> s;
};

struct stp_qpgraph_list {
    std::vector<stp::qgenum::Candidate> c;
};

struct stp_qpgraph_search {
    stp_qpgraph_list cands;
    stp::QgSearchResult r;
};

namespace {

thread_local std::string g_last_error;

struct FormatError : std::runtime_error {
    explicit FormatError(const std::string& m) : std::runtime_error(m) {}
};
struct IoError : std::runtime_error {
    explicit IoError(const std::string& m) : std::runtime_error(m) {}
};

template <class F>
int guard(F&& f) {
    try {
        f();
        g_last_error.clear();
        return STP_OK;
    } catch (const stp::OomError& e) {
        g_last_error = e.what();
        return STP_ERR_OOM;
    } catch (const stp::DeviceError& e) {
        g_last_error = e.what();
        return STP_ERR_DEVICE;
    } catch (const stp::ArgError& e) {
        g_last_error = e.what();
        return STP_ERR_INVALID;
    } catch (const FormatError& e) {
        g_last_error = e.what();
        return STP_ERR_FORMAT;
    } catch (const IoError& e) {
        g_last_error = e.what();
        return STP_ERR_IO;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return STP_ERR_OOM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return STP_ERR_INTERNAL;
    }
}

stp::HipBackend& backend(stp_context* ctx) {
    STP_REQUIRE(ctx && ctx->be, "null context");
    return *ctx->be;
}

const stp::F2Blocks& blocks_of(const stp_f2_handle* h) {
    STP_REQUIRE(h && h->blocks, "null f2 handle");
    return *h->blocks;
}

// STPF2BK1 header (src/device/f2_disk_format.hpp:17-50): 64 bytes, little-endian.
using stp::F2BinHeader;
constexpr const char* kMagic = stp::kF2BinMagic;

F2BinHeader read_header(std::ifstream& f, const std::string& path) {
    F2BinHeader h{};
    f.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (!f) throw FormatError("f2.bin: truncated header: " + path);
    if (std::memcmp(h.magic, kMagic, 8) != 0) throw FormatError("f2.bin: bad magic (not STPF2BK1): " + path);
    if (h.version != 1u) throw FormatError("f2.bin: unsupported version " + std::to_string(h.version));
    if (h.dtype != 1u) throw FormatError("f2.bin: unsupported dtype " + std::to_string(h.dtype));
    if (h.P <= 0 || h.n_block <= 0) throw FormatError("f2.bin: degenerate shape");
    const std::uint64_t slab = static_cast<std::uint64_t>(h.P) * h.P * h.n_block * sizeof(double);
    if (h.f2_offset != 64 || h.vpair_offset != 64 + slab || h.block_sizes_offset != 64 + 2 * slab)
        throw FormatError("f2.bin: section offsets do not match P and n_block");
    f.seekg(0, std::ios::end);
    const std::uint64_t size = static_cast<std::uint64_t>(f.tellg());
    if (size < h.block_sizes_offset + 4ull * h.n_block) throw FormatError("f2.bin: truncated file: " + path);
    return h;
}

}  // namespace

extern "C" {

const char* stp_last_error(void) { return g_last_error.c_str(); }

int stp_open(int device, stp_context** out) {
    return guard([&] {
        STP_REQUIRE(out != nullptr, "stp_open: out is null");
        auto c = std::make_unique<stp_context>();
        c->be = std::make_unique<stp::HipBackend>(device);
        *out = c.release();
    });
}

int stp_close(stp_context* ctx) {
    return guard([&] { delete ctx; });
}

int stp_device_info(stp_context* ctx, char* name, int name_len, int32_t* n_cu) {
    return guard([&] {
        auto& be = backend(ctx);
        if (name && name_len > 0) {
            std::strncpy(name, be.name().c_str(), static_cast<std::size_t>(name_len - 1));
            name[name_len - 1] = '\0';
        }
        if (n_cu) *n_cu = be.n_cu();
    });
}

int stp_detect_ploidy(stp_context* ctx, const stp_tile* tile, int32_t* out_ploidy) {
    return guard([&] {
        STP_REQUIRE(tile && out_ploidy, "stp_detect_ploidy: null argument");
        const auto p = backend(ctx).detect_ploidy(*tile);
        for (std::size_t g = 0; g < p.size(); ++g) out_ploidy[g] = p[g];
    });
}

int stp_decode_counts(stp_context* ctx, const stp_tile* tile, uint16_t* d_a2, uint16_t* d_n, int64_t ld) {
    return guard([&] {
        STP_REQUIRE(tile && d_a2 && d_n, "stp_decode_counts: null argument");
        backend(ctx).decode_counts(*tile, d_a2, d_n, ld);
    });
}

int stp_decode_counts_noalloc(stp_context* ctx, const stp_tile* tile, uint16_t* d_a2, uint16_t* d_n, int64_t ld) {
    return guard([&] {
        STP_REQUIRE(tile && d_a2 && d_n, "stp_decode_counts_noalloc: null argument");
        backend(ctx).decode_counts_noalloc(*tile, d_a2, d_n, ld);
    });
}

int stp_decode_af(stp_context* ctx, const stp_tile* tile, double* q, double* v, double* n) {
    return guard([&] {
        STP_REQUIRE(tile && q && v && n, "stp_decode_af: null argument");
        backend(ctx).decode_af(*tile, q, v, n);
    });
}

int stp_f2_blocks_from_counts(stp_context* ctx, const uint16_t* d_a2, const uint16_t* d_n, int64_t ld,
                              int32_t P, int64_t M, const int32_t* block_id, int32_t n_block,
                              int32_t precision, double* d_f2, double* d_vpair) {
    return guard([&] {
        STP_REQUIRE(d_a2 && d_n && d_f2 && d_vpair, "stp_f2_blocks_from_counts: null argument");
        backend(ctx).f2_from_counts(d_a2, d_n, ld, P, M, block_id, n_block, precision, d_f2, d_vpair);
    });
}

int stp_f2_autotune(stp_context* ctx, const uint16_t* d_a2, const uint16_t* d_n, int64_t ld, int32_t P, int64_t M,
                    const int32_t* block_id, int32_t n_block, int32_t precision, double* d_f2, double* d_vpair,
                    char* report, int32_t report_len, int32_t* report_needed) {
    return guard([&] {
        STP_REQUIRE(d_a2 && d_n && d_f2 && d_vpair, "stp_f2_autotune: null argument");
        const std::string r =
            backend(ctx).f2_autotune(d_a2, d_n, ld, P, M, block_id, n_block, precision, d_f2, d_vpair);
        if (report_needed) *report_needed = static_cast<int32_t>(r.size() + 1);
        if (report && report_len > 0) {
            const std::size_t k = std::min<std::size_t>(r.size(), static_cast<std::size_t>(report_len - 1));
            std::memcpy(report, r.data(), k);
            report[k] = '\0';
        }
    });
}

int stp_compute_f2_blocks(stp_context* ctx, const double* Q, const double* V, const double* N, int32_t P,
                          int64_t M, const int32_t* block_id, int32_t n_block, int32_t precision, double* f2,
                          double* vpair, int32_t* block_sizes) {
    return guard([&] {
        STP_REQUIRE(f2 && vpair && block_sizes, "stp_compute_f2_blocks: null output");
        STP_REQUIRE(M == 0 || (Q && V && N), "stp_compute_f2_blocks: null Q/V/N");
        backend(ctx).f2_from_qvn(Q, V, N, P, M, block_id, n_block, precision, f2, vpair, block_sizes);
    });
}

int stp_extract_f2(stp_context* ctx, const stp_tile* tile, const int32_t* block_id, int32_t n_block,
                   int32_t precision, stp_f2_handle** out) {
    return guard([&] {
        STP_REQUIRE(tile && out, "stp_extract_f2: null argument");
        auto h = std::make_unique<stp_f2_handle>();
        h->ctx = ctx;
        h->blocks.reset(backend(ctx).extract_f2(*tile, block_id, n_block, precision));
        *out = h.release();
    });
}

int stp_f2_graph_create(stp_context* ctx, const stp_tile* tile, const int32_t* block_id, int32_t n_block,
                        int32_t precision, uint16_t* d_a2, uint16_t* d_n, int64_t ld, double* d_f2, double* d_vpair,
                        stp_f2_graph** out) {
    return guard([&] {
        STP_REQUIRE(tile && block_id && out, "stp_f2_graph_create: null argument");
        auto g = std::make_unique<stp_f2_graph>();
        g->ctx = ctx;
        g->g.reset(backend(ctx).f2_graph_create(*tile, block_id, n_block, precision, d_a2, d_n, ld, d_f2, d_vpair));
        *out = g.release();
    });
}

int stp_f2_graph_launch(stp_f2_graph* g) {
    return guard([&] {
        STP_REQUIRE(g && g->g, "stp_f2_graph_launch: null graph");
        backend(g->ctx).f2_graph_launch(*g->g);
    });
}

int stp_f2_graph_free(stp_f2_graph* g) {
    return guard([&] {
        if (g) backend(g->ctx).activate();
        delete g;
    });
}

int stp_extract_f2_tiered(stp_context* ctx, const stp_tile* tile, const int32_t* block_id, int32_t n_block,
                          int32_t precision, int32_t tier, const char* disk_path, int64_t chunk_blocks,
                          stp_f2_handle** out) {
    return guard([&] {
        STP_REQUIRE(tile && out, "stp_extract_f2_tiered: null argument");
        stp::TierRequest tr;
        tr.tier = tier;
        tr.disk_path = disk_path ? disk_path : "";
        tr.chunk_blocks = chunk_blocks;
        auto h = std::make_unique<stp_f2_handle>();
        h->ctx = ctx;
        h->blocks.reset(backend(ctx).extract_f2(*tile, block_id, n_block, precision, tr));
        *out = h.release();
    });
}

int stp_f2_handle_tier(const stp_f2_handle* h, int32_t* tier) {
    return guard([&] {
        STP_REQUIRE(tier, "null output");
        *tier = blocks_of(h).tier;
    });
}

int stp_extract_f2_files(stp_context* ctx, const stp_extract_request* req, stp_f2_handle** out,
                         stp_extract_info* info) {
    return guard([&] {
        STP_REQUIRE(req && out, "stp_extract_f2_files: null argument");
        STP_REQUIRE(req->geno && req->snp && req->ind, "stp_extract_f2_files: the .geno/.snp/.ind paths are required");
        stp::ExtractRequest r;
        r.geno = req->geno;
        r.snp = req->snp;
        r.ind = req->ind;
        STP_REQUIRE(req->pop_mode >= 0 && req->pop_mode <= 2, "stp_extract_f2_files: unknown pop_mode");
        r.pops.mode = static_cast<stp::PopSelect::Mode>(req->pop_mode);
        if (req->pop_mode == STP_POPS_EXPLICIT) {
            STP_REQUIRE(req->n_pops > 0 && req->pops, "extract_f2: pops needs at least one population name");
            for (int i = 0; i < req->n_pops; ++i) {
                STP_REQUIRE(req->pops[i], "stp_extract_f2_files: null population name");
                r.pops.labels.emplace_back(req->pops[i]);
            }
        }
        STP_REQUIRE(req->top_k >= 0 && req->min_n >= 0, "stp_extract_f2_files: negative top_k / min_n");
        r.pops.k = static_cast<std::size_t>(req->top_k);
        r.pops.min_n = static_cast<std::size_t>(req->min_n);
        r.maf_min = req->maf_min;
        r.maxmiss = req->maxmiss;
        r.autosomes_only = req->autosomes_only != 0;
        r.drop_monomorphic = req->drop_monomorphic != 0;
        r.transversions_only = req->transversions_only != 0;
        r.strand_mode = req->strand_mode;
        r.blgsize = req->blgsize;
        r.ploidy = req->ploidy;
        r.precision = req->precision;
        r.tier.tier = req->tier;
        r.tier.disk_path = req->disk_path ? req->disk_path : "";
        r.tier.chunk_blocks = req->chunk_blocks;
        stp::ExtractInfo inf;
        auto h = std::make_unique<stp_f2_handle>();
        h->ctx = ctx;
        h->blocks.reset(backend(ctx).extract_f2_files(r, inf));
        if (info) {
            info->n_snp_total = inf.n_snp_total;
            info->n_snp_kept = inf.n_snp_kept;
            info->n_pseudo_haploid = inf.n_pseudo_haploid;
            info->n_diploid = inf.n_diploid;
            info->P = h->blocks->P;
            info->n_block = h->blocks->n_block;
        }
        *out = h.release();
    });
}

int stp_f2_handle_label(const stp_f2_handle* h, int32_t i, char* buf, int32_t len, int32_t* n) {
    return guard([&] {
        STP_REQUIRE(h && h->blocks, "stp_f2_handle_label: null handle");
        STP_REQUIRE(i >= 0 && i < h->blocks->P, "stp_f2_handle_label: index out of range");
        const std::string s = i < static_cast<int32_t>(h->blocks->labels.size()) ? h->blocks->labels[i] : std::string();
        if (n) *n = static_cast<int32_t>(s.size());
        if (buf && len > 0) {
            const std::size_t k = std::min<std::size_t>(s.size(), static_cast<std::size_t>(len - 1));
            std::memcpy(buf, s.data(), k);
            buf[k] = '\0';
        }
    });
}

int stp_f2_handle_upload(stp_context* ctx, const double* f2, const double* vpair, const int32_t* block_sizes,
                         int32_t P, int32_t n_block, stp_f2_handle** out) {
    return guard([&] {
        STP_REQUIRE(out, "stp_f2_handle_upload: null out");
        auto h = std::make_unique<stp_f2_handle>();
        h->ctx = ctx;
        h->blocks.reset(backend(ctx).upload(f2, vpair, block_sizes, P, n_block));
        *out = h.release();
    });
}

int stp_f2_handle_shape(const stp_f2_handle* h, int32_t* P, int32_t* n_block) {
    return guard([&] {
        const auto& b = blocks_of(h);
        if (P) *P = b.P;
        if (n_block) *n_block = b.n_block;
    });
}

int stp_f2_handle_block_sizes(const stp_f2_handle* h, int32_t* block_sizes) {
    return guard([&] {
        const auto& b = blocks_of(h);
        STP_REQUIRE(block_sizes, "null output");
        std::memcpy(block_sizes, b.block_sizes.data(), b.block_sizes.size() * sizeof(int32_t));
    });
}

const double* stp_f2_handle_f2_device(const stp_f2_handle* h) {
    return (h && h->blocks && h->blocks->tier == stp::kTierResident) ? h->blocks->f2.get() : nullptr;
}

const double* stp_f2_handle_vpair_device(const stp_f2_handle* h) {
    return (h && h->blocks && h->blocks->tier == stp::kTierResident) ? h->blocks->vpair.get() : nullptr;
}

int stp_f2_handle_download(const stp_f2_handle* h, double* f2, double* vpair) {
    return guard([&] {
        backend(h->ctx).download(blocks_of(h), f2, vpair);
    });
}

int stp_f2_handle_free(stp_f2_handle* h) {
    return guard([&] {
        if (h && h->ctx && h->ctx->be) h->ctx->be->activate();
        delete h;
    });
}

int stp_qpadm_batch(stp_context* ctx, const stp_f2_handle* f2, int32_t n_models, const int32_t* target,
                    const int32_t* left_offsets, const int32_t* left, const int32_t* right_offsets,
                    const int32_t* right, const stp_qpadm_options* options, stp_qpadm_result* out) {
    return guard([&] {
        STP_REQUIRE(f2 && f2->blocks, "stp_qpadm_batch: null f2 handle");
        STP_REQUIRE(options, "stp_qpadm_batch: null options");
        STP_REQUIRE(f2->ctx == ctx, "stp_qpadm_batch: the f2 handle belongs to another context");
        backend(ctx).qpadm_batch(*f2->blocks, n_models, target, left_offsets, left, right_offsets, right, *options,
                                 out);
    });
}

int stp_f4(stp_context* ctx, const stp_f2_handle* f2, const int32_t* quartets, int64_t n, double* est, double* se,
           double* z, double* p) {
    return guard([&] {
        STP_REQUIRE(f2 && f2->blocks, "stp_f4: null f2 handle");
        STP_REQUIRE(f2->ctx == ctx, "stp_f4: the f2 handle belongs to another context");
        backend(ctx).fstat(*f2->blocks, 4, n, quartets, est, se, z, p);
    });
}

int stp_f3(stp_context* ctx, const stp_f2_handle* f2, const int32_t* triples, int64_t n, double* est, double* se,
           double* z, double* p) {
    return guard([&] {
        STP_REQUIRE(f2 && f2->blocks, "stp_f3: null f2 handle");
        STP_REQUIRE(f2->ctx == ctx, "stp_f3: the f2 handle belongs to another context");
        backend(ctx).fstat(*f2->blocks, 3, n, triples, est, se, z, p);
    });
}

int stp_f4ratio(stp_context* ctx, const stp_f2_handle* f2, const int32_t* tuples, int64_t n, double* alpha,
                double* se, double* z) {
    return guard([&] {
        STP_REQUIRE(f2 && f2->blocks, "stp_f4ratio: null f2 handle");
        STP_REQUIRE(f2->ctx == ctx, "stp_f4ratio: the f2 handle belongs to another context");
        backend(ctx).f4ratio(*f2->blocks, n, tuples, alpha, se, z);
    });
}

int stp_fstat_sweep(stp_context* ctx, const stp_f2_handle* f2, const stp_sweep_request* req, int64_t capacity,
                    int32_t* keys, double* est, double* se, double* z, double* p, int64_t* n_out,
                    uint64_t* enumerated, int32_t* capped, int32_t* status) {
    return guard([&] {
        STP_REQUIRE(f2 && f2->blocks, "stp_fstat_sweep: null f2 handle");
        STP_REQUIRE(f2->ctx == ctx, "stp_fstat_sweep: the f2 handle belongs to another context");
        STP_REQUIRE(req && n_out && enumerated && capped && status, "stp_fstat_sweep: null argument");
        STP_REQUIRE(capacity >= 0, "stp_fstat_sweep: negative capacity");
        STP_REQUIRE(capacity == 0 || (keys && est && se && z && p), "stp_fstat_sweep: null output array");
        STP_REQUIRE(req->n_subset >= 0 && (req->n_subset == 0 || req->subset), "stp_fstat_sweep: bad subset");
        stp::FstatSweepRequest r;
        r.k = req->k;
        r.filter = req->filter;
        r.min_z = req->min_z;
        r.top_k = req->top_k;
        r.sure = req->sure != 0;
        r.method = req->method;
        if (req->n_subset > 0) r.subset.assign(req->subset, req->subset + req->n_subset);
        stp::FstatSweepOut o;
        backend(ctx).fstat_sweep(*f2->blocks, r, o);
        const std::int64_t ns = static_cast<std::int64_t>(o.est.size());
        const std::int64_t w = std::min(ns, capacity);
        std::copy(o.keys.begin(), o.keys.begin() + 4 * w, keys);
        std::copy(o.est.begin(), o.est.begin() + w, est);
        std::copy(o.se.begin(), o.se.begin() + w, se);
        std::copy(o.z.begin(), o.z.begin() + w, z);
        std::copy(o.p.begin(), o.p.begin() + w, p);
        *n_out = ns;
        *enumerated = o.enumerated;
        *capped = o.capped ? 1 : 0;
        *status = o.status;
    });
}

int stp_qpgraph_parse(int32_t n_edges, const char* const* from, const char* const* to, int32_t n_names,
                      const char* const* names, stp_qpgraph_topology** out) {
    return guard([&] {
        STP_REQUIRE(out && from && to && (names || n_names == 0), "stp_qpgraph_parse: null argument");
        STP_REQUIRE(n_edges >= 0 && n_names >= 0, "stp_qpgraph_parse: negative count");
        std::vector<std::pair<std::string, std::string>> e;
        for (int i = 0; i < n_edges; ++i) {
            STP_REQUIRE(from[i] && to[i], "stp_qpgraph_parse: null edge endpoint");
            e.emplace_back(from[i], to[i]);
        }
        std::vector<std::string> nm;
        for (int i = 0; i < n_names; ++i) nm.emplace_back(names[i] ? names[i] : "");
        auto t = std::make_unique<stp_qpgraph_topology>();
        t->t = stp::parse_qpgraph(e, nm);
        *out = t.release();
    });
}

int stp_qpgraph_topology_free(stp_qpgraph_topology* t) {
    return guard([&] { delete t; });
}

int stp_qpgraph_topology_dims(const stp_qpgraph_topology* t, int32_t* dims) {
    return guard([&] {
        STP_REQUIRE(t && dims, "stp_qpgraph_topology_dims: null argument");
        const auto& m = t->t;
        const int32_t d[6] = {m.npop, m.ne, m.nadmix, m.npair, m.npath, m.base};
        for (int i = 0; i < 6; ++i) dims[i] = d[i];
    });
}

namespace {
void topology_label(const stp::QgTopology& m, int32_t kind, int32_t index, char* buf, int32_t len) {
    STP_REQUIRE(buf && len > 0, "qpgraph label: null argument");
    const std::vector<std::string>* v = nullptr;
    switch (kind) {
        case 0: v = &m.leaves; break;
        case 1: v = &m.admix_from; break;
        case 2: v = &m.admix_to; break;
        case 3: v = &m.edge_from; break;
        case 4: v = &m.edge_to; break;
        default: throw stp::ArgError("qpgraph label: kind must be 0..4");
    }
    STP_REQUIRE(index >= 0 && index < static_cast<int32_t>(v->size()), "qpgraph label: index out of range");
    std::strncpy(buf, (*v)[index].c_str(), static_cast<std::size_t>(len - 1));
    buf[len - 1] = '\0';
}

std::vector<std::pair<std::string, std::string>> edge_list(int32_t n, const char* const* from, const char* const* to) {
    STP_REQUIRE(n >= 0 && (n == 0 || (from && to)), "qpgraph: null edge list");
    std::vector<std::pair<std::string, std::string>> e;
    for (int i = 0; i < n; ++i) {
        STP_REQUIRE(from[i] && to[i], "qpgraph: null edge endpoint");
        e.emplace_back(from[i], to[i]);
    }
    return e;
}

std::vector<std::string> label_list(int32_t n, const char* const* names) {
    STP_REQUIRE(n >= 0 && (n == 0 || names), "qpgraph: null label list");
    std::vector<std::string> v;
    for (int i = 0; i < n; ++i) {
        STP_REQUIRE(names[i], "qpgraph: null label");
        v.emplace_back(names[i]);
    }
    return v;
}
}  // namespace

int stp_qpgraph_topology_label(const stp_qpgraph_topology* t, int32_t kind, int32_t index, char* buf,
                               int32_t len) {
    return guard([&] {
        STP_REQUIRE(t, "stp_qpgraph_topology_label: null argument");
        topology_label(t->t, kind, index, buf, len);
    });
}

// ---- topology enumeration and search ----------------------------------------------------
int stp_qpgraph_enumerate(int32_t n_leaves, const char* const* leaves, int32_t max_nadmix, stp_qpgraph_list** out) {
    return guard([&] {
        STP_REQUIRE(out, "stp_qpgraph_enumerate: null output");
        STP_REQUIRE(n_leaves >= 1 && n_leaves <= 9, "stp_qpgraph_enumerate: 1..9 leaves");
        auto l = std::make_unique<stp_qpgraph_list>();
        l->c = stp::qgenum::enumerate_space(label_list(n_leaves, leaves), max_nadmix);
        *out = l.release();
    });
}

int stp_qpgraph_neighbors(int32_t n_edges, const char* const* from, const char* const* to, int32_t nadmix,
                          uint64_t hash, int32_t n_leaves, const char* const* leaves, int32_t max_nadmix,
                          stp_qpgraph_list** out) {
    return guard([&] {
        STP_REQUIRE(out, "stp_qpgraph_neighbors: null output");
        STP_REQUIRE(nadmix == 0 || nadmix == 1, "stp_qpgraph_neighbors: moves are defined for 0 or 1 admixture events");
        stp::qgenum::Candidate cur;
        cur.edges = edge_list(n_edges, from, to);
        cur.nadmix = nadmix;
        cur.hash = hash;
        auto l = std::make_unique<stp_qpgraph_list>();
        l->c = stp::qgenum::neighbors(cur, label_list(n_leaves, leaves), max_nadmix);
        *out = l.release();
    });
}

int stp_qpgraph_graph_hash(int32_t n_edges, const char* const* from, const char* const* to, uint64_t* hash) {
    return guard([&] {
        STP_REQUIRE(hash, "stp_qpgraph_graph_hash: null output");
        *hash = stp::qgenum::graph_hash(edge_list(n_edges, from, to));
    });
}

int stp_qpgraph_list_size(const stp_qpgraph_list* l, int32_t* n) {
    return guard([&] {
        STP_REQUIRE(l && n, "stp_qpgraph_list_size: null argument");
        *n = static_cast<int32_t>(l->c.size());
    });
}

int stp_qpgraph_list_get(const stp_qpgraph_list* l, int32_t i, int32_t* nadmix, int32_t* id, uint64_t* hash,
                         int32_t* n_edges) {
    return guard([&] {
        STP_REQUIRE(l, "stp_qpgraph_list_get: null list");
        STP_REQUIRE(i >= 0 && i < static_cast<int32_t>(l->c.size()), "stp_qpgraph_list_get: index out of range");
        const auto& c = l->c[i];
        if (nadmix) *nadmix = c.nadmix;
        if (id) *id = c.id;
        if (hash) *hash = c.hash;
        if (n_edges) *n_edges = static_cast<int32_t>(c.edges.size());
    });
}

int stp_qpgraph_list_edge(const stp_qpgraph_list* l, int32_t i, int32_t k, char* from, char* to, int32_t len) {
    return guard([&] {
        STP_REQUIRE(l && from && to && len > 0, "stp_qpgraph_list_edge: null argument");
        STP_REQUIRE(i >= 0 && i < static_cast<int32_t>(l->c.size()), "stp_qpgraph_list_edge: index out of range");
        const auto& e = l->c[i].edges;
        STP_REQUIRE(k >= 0 && k < static_cast<int32_t>(e.size()), "stp_qpgraph_list_edge: edge out of range");
        std::strncpy(from, e[k].first.c_str(), static_cast<std::size_t>(len - 1));
        from[len - 1] = '\0';
        std::strncpy(to, e[k].second.c_str(), static_cast<std::size_t>(len - 1));
        to[len - 1] = '\0';
    });
}

int stp_qpgraph_list_free(stp_qpgraph_list* l) {
    return guard([&] { delete l; });
}

int stp_qpgraph_search_run(stp_context* ctx, const stp_f2_handle* f2, int32_t n_names, const char* const* names,
                           int32_t n_pops, const char* const* pops, const stp_qpgraph_search_options* options,
                           stp_qpgraph_search** out) {
    return guard([&] {
        STP_REQUIRE(out && f2 && f2->blocks && options, "stp_qpgraph_search_run: null argument");
        STP_REQUIRE(f2->ctx == ctx, "stp_qpgraph_search_run: the f2 handle belongs to another context");
        STP_REQUIRE(n_pops <= 8, "qpgraph-search: at most 8 populations (the bounded space grows factorially)");
        STP_REQUIRE(options->numstart >= 1 && options->maxit >= 0 && options->tol >= 0.0,
                    "qpgraph-search: numstart must be >= 1, maxit and tol >= 0");
        stp::QgSearchOptions o;
        o.max_nadmix = options->max_nadmix;
        o.numstart = options->numstart;
        o.diag_f3 = options->diag_f3;
        o.fit.maxit = options->maxit;
        o.fit.gtol = options->tol * 1e-2;
        o.fit.xtol = options->tol * 1e-3;
        o.fit.fudge = options->fudge;
        o.fit.constrained = options->constrained != 0;
        o.run_heuristic = options->run_heuristic != 0;
        o.heuristic_seeds = options->heuristic_seeds;
        auto s = std::make_unique<stp_qpgraph_search>();
        s->r = backend(ctx).qpgraph_search(*f2->blocks, label_list(n_names, names), label_list(n_pops, pops), o);
        s->cands.c = s->r.cands;
        *out = s.release();
    });
}

int stp_qpgraph_search_get_summary(const stp_qpgraph_search* s, stp_qpgraph_search_summary* out) {
    return guard([&] {
        STP_REQUIRE(s && out, "stp_qpgraph_search_get_summary: null argument");
        const auto& r = s->r;
        stp_qpgraph_search_summary m{};
        m.status = r.status;
        m.n_trees = r.n_trees;
        m.n_admix1 = r.n_admix1;
        m.n_candidates = r.n_candidates;
        m.n_fitted = static_cast<int32_t>(r.fitted.size());
        m.best = r.best;
        m.best_candidate = r.best >= 0 ? r.fitted[r.best] : -1;
        m.best_nadmix = r.best >= 0 ? r.cands[r.fitted[r.best]].nadmix : 0;
        m.best_hash = r.best >= 0 ? r.cands[r.fitted[r.best]].hash : 0;
        m.best_score = r.best_score;
        m.second_best_score = r.second_best_score;
        m.heuristic_recovered = r.heuristic_recovered ? 1 : 0;
        m.n_seeds = static_cast<int32_t>(r.seed_hashes.size());
        m.fit_all_wall_ms = r.fit_all_wall_ms;
        m.topologies_per_s = r.topologies_per_s;
        *out = m;
    });
}

const stp_qpgraph_list* stp_qpgraph_search_candidates(const stp_qpgraph_search* s) { return s ? &s->cands : nullptr; }

int stp_qpgraph_search_scores(const stp_qpgraph_search* s, int32_t* candidate, double* score, double* restart_spread) {
    return guard([&] {
        STP_REQUIRE(s, "stp_qpgraph_search_scores: null search");
        const auto& r = s->r;
        for (std::size_t k = 0; k < r.fitted.size(); ++k) {
            if (candidate) candidate[k] = r.fitted[k];
            if (score) score[k] = k < r.score.size() ? r.score[k] : std::nan("");
            if (restart_spread) restart_spread[k] = k < r.restart_spread.size() ? r.restart_spread[k] : 0.0;
        }
    });
}

int stp_qpgraph_search_seeds(const stp_qpgraph_search* s, uint64_t* hashes) {
    return guard([&] {
        STP_REQUIRE(s && hashes, "stp_qpgraph_search_seeds: null argument");
        std::copy(s->r.seed_hashes.begin(), s->r.seed_hashes.end(), hashes);
    });
}

int stp_qpgraph_search_best_fit(const stp_qpgraph_search* s, int32_t* dims, double* weight, double* weight_lo,
                                double* weight_hi, double* edge_length, double* f3_fit, double* f_obs,
                                double* qdiag) {
    return guard([&] {
        STP_REQUIRE(s, "stp_qpgraph_search_best_fit: null search");
        const auto& r = s->r;
        STP_REQUIRE(r.best >= 0, "qpgraph-search: no graph was fitted");
        const auto& m = r.best_topo;
        if (dims) {
            const int32_t d[6] = {m.npop, m.ne, m.nadmix, m.npair, m.npath, m.base};
            std::copy(d, d + 6, dims);
        }
        auto put = [](double* dst, const std::vector<double>& v) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        put(weight, r.theta);
        put(weight_lo, r.theta_lo);
        put(weight_hi, r.theta_hi);
        put(edge_length, r.edge_length);
        put(f3_fit, r.f3_fit);
        put(f_obs, r.f_obs);
        put(qdiag, r.qdiag);
    });
}

int stp_qpgraph_search_best_label(const stp_qpgraph_search* s, int32_t kind, int32_t index, char* buf, int32_t len) {
    return guard([&] {
        STP_REQUIRE(s && s->r.best >= 0, "stp_qpgraph_search_best_label: no best graph");
        topology_label(s->r.best_topo, kind, index, buf, len);
    });
}

int stp_qpgraph_search_free(stp_qpgraph_search* s) {
    return guard([&] { delete s; });
}

int stp_qpgraph_set_create(stp_context* ctx, int32_t n_graphs, const stp_qpgraph_topology* const* topologies,
                           const stp_f2_handle* const* f2, double diag_f3, stp_qpgraph_set** out) {
    return guard([&] {
        STP_REQUIRE(out && topologies && f2 && n_graphs > 0, "stp_qpgraph_set_create: bad arguments");
        std::vector<const stp::QgTopology*> tp(n_graphs);
        std::vector<const stp::F2Blocks*> fb(n_graphs);
        for (int g = 0; g < n_graphs; ++g) {
            STP_REQUIRE(topologies[g] && f2[g] && f2[g]->blocks, "stp_qpgraph_set_create: null topology or handle");
            STP_REQUIRE(f2[g]->ctx == ctx, "stp_qpgraph_set_create: an f2 handle belongs to another context");
            tp[g] = &topologies[g]->t;
            fb[g] = f2[g]->blocks.get();
        }
        auto s = std::make_unique<stp_qpgraph_set>();
        s->ctx = ctx;
        s->s.reset(backend(ctx).qpgraph_set_create(tp, fb.data(), diag_f3));
        *out = s.release();
    });
}

int stp_qpgraph_set_free(stp_qpgraph_set* s) {
    return guard([&] {
        if (s && s->ctx && s->ctx->be) s->ctx->be->activate();
        delete s;
    });
}

int stp_qpgraph_set_status(const stp_qpgraph_set* s, int32_t* status) {
    return guard([&] {
        STP_REQUIRE(s && s->s && status, "stp_qpgraph_set_status: null argument");
        for (int g = 0; g < s->s->n; ++g) status[g] = s->s->host_status[g];
    });
}

int stp_qpgraph_set_strides(const stp_qpgraph_set* s, int32_t* dmax, int32_t* emax, int32_t* pmax) {
    return guard([&] {
        STP_REQUIRE(s && s->s && dmax && emax, "stp_qpgraph_set_strides: null argument");
        *dmax = std::max(s->s->dmax, 1);
        *emax = s->s->ne_max;
        if (pmax) *pmax = s->s->np_max;
    });
}

int stp_qpgraph_set_basis(stp_context* ctx, const stp_qpgraph_set* s, int32_t graph, double* f_obs, double* qdiag,
                          double* qinv) {
    return guard([&] {
        STP_REQUIRE(s && s->s && s->ctx == ctx, "stp_qpgraph_set_basis: bad set");
        backend(ctx).qpgraph_set_basis(*s->s, graph, f_obs, qdiag, qinv);
    });
}

int stp_qpgraph_eval(stp_context* ctx, stp_qpgraph_set* s, int32_t n_inst, const int32_t* inst_graph,
                     const double* theta, const double* bl, int32_t mode, double fudge, int32_t constrained,
                     double* score, double* grad_theta, double* grad_bl, double* bl_out, double* fit_out) {
    return guard([&] {
        STP_REQUIRE(s && s->s && s->ctx == ctx, "stp_qpgraph_eval: bad set");
        STP_REQUIRE(n_inst >= 0 && (n_inst == 0 || (inst_graph && score)), "stp_qpgraph_eval: null argument");
        backend(ctx).qpgraph_set_eval(*s->s, n_inst, inst_graph, theta, bl, mode, fudge, constrained, score,
                                      grad_theta, grad_bl, bl_out, fit_out);
    });
}

int stp_qpgraph_set_create_basis(stp_context* ctx, int32_t n_graphs, const stp_qpgraph_topology* const* topologies,
                                 const double* f_obs, const double* qinv, stp_qpgraph_set** out) {
    return guard([&] {
        STP_REQUIRE(out && topologies && f_obs && qinv && n_graphs > 0, "stp_qpgraph_set_create_basis: bad arguments");
        *out = nullptr;
        std::vector<const stp::QgTopology*> tp(n_graphs);
        for (int g = 0; g < n_graphs; ++g) {
            STP_REQUIRE(topologies[g], "stp_qpgraph_set_create_basis: null topology");
            tp[g] = &topologies[g]->t;
        }
        auto s = std::make_unique<stp_qpgraph_set>();
        s->ctx = ctx;
        s->s.reset(backend(ctx).qpgraph_set_from_basis(tp, f_obs, qinv));
        *out = s.release();
    });
}

namespace {
stp::QgFitOptions fit_options(const stp_qpgraph_fit_options* o) {
    stp::QgFitOptions f;
    if (o) {
        STP_REQUIRE(o->maxit >= 0 && o->tol >= 0.0, "qpgraph: maxit and tol must be >= 0");
        f.maxit = o->maxit;
        f.gtol = o->tol * 1e-2;
        f.xtol = o->tol * 1e-3;
        f.fudge = o->fudge;
        f.constrained = o->constrained != 0;
    }
    return f;
}
}  // namespace

int stp_qpgraph_fit(stp_context* ctx, stp_qpgraph_set* s, int32_t n_inst, const int32_t* inst_graph,
                    const double* theta0, const stp_qpgraph_fit_options* options, double* theta_out,
                    double* score_out, int32_t* iterations) {
    return guard([&] {
        STP_REQUIRE(s && s->s && s->ctx == ctx, "stp_qpgraph_fit: bad set");
        STP_REQUIRE(n_inst >= 0 && (n_inst == 0 || (inst_graph && theta0 && score_out)), "stp_qpgraph_fit: null argument");
        backend(ctx).qpgraph_set_fit(*s->s, n_inst, inst_graph, theta0, fit_options(options), theta_out, score_out,
                                     iterations);
    });
}

int stp_qpgraph_fit_fleet(stp_context* ctx, stp_qpgraph_set* s, const stp_qpgraph_fit_options* options, double* score,
                          double* restart_spread, double* theta, double* theta_lo, double* theta_hi,
                          double* edge_length, double* f3_fit, int32_t* iterations, int32_t* status) {
    return guard([&] {
        STP_REQUIRE(s && s->s && s->ctx == ctx, "stp_qpgraph_fit_fleet: bad set");
        STP_REQUIRE(score && status, "stp_qpgraph_fit_fleet: null argument");
        stp::QgFleetOut r;
        backend(ctx).qpgraph_set_fit_fleet(*s->s, options ? options->numstart : 10, fit_options(options), r);
        auto put = [](const auto& v, auto* dst) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        put(r.score, score);
        put(r.restart_spread, restart_spread);
        put(r.theta, theta);
        put(r.theta_lo, theta_lo);
        put(r.theta_hi, theta_hi);
        put(r.edge_length, edge_length);
        put(r.f3_fit, f3_fit);
        put(r.iterations, iterations);
        put(r.status, status);
    });
}

int stp_f2bin_write(const char* path, const double* f2, const double* vpair, const int32_t* block_sizes,
                    int32_t P, int32_t n_block) {
    return guard([&] {
        STP_REQUIRE(path && f2 && vpair && block_sizes, "stp_f2bin_write: null argument");
        STP_REQUIRE(P > 0 && n_block > 0, "stp_f2bin_write: degenerate f2 tensor");
        const std::uint64_t slab = static_cast<std::uint64_t>(P) * P * n_block * sizeof(double);
        const F2BinHeader h = stp::f2bin_header(P, n_block);
        std::ofstream o(path, std::ios::binary | std::ios::trunc);
        if (!o) throw IoError(std::string("cannot open for write: ") + path);
        o.write(reinterpret_cast<const char*>(&h), sizeof(h));
        o.write(reinterpret_cast<const char*>(f2), static_cast<std::streamsize>(slab));
        o.write(reinterpret_cast<const char*>(vpair), static_cast<std::streamsize>(slab));
        o.write(reinterpret_cast<const char*>(block_sizes), static_cast<std::streamsize>(4 * n_block));
        o.flush();
        if (!o) throw IoError(std::string("f2.bin write failed (disk full?): ") + path);
    });
}

int stp_f2bin_read_header(const char* path, int32_t* P, int32_t* n_block) {
    return guard([&] {
        STP_REQUIRE(path, "null path");
        std::ifstream f(path, std::ios::binary);
        if (!f) throw IoError(std::string("cannot open: ") + path);
        const F2BinHeader h = read_header(f, path);
        if (P) *P = h.P;
        if (n_block) *n_block = h.n_block;
    });
}

int stp_f2bin_read(const char* path, double* f2, double* vpair, int32_t* block_sizes) {
    return guard([&] {
        STP_REQUIRE(path, "null path");
        std::ifstream f(path, std::ios::binary);
        if (!f) throw IoError(std::string("cannot open: ") + path);
        const F2BinHeader h = read_header(f, path);
        const std::uint64_t slab = static_cast<std::uint64_t>(h.P) * h.P * h.n_block * sizeof(double);
        if (f2) {
            f.seekg(static_cast<std::streamoff>(h.f2_offset));
            f.read(reinterpret_cast<char*>(f2), static_cast<std::streamsize>(slab));
        }
        if (vpair) {
            f.seekg(static_cast<std::streamoff>(h.vpair_offset));
            f.read(reinterpret_cast<char*>(vpair), static_cast<std::streamsize>(slab));
        }
        if (block_sizes) {
            f.seekg(static_cast<std::streamoff>(h.block_sizes_offset));
            f.read(reinterpret_cast<char*>(block_sizes), static_cast<std::streamsize>(4 * h.n_block));
        }
        if (!f) throw FormatError(std::string("f2.bin: read failed: ") + path);
    });
}

int stpx_f2_tier_choice(int32_t P, int64_t M, const int32_t* block_id, int32_t n_block, int32_t precision,
                        uint64_t free_vram, uint64_t free_host, uint64_t* working_bytes, int32_t* tier) {
    return guard([&] {
        STP_REQUIRE(working_bytes && tier, "stpx_f2_tier_choice: null output");
        const stp::F2Plan plan = stp::make_f2_plan(P, M, block_id, n_block);
        const std::size_t result = 2 * static_cast<std::size_t>(P) * P * n_block * sizeof(double);
        *working_bytes = stp::f2_working_set_bytes(plan, precision);
        *tier = stp::choose_f2_tier(result, *working_bytes, free_vram, free_host);
    });
}

int stpx_timing(stp_context* ctx, int32_t enable) {
    return guard([&] { backend(ctx).set_timing(enable != 0); });
}

int stpx_timing_read(stp_context* ctx, double* ms, int64_t* count, int32_t reset) {
    return guard([&] { backend(ctx).read_timing(ms, count, reset != 0); });
}

int stpx_synth_panel(stp_context* ctx, uint64_t seed, int32_t P, int64_t M, int64_t n_ind,
                     int64_t bytes_per_record, const float* fst, const float* pop_missing,
                     const int32_t* ind_pop, const int8_t* ind_haploid, uint8_t* d_packed) {
    return guard([&] {
        STP_REQUIRE(fst && pop_missing && ind_pop && ind_haploid && d_packed, "stpx_synth_panel: null argument");
        backend(ctx).synth_panel(seed, P, M, n_ind, bytes_per_record, fst, pop_missing, ind_pop, ind_haploid,
                                 d_packed);
    });
}

}  // extern "C"
