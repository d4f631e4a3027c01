// oracle/ref_harness.cpp — a C-ABI shim over the REAL reference CpuBackend.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with the
// reference's own CPU sources where they lie under /root/reference (no copy of them
// is kept in this repository) into oracle/_ref/libsteppe_ref.so.  Tests use it to
// pin the restated oracle (cpu_ref.cpp) bit-for-bit, and bench.py may time it as
// the "reference" CPU baseline.  The product library never links it.
//
// Entry points (each calls the reference function named beside it):
//   ref_decode_af        ComputeBackend::decode_af          src/device/cpu/cpu_backend.cpp:267
//   ref_detect_ploidy    detect_sample_ploidy_device        src/device/cpu/cpu_backend.cpp:1200
//   ref_assign_blocks    core::assign_blocks                src/core/domain/block_partition_rule.cpp:56
//   ref_f2_blocks        ComputeBackend::compute_f2_blocks  src/device/cpu/cpu_backend.cpp:217
//   ref_qpadm            steppe::run_qpadm (host tensor)    src/core/qpadm/qpadm_fit.cpp:216
//   ref_qpgraph_fit      steppe::run_qpgraph (host tensor)  src/core/qpadm/qpgraph_fit.cpp:127
//   ref_qpgraph_score    core::qpadm::qpgraph_score         src/core/qpadm/qpgraph_objective.hpp:163
//   ref_f4               steppe::run_f4 (host tensor)       src/core/qpadm/f4.cpp:106
//   ref_f3               steppe::run_f3 (host tensor)       src/core/qpadm/f3.cpp:105
//   ref_f4ratio          steppe::run_f4ratio (host tensor)  src/core/qpadm/f4ratio.cpp:94
//   ref_qg_enumerate     core::qpadm::enumerate_bounded_space  src/core/qpadm/qpgraph_enumerate.cpp:229
//   ref_qg_neighbors     core::qpadm::topology_neighbors       src/core/qpadm/qpgraph_enumerate.cpp:383
//   ref_qg_graph_hash    core::qpadm::graph_hash               src/core/qpadm/qpgraph_enumerate.cpp:183
//   ref_qg_search        steppe::run_qpgraph_search (host tensor)  src/core/qpadm/qpgraph_search.cpp:290
//   ref_extract_*        the CPU branch of steppe::run_extract_f2 (src/app/extract_f2_core.cpp:
//                        96-407, on_device == false) over the reference's own GenoReader,
//                        read_ind, read_snp, ComputeBackend::transpose_to_canonical,
//                        detect_sample_ploidy_device, decode_af, build_snp_keep_mask,
//                        assign_blocks and compute_f2_blocks; the glue between those calls
//                        follows that branch line by line

#include <array>
#include <cstdint>
#include <cstring>
#include <exception>
#include <span>
#include <string>
#include <vector>

#include "core/domain/block_partition_rule.hpp"
#include "io/filter/include_exclude.hpp"
#include "io/filter/snp_filter.hpp"
#include "io/geno_reader.hpp"
#include "io/genotype_source.hpp"
#include "io/ind_reader.hpp"
#include "io/snp_reader.hpp"
#include "core/qpadm/qpgraph_enumerate.hpp"
#include "core/qpadm/qpgraph_model.hpp"
#include "core/qpadm/qpgraph_objective.hpp"
#include "core/internal/views.hpp"
#include "device/backend.hpp"
#include "device/backend_factory.hpp"
#include "device/resources.hpp"
#include "steppe/f3.hpp"
#include "steppe/f4.hpp"
#include "steppe/f4ratio.hpp"
#include "steppe/fstats.hpp"
#include "steppe/qpadm.hpp"
#include "steppe/qpgraph.hpp"
#include "steppe/qpgraph_search.hpp"

#include "stp_qpadm_result.h"

namespace {

steppe::device::Resources cpu_resources() {
    steppe::device::Resources res;
    steppe::device::PerGpuResources cpu;
    cpu.device_id = -1;
    cpu.backend = steppe::device::make_cpu_backend();
    res.gpus.push_back(std::move(cpu));
    return res;
}

steppe::DecodeTileView tile_view(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                                 std::size_t n_ind, const std::size_t* offs, int n_pop,
                                 const int* sample_ploidy, int ploidy, int detect) {
    steppe::DecodeTileView t;
    t.packed = packed;
    t.bytes_per_record = bpr;
    t.n_snp = n_snp;
    t.n_individuals = n_ind;
    t.pop_offsets = offs;
    t.n_pop = n_pop;
    t.sample_ploidy = sample_ploidy;
    t.ploidy = ploidy;
    t.detect_ploidy_on_device = detect != 0;
    return t;
}

steppe::F2BlockTensor host_tensor(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb) {
    steppe::F2BlockTensor t;
    t.P = P;
    t.n_block = nb;
    const std::size_t tot = static_cast<std::size_t>(P) * P * nb;
    t.f2.assign(f2, f2 + tot);
    if (vpair) t.vpair.assign(vpair, vpair + tot);
    t.block_sizes.assign(bsz, bsz + nb);
    return t;
}

template <class T>
void put(T* dst, const std::vector<T>& v, std::size_t cap) {
    for (std::size_t k = 0; k < v.size() && k < cap; ++k) dst[k] = v[k];
}

}  // namespace

extern "C" {

int ref_decode_af(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                  std::size_t n_ind, const std::uint64_t* pop_offsets, int n_pop,
                  const std::int32_t* sample_ploidy, int ploidy, int detect, double* q,
                  double* v, double* n) {
    try {
        std::vector<std::size_t> offs(pop_offsets, pop_offsets + n_pop + 1);
        auto be = steppe::device::make_cpu_backend();
        const auto r = be->decode_af(tile_view(packed, bpr, n_snp, n_ind, offs.data(), n_pop,
                                               sample_ploidy, ploidy, detect));
        std::memcpy(q, r.q.data(), r.q.size() * sizeof(double));
        std::memcpy(v, r.v.data(), r.v.size() * sizeof(double));
        std::memcpy(n, r.n.data(), r.n.size() * sizeof(double));
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_detect_ploidy(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp,
                      std::size_t n_ind, std::int32_t* out) {
    try {
        auto be = steppe::device::make_cpu_backend();
        const auto p = be->detect_sample_ploidy_device(
            tile_view(packed, bpr, n_snp, n_ind, nullptr, 0, nullptr, 2, 1));
        for (std::size_t g = 0; g < p.size(); ++g) out[g] = p[g];
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_assign_blocks(const std::int32_t* chrom, const double* genpos, std::int64_t M,
                      double blgsize, const double* physpos, double bp_window,
                      std::int32_t* out_block_id) {
    try {
        std::vector<int> c(chrom, chrom + M);
        std::span<const double> phys;
        if (physpos) phys = std::span<const double>(physpos, static_cast<std::size_t>(M));
        const auto bp = steppe::core::assign_blocks(
            c, std::span<const double>(genpos, static_cast<std::size_t>(M)), blgsize, phys,
            bp_window);
        for (std::size_t s = 0; s < bp.block_id.size(); ++s) out_block_id[s] = bp.block_id[s];
        return bp.n_block;
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_f2_blocks(const double* Q, const double* V, const double* N, int P, std::int64_t M,
                  const std::int32_t* block_id, int nb, double* f2, double* vpair,
                  std::int32_t* block_sizes) {
    try {
        auto be = steppe::device::make_cpu_backend();
        const steppe::core::MatView q{Q, P, static_cast<long>(M)};
        const steppe::core::MatView v{V, P, static_cast<long>(M)};
        const steppe::core::MatView n{N, P, static_cast<long>(M)};
        std::vector<int> bid(block_id, block_id + M);
        const auto t = be->compute_f2_blocks(q, v, n, bid.data(), nb, steppe::Precision{});
        std::memcpy(f2, t.f2.data(), t.f2.size() * sizeof(double));
        std::memcpy(vpair, t.vpair.data(), t.vpair.size() * sizeof(double));
        for (int b = 0; b < nb; ++b) block_sizes[b] = t.block_sizes[static_cast<std::size_t>(b)];
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// One qpAdm model through the reference's public driver on a host f2 tensor.
// left excludes the target (left_idx = [target] + left inside the driver).
int ref_qpadm(const double* f2, const double* vpair, const std::int32_t* block_sizes, int P,
              int nb, int target, const std::int32_t* left, int n_left,
              const std::int32_t* right, int n_right, double fudge, int als_iterations,
              int rank, double rank_alpha, int jackknife, stp_qpadm_result* out) {
    try {
        std::memset(out, 0, sizeof(*out));
        steppe::F2BlockTensor t;
        t.P = P;
        t.n_block = nb;
        const std::size_t tot = static_cast<std::size_t>(P) * P * nb;
        t.f2.assign(f2, f2 + tot);
        if (vpair) t.vpair.assign(vpair, vpair + tot);
        t.block_sizes.assign(block_sizes, block_sizes + nb);
        steppe::QpAdmModel m;
        m.target = target;
        m.left.assign(left, left + n_left);
        m.right.assign(right, right + n_right);
        m.model_index = 0;
        steppe::QpAdmOptions o;
        o.fudge = fudge;
        o.als_iterations = als_iterations;
        o.rank = rank;
        o.rank_alpha = rank_alpha;
        o.jackknife = static_cast<steppe::JackknifePolicy>(jackknife);
        auto res = cpu_resources();
        const steppe::QpAdmResult r = steppe::run_qpadm(t, m, o, res);
        out->status = static_cast<int>(r.status);
        out->nl = static_cast<int>(r.weight.size());
        out->dof = r.dof;
        out->est_rank = r.est_rank;
        out->f4rank = r.f4rank;
        out->p = r.p;
        out->chisq = r.chisq;
        put(out->weight, r.weight, STP_QPADM_MAX_NL);
        put(out->se, r.se, STP_QPADM_MAX_NL);
        put(out->z, r.z, STP_QPADM_MAX_NL);
        out->n_rank = static_cast<int>(r.rank_chisq.size());
        put(out->rank_chisq, r.rank_chisq, STP_QPADM_MAX_RANK);
        put(out->rank_dof, r.rank_dof, STP_QPADM_MAX_RANK);
        put(out->rankdrop_f4rank, r.rankdrop_f4rank, STP_QPADM_MAX_RANK);
        put(out->rankdrop_dof, r.rankdrop_dof, STP_QPADM_MAX_RANK);
        put(out->rankdrop_dofdiff, r.rankdrop_dofdiff, STP_QPADM_MAX_RANK);
        put(out->rankdrop_chisq, r.rankdrop_chisq, STP_QPADM_MAX_RANK);
        put(out->rankdrop_p, r.rankdrop_p, STP_QPADM_MAX_RANK);
        put(out->rankdrop_chisqdiff, r.rankdrop_chisqdiff, STP_QPADM_MAX_RANK);
        put(out->rankdrop_p_nested, r.rankdrop_p_nested, STP_QPADM_MAX_RANK);
        out->n_popdrop = static_cast<int>(r.popdrop_pat.size());
        for (std::size_t k = 0; k < r.popdrop_pat.size() && k < STP_QPADM_MAX_POPDROP; ++k) {
            std::strncpy(out->popdrop_pat[k], r.popdrop_pat[k].c_str(), STP_QPADM_MAX_NL);
            out->popdrop_wt[k] = r.popdrop_wt[k];
            out->popdrop_dof[k] = r.popdrop_dof[k];
            out->popdrop_f4rank[k] = r.popdrop_f4rank[k];
            out->popdrop_chisq[k] = r.popdrop_chisq[k];
            out->popdrop_p[k] = r.popdrop_p[k];
            out->popdrop_feasible[k] = r.popdrop_feasible[k] != 0;
        }
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

static std::vector<steppe::QpGraphEdge> qg_edges(int n, const char* const* from, const char* const* to) {
    std::vector<steppe::QpGraphEdge> e;
    for (int i = 0; i < n; ++i) e.push_back(steppe::QpGraphEdge{from[i], to[i]});
    return e;
}

static std::vector<std::string> qg_names(int n, const char* const* names) {
    std::vector<std::string> v;
    for (int i = 0; i < n; ++i) v.emplace_back(names[i]);
    return v;
}

// Same outputs as cref_qpgraph_fit (oracle/cpu_ref_qpgraph.cpp).
int ref_qpgraph_fit(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, int n_edges,
                    const char* const* from, const char* const* to, int n_names, const char* const* names,
                    double fudge, double diag_f3, int numstart, int constrained, int maxit, double tol,
                    double* scal, double* theta, double* lo, double* hi, double* bl, double* fit) {
    try {
        for (int i = 0; i < 5; ++i) scal[i] = 0.0;
        steppe::F2BlockTensor t;
        t.P = P;
        t.n_block = nb;
        const std::size_t tot = static_cast<std::size_t>(P) * P * nb;
        t.f2.assign(f2, f2 + tot);
        if (vpair) t.vpair.assign(vpair, vpair + tot);
        t.block_sizes.assign(bsz, bsz + nb);
        steppe::QpGraphOptions o;
        o.fudge = fudge;
        o.diag_f3 = diag_f3;
        o.numstart = numstart;
        o.constrained = constrained != 0;
        o.maxit = maxit;
        o.tol = tol;
        auto res = cpu_resources();
        const steppe::QpGraphResult r = steppe::run_qpgraph(t, qg_edges(n_edges, from, to), qg_names(n_names, names), o, res);
        scal[4] = static_cast<double>(static_cast<int>(r.status));
        scal[0] = r.score;
        scal[1] = r.restart_spread;
        scal[2] = r.worst_residual_z;
        put(theta, r.weight, r.weight.size());
        put(lo, r.weight_lo, r.weight_lo.size());
        put(hi, r.weight_hi, r.weight_hi.size());
        put(bl, r.edge_length, r.edge_length.size());
        (void)fit;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

double ref_qpgraph_score(int n_edges, const char* const* from, const char* const* to, int n_names,
                         const char* const* names, const double* theta, const double* f_obs, const double* qinv,
                         double fudge, int constrained, double* bl, double* fit) {
    const auto m = steppe::core::qpadm::parse_qpgraph(qg_edges(n_edges, from, to), qg_names(n_names, names));
    if (!m.ok()) return std::nan("");
    const std::vector<double> f(f_obs, f_obs + m.npair);
    const std::vector<double> qi(qinv, qinv + static_cast<std::size_t>(m.npair) * m.npair);
    std::vector<double> blv, fitv(static_cast<std::size_t>(m.npair), 0.0);
    const double s = steppe::core::qpadm::qpgraph_score(m, theta, f, qi, fudge, constrained != 0, &blv, &fitv);
    if (bl) put(bl, blv, blv.size());
    if (fit) put(fit, fitv, fitv.size());
    return s;
}

// run_f4 / run_f3 over N index tuples (4 or 3 ints each); est/se/z/p per tuple.
int ref_f4(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* quartets,
           int N, double* est, double* se, double* z, double* p) {
    try {
        const auto t = host_tensor(f2, vpair, bsz, P, nb);
        std::vector<std::array<int, 4>> q(N);
        for (int k = 0; k < N; ++k) q[k] = {quartets[4 * k], quartets[4 * k + 1], quartets[4 * k + 2], quartets[4 * k + 3]};
        auto res = cpu_resources();
        const steppe::F4Result r = steppe::run_f4(t, q, steppe::QpAdmOptions{}, res);
        put(est, r.est, N);
        put(se, r.se, N);
        put(z, r.z, N);
        put(p, r.p, N);
        return static_cast<int>(r.status);
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_f3(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, const std::int32_t* triples,
           int N, double* est, double* se, double* z, double* p) {
    try {
        const auto t = host_tensor(f2, vpair, bsz, P, nb);
        std::vector<std::array<int, 3>> q(N);
        for (int k = 0; k < N; ++k) q[k] = {triples[3 * k], triples[3 * k + 1], triples[3 * k + 2]};
        auto res = cpu_resources();
        const steppe::F3Result r = steppe::run_f3(t, q, steppe::QpAdmOptions{}, res);
        put(est, r.est, N);
        put(se, r.se, N);
        put(z, r.z, N);
        put(p, r.p, N);
        return static_cast<int>(r.status);
    } catch (const std::exception&) {
        return -1;
    }
}

int ref_f4ratio(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb,
                const std::int32_t* tuples, int N, double* alpha, double* se, double* z) {
    try {
        const auto t = host_tensor(f2, vpair, bsz, P, nb);
        std::vector<std::array<int, 5>> q(N);
        for (int k = 0; k < N; ++k)
            q[k] = {tuples[5 * k], tuples[5 * k + 1], tuples[5 * k + 2], tuples[5 * k + 3], tuples[5 * k + 4]};
        auto res = cpu_resources();
        const steppe::F4RatioResult r = steppe::run_f4ratio(t, q, steppe::QpAdmOptions{}, res);
        put(alpha, r.alpha, N);
        put(se, r.se, N);
        put(z, r.z, N);
        return static_cast<int>(r.status);
    } catch (const std::exception&) {
        return -1;
    }
}

namespace {
struct RefExtract {
    int P = 0, nb = 0;
    long long total = 0, kept = 0, n_ph = 0, n_dip = 0;
    std::vector<std::string> labels;
    steppe::F2BlockTensor t;
};
}  // namespace

void* ref_extract_open(const char* geno, const char* snp, const char* ind, int pop_mode, const char* const* pops,
                       int n_pops, long long top_k, long long min_n, double maf_min, double maxmiss,
                       int autosomes_only, int drop_monomorphic, int transversions_only, int strand_mode,
                       double blgsize, int ploidy, char* err, int err_len) {
    try {
        namespace io = steppe::io;
        namespace flt = steppe::io::filter;
        io::GenoReader reader(geno);
        io::PopSelection sel;
        sel.mode = pop_mode == 0 ? io::PopSelection::Mode::Explicit
                 : pop_mode == 1 ? io::PopSelection::Mode::AutoTopK : io::PopSelection::Mode::MinN;
        for (int i = 0; i < n_pops; ++i) sel.labels.emplace_back(pops[i]);
        sel.k = static_cast<std::size_t>(top_k);
        sel.min_n = static_cast<std::size_t>(min_n);
        const io::IndPartition part = io::read_ind_partition(reader.header().format, ind, sel, reader.records_present());
        if (pop_mode == 0)
            for (const auto& want : sel.labels) {
                bool found = false;
                for (const auto& g : part.groups) found = found || g.label == want;
                if (!found) throw std::invalid_argument("unknown population " + want);
            }
        const io::SnpTable snptab = io::read_snp_table(reader.header().format, snp, SIZE_MAX);
        const std::size_t M0 = std::min(reader.header().n_snp, snptab.count);
        auto be = steppe::device::make_cpu_backend();
        io::GenotypeTile tile;
        if (reader.header().format == io::GenoFormat::Tgeno) {
            tile = reader.read_tile(part, 0, M0);
        } else {
            const io::GenoFormat fmt = reader.header().format;
            const io::SnpMajorTile src = fmt == io::GenoFormat::Plink ? reader.read_plink_snp_major_tile(part, 0, M0)
                                       : fmt == io::GenoFormat::Eigenstrat
                                           ? reader.read_eigenstrat_snp_major_tile(part, 0, M0)
                                       : fmt == io::GenoFormat::Ancestrymap
                                           ? reader.read_ancestrymap_snp_major_tile(part, 0, M0)
                                           : reader.read_snp_major_tile(part, 0, M0);
            steppe::SnpMajorTileView v;
            v.snp_major = src.snp_major.data();
            v.src_bytes_per_record = src.src_bytes_per_record;
            v.n_snp = src.n_snp;
            v.sel_rows = src.sel_rows.data();
            v.n_individuals = src.n_individuals;
            v.pop_offsets = src.pop_offsets.data();
            v.n_pop = static_cast<int>(src.n_pop());
            steppe::CanonicalTile c = be->transpose_to_canonical(v);
            tile.packed = std::move(c.packed);
            tile.bytes_per_record = c.bytes_per_record;
            tile.n_snp = c.n_snp;
            tile.n_individuals = c.n_individuals;
            tile.pop_offsets = std::move(c.pop_offsets);
            tile.pop_labels = src.pop_labels;
        }
        const int P = static_cast<int>(tile.n_pop());
        const long M = static_cast<long>(tile.n_snp);
        steppe::DecodeTileView view;
        view.packed = tile.packed.data();
        view.bytes_per_record = tile.bytes_per_record;
        view.n_snp = tile.n_snp;
        view.n_individuals = tile.n_individuals;
        view.pop_offsets = tile.pop_offsets.data();
        view.n_pop = P;
        view.ploidy = 2;
        std::vector<int> sample_ploidy;
        if (ploidy == 0) {
            view.detect_ploidy_on_device = true;
            sample_ploidy = be->detect_sample_ploidy_device(view);
        } else {
            sample_ploidy.assign(tile.n_individuals, ploidy);
        }
        auto* out = new RefExtract;
        for (int pl : sample_ploidy) (pl == 1 ? out->n_ph : out->n_dip)++;
        view.sample_ploidy = sample_ploidy.data();
        view.detect_ploidy_on_device = false;
        const steppe::DecodeResult dec = be->decode_af(view);
        std::vector<std::size_t> pop_individuals(static_cast<std::size_t>(P));
        for (int p = 0; p < P; ++p) pop_individuals[p] = tile.pop_offsets[p + 1] - tile.pop_offsets[p];
        flt::DecodedTileSummaryInput fin;
        fin.q = dec.q.data();
        fin.v = dec.v.data();
        fin.n = dec.n.data();
        fin.P = P;
        fin.M = M;
        fin.pop_individuals = pop_individuals;
        fin.ploidy = 2;
        steppe::FilterConfig filter;
        filter.maf_min = maf_min;
        filter.geno_max_missing = maxmiss;
        filter.autosomes_only = autosomes_only != 0;
        filter.drop_monomorphic = drop_monomorphic != 0;
        filter.transversions_only = transversions_only != 0;
        filter.strand_mode = strand_mode == 0 ? steppe::StrandMode::Drop
                           : strand_mode == 1 ? steppe::StrandMode::Keep : steppe::StrandMode::Flip;
        steppe::FilterConfig class_filter = filter;
        class_filter.geno_max_missing = 1.0;
        flt::SnpMembership mem(class_filter);
        std::vector<bool> keep = flt::build_snp_keep_mask(fin, snptab, class_filter, mem);
        if (maxmiss < 1.0)
            for (long s = 0; s < M; ++s) {
                if (!keep[s]) continue;
                int n_missing = 0;
                for (int p = 0; p < P; ++p)
                    if (dec.n[static_cast<std::size_t>(p) + static_cast<std::size_t>(P) * s] <= 0.0) ++n_missing;
                if (static_cast<double>(n_missing) / static_cast<double>(P) > maxmiss) keep[s] = false;
            }
        std::vector<double> Qk, Vk, Nk, gk, pk;
        std::vector<int> ck;
        for (long s = 0; s < M; ++s) {
            if (!keep[s]) continue;
            for (int i = 0; i < P; ++i) {
                const std::size_t o = static_cast<std::size_t>(i) + static_cast<std::size_t>(P) * s;
                Qk.push_back(dec.q[o]);
                Vk.push_back(dec.v[o]);
                Nk.push_back(dec.n[o]);
            }
            ck.push_back(snptab.chrom[s]);
            gk.push_back(snptab.genpos_morgans[s]);
            pk.push_back(snptab.physpos[s]);
        }
        const long Mk = static_cast<long>(ck.size());
        if (Mk <= 0) throw std::invalid_argument("every SNP was filtered out");
        const auto part_b = steppe::core::assign_blocks(ck, gk, blgsize, pk);
        if (part_b.n_block <= 0) throw std::invalid_argument("no blocks");
        out->t = be->compute_f2_blocks(steppe::core::MatView{Qk.data(), P, Mk}, steppe::core::MatView{Vk.data(), P, Mk},
                                       steppe::core::MatView{Nk.data(), P, Mk}, part_b.block_id.data(),
                                       part_b.n_block, steppe::Precision{});
        out->P = P;
        out->nb = part_b.n_block;
        out->total = M;
        out->kept = Mk;
        out->labels = tile.pop_labels;
        return out;
    } catch (const std::exception& e) {
        if (err && err_len > 0) {
            std::strncpy(err, e.what(), static_cast<std::size_t>(err_len - 1));
            err[err_len - 1] = 0;
        }
        return nullptr;
    }
}

int ref_extract_dims(void* h, int* P, int* nb, long long* info) {
    const auto* x = static_cast<RefExtract*>(h);
    *P = x->P;
    *nb = x->nb;
    info[0] = x->total;
    info[1] = x->kept;
    info[2] = x->n_ph;
    info[3] = x->n_dip;
    return 0;
}

int ref_extract_fetch(void* h, double* f2, double* vpair, std::int32_t* bsz) {
    const auto* x = static_cast<RefExtract*>(h);
    std::memcpy(f2, x->t.f2.data(), x->t.f2.size() * sizeof(double));
    std::memcpy(vpair, x->t.vpair.data(), x->t.vpair.size() * sizeof(double));
    for (int b = 0; b < x->nb; ++b) bsz[b] = x->t.block_sizes[static_cast<std::size_t>(b)];
    return 0;
}

// ---- qpGraph topology enumeration and search --------------------------------------
struct RefTopoList {
    std::vector<int> nadmix, id;
    std::vector<unsigned long long> hash;
    std::vector<std::string> edges;          // "from>to;from>to;..."
    std::vector<double> score, spread;       // search candidates only
};

static std::string join_edges(const std::vector<steppe::QpGraphEdge>& e) {
    std::string s;
    for (std::size_t k = 0; k < e.size(); ++k) {
        if (k) s += ';';
        s += e[k].from + ">" + e[k].to;
    }
    return s;
}

static std::vector<steppe::QpGraphEdge> split_edges(const char* str) {
    std::vector<steppe::QpGraphEdge> out;
    std::string s(str);
    std::size_t at = 0;
    while (at < s.size()) {
        std::size_t end = s.find(';', at);
        if (end == std::string::npos) end = s.size();
        const std::string tok = s.substr(at, end - at);
        const std::size_t gt = tok.find('>');
        if (gt != std::string::npos) out.push_back({tok.substr(0, gt), tok.substr(gt + 1)});
        at = end + 1;
    }
    return out;
}

static RefTopoList* to_list(const std::vector<steppe::core::qpadm::EnumeratedTopology>& v) {
    auto* L = new RefTopoList;
    for (const auto& t : v) {
        L->nadmix.push_back(t.nadmix);
        L->id.push_back(t.id);
        L->hash.push_back(t.hash);
        L->edges.push_back(join_edges(t.edges));
    }
    return L;
}

void* ref_qg_enumerate(int n, const char* const* leaves, int max_nadmix) {
    return to_list(steppe::core::qpadm::enumerate_bounded_space(qg_names(n, leaves), max_nadmix));
}

void* ref_qg_neighbors(const char* edges, int nadmix, unsigned long long hash, int n, const char* const* leaves,
                       int max_nadmix) {
    steppe::core::qpadm::EnumeratedTopology cur;
    cur.edges = split_edges(edges);
    cur.nadmix = nadmix;
    cur.hash = hash;
    return to_list(steppe::core::qpadm::topology_neighbors(cur, qg_names(n, leaves), max_nadmix));
}

unsigned long long ref_qg_graph_hash(const char* edges) {
    return steppe::core::qpadm::graph_hash(split_edges(edges));
}

// scal: best_score, second_best_score, status, heuristic_recovered, best_nadmix,
//       n_trees, n_admix1, n_candidates, fit_all_wall_ms; best_hash[0]
void* ref_qg_search(const double* f2, const double* vpair, const std::int32_t* bsz, int P, int nb, int n_names,
                    const char* const* names, int n_pops, const char* const* pops, int max_nadmix, int numstart,
                    double fudge, double diag_f3, int constrained, int run_heuristic, double* scal,
                    unsigned long long* best_hash) {
    try {
        steppe::F2BlockTensor t = host_tensor(f2, vpair, bsz, P, nb);
        steppe::QpGraphSearchOptions o;
        o.pops = qg_names(n_pops, pops);
        o.max_nadmix = max_nadmix;
        o.fit.numstart = numstart;
        o.fit.fudge = fudge;
        o.fit.diag_f3 = diag_f3;
        o.fit.constrained = constrained != 0;
        o.run_heuristic = run_heuristic != 0;
        auto res = cpu_resources();
        const steppe::QpGraphSearchResult r = steppe::run_qpgraph_search(t, qg_names(n_names, names), o, res);
        scal[0] = r.best.score;
        scal[1] = r.second_best_score;
        scal[2] = static_cast<double>(static_cast<int>(r.status));
        scal[3] = r.heuristic_recovered ? 1.0 : 0.0;
        scal[4] = r.best.nadmix;
        scal[5] = r.n_trees;
        scal[6] = r.n_admix1;
        scal[7] = r.n_candidates;
        scal[8] = r.fit_all_wall_ms;
        best_hash[0] = r.best.hash;
        auto* L = new RefTopoList;
        for (const auto& c : r.candidates) {
            L->nadmix.push_back(c.nadmix);
            L->id.push_back(c.id);
            L->hash.push_back(c.hash);
            L->edges.push_back(join_edges(c.edges));
            L->score.push_back(c.score);
            L->spread.push_back(c.restart_spread);
        }
        return L;
    } catch (const std::exception&) {
        return nullptr;
    }
}

int ref_qg_list_size(void* h) { return static_cast<int>(static_cast<RefTopoList*>(h)->hash.size()); }

int ref_qg_list_get(void* h, int i, int* nadmix, int* id, unsigned long long* hash, double* score, double* spread) {
    const auto* L = static_cast<RefTopoList*>(h);
    *nadmix = L->nadmix[i];
    *id = L->id[i];
    *hash = L->hash[i];
    *score = L->score.empty() ? 0.0 : L->score[i];
    *spread = L->spread.empty() ? 0.0 : L->spread[i];
    return 0;
}

const char* ref_qg_list_edges(void* h, int i) { return static_cast<RefTopoList*>(h)->edges[i].c_str(); }

void ref_qg_list_free(void* h) { delete static_cast<RefTopoList*>(h); }

const char* ref_extract_label(void* h, int i) { return static_cast<RefExtract*>(h)->labels[i].c_str(); }

void ref_extract_free(void* h) { delete static_cast<RefExtract*>(h); }

}  // extern "C"
