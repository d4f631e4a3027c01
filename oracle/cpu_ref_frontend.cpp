// oracle/cpu_ref_frontend.cpp — CPU restatement of the reference's genotype-file extract:
// packed .geno (TGENO / GENO) or EIGENSTRAT text .geno + .snp + .ind, or PLINK .bed + .bim +
// .fam -> f2 block tensor.
//
// TEST INFRASTRUCTURE ONLY: the parity tests call it as the checker; the product
// library never links or calls it.
//
// Restated from (paths relative to the reference repository root):
//   .geno header            src/io/eigenstrat_format.cpp:17-88 (magic, n_ind, n_snp, strides)
//   records present         src/io/geno_reader.cpp:195-271
//   .ind + selection        src/io/ind_reader.cpp:13-50, src/io/detail/pop_select.hpp:25-76
//   .snp                    src/io/snp_reader.cpp:20-63, src/io/detail/snp_text_parse.hpp
//   TGENO gather            src/io/geno_reader.cpp:274-372
//   GENO transpose          src/io/geno_reader.cpp:375-505 + src/device/backend.hpp:790-830
//                           (ComputeBackend::transpose_to_canonical)
//   EIGENSTRAT text .geno   src/io/geno_reader.cpp:83-120 (geometry), 516-583 (0/1/2/9 per char)
//   ANCESTRYMAP .geno       src/io/geno_reader.cpp:122-177, 648-744 (<snp> <ind> <0/1/2/-1> lines,
//                           by position; geometry from the .snp / .ind beside the .geno)
//   PLINK .bed              src/io/geno_reader.cpp:48-72, 208-238 (magic 6c 1b 01, geometry from
//                           the .bim / .fam beside the .bed), 585-634 (LSB-first codes,
//                           kBedToCanon = {2, 3, 1, 0}, src/io/eigenstrat_format.hpp:66-77)
//   .bim / .fam             src/io/plink_reader.cpp:53-151 (A1 is the counted allele, '0' -> 'N';
//                           the population is the .fam phenotype column, 1 Control, 2 Case,
//                           9 / -9 / 0 skipped), src/io/genotype_source.cpp:33-44
//   front end               src/core/stats/genotype_front_end.cpp:9-37 (M0 = min(n_snp, rows))
//   decode / ploidy         cpu_ref.cpp (cref_decode_af, cref_detect_ploidy)
//   keep mask               src/io/filter/snp_filter.cpp:11-117, snp_summary_reduce.hpp:36-77,
//                           filter_decision.hpp:17-91, with the population-axis test of
//                           src/app/extract_f2_core.cpp:217-232 (the reference's CPU branch)
//   blocks / f2             cpu_ref.cpp (cref_assign_blocks, cref_f2_blocks)
#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

extern "C" {
int cref_detect_ploidy(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp, std::size_t n_ind,
                       std::int32_t* out);
int cref_decode_af(const std::uint8_t* packed, std::size_t bpr, std::size_t n_snp, std::size_t n_ind,
                   const std::uint64_t* pop_offsets, int n_pop, const std::int32_t* sample_ploidy, int ploidy,
                   int detect, double* out_q, double* out_v, double* out_n);
int cref_assign_blocks(const std::int32_t* chrom, const double* genpos, std::int64_t M, double blgsize,
                       const double* physpos, double bp_window, std::int32_t* out_block_id);
int cref_f2_blocks(const double* Q, const double* V, const double* N, int P, std::int64_t M,
                   const std::int32_t* block_id, int nb, double* f2, double* vpair, std::int32_t* block_sizes,
                   int n_threads);
}

namespace {

struct Extract {
    int P = 0, nb = 0;
    long long total = 0, kept = 0, n_ph = 0, n_dip = 0;
    std::vector<std::string> labels;
    std::vector<double> f2, vpair;
    std::vector<std::int32_t> bsz;
};

std::vector<std::string> split(const std::string& line) {
    std::vector<std::string> t;
    std::istringstream ls(line);
    std::string w;
    while (ls >> w) t.push_back(w);
    return t;
}

template <class T>
bool full(const std::string& s, T& v) {
    const auto [p, ec] = std::from_chars(s.data(), s.data() + s.size(), v);
    return ec == std::errc{} && p == s.data() + s.size();
}

char upper_acgt(char a) {
    const char u = (a >= 'a' && a <= 'z') ? static_cast<char>(a - 32) : a;
    return (u == 'A' || u == 'C' || u == 'G' || u == 'T') ? u : 0;
}

bool keep_snp(double ref_af, double minor, double missing, char r, char a, int chrom, double maf_min,
              bool strand_drop, bool drop_mono, bool tv_only, bool auto_only) {
    const char x = upper_acgt(r), y = upper_acgt(a);
    if (x == 0 || y == 0 || x == y) return false;                                   // multiallelic
    auto comp = [](char c) { return c == 'A' ? 'T' : c == 'T' ? 'A' : c == 'C' ? 'G' : 'C'; };
    if (strand_drop && x == comp(y)) return false;                                  // A/T, C/G
    if (!(minor >= maf_min)) return false;
    if (!(missing <= 1.0)) return false;                                            // lifted to 1
    if (drop_mono && std::min(ref_af, 1.0 - ref_af) == 0.0) return false;
    if (tv_only) {
        const bool px = x == 'A' || x == 'G', py = y == 'A' || y == 'G';
        if (px == py) return false;
    }
    if (auto_only && !(chrom >= 1 && chrom <= 22)) return false;
    return true;
}

Extract* run(const char* geno, const char* snp, const char* ind, int pop_mode, const char* const* pops, int n_pops,
             long long top_k, long long min_n, double maf_min, double maxmiss, int autosomes_only,
             int drop_monomorphic, int transversions_only, int strand_mode, double blgsize, int ploidy) {
    // ---- .geno header ----
    std::ifstream g(geno, std::ios::binary);
    if (!g) throw std::runtime_error("cannot open .geno");
    char head[48] = {};
    g.read(head, 48);
    std::string text;
    for (int i = 0; i < 48 && head[i]; ++i) text.push_back(head[i]);
    const auto tok = split(text);
    const bool packed_hdr = !tok.empty() && (tok[0] == "TGENO" || tok[0] == "GENO");
    const bool plink = !packed_hdr && g.gcount() >= 3 && static_cast<unsigned char>(head[0]) == 0x6c &&
                       static_cast<unsigned char>(head[1]) == 0x1b;
    // EIGENSTRAT text: equal-length lines of 0 / 1 / 2 / 9, an empty line only at the end
    std::vector<std::string> text_lines;
    bool ascii = false;
    if (!packed_hdr && !plink) {
        std::ifstream t(geno, std::ios::binary);
        std::string l;
        bool ok = true;
        while (ok && std::getline(t, l)) {
            if (!l.empty() && l.back() == '\r') l.pop_back();
            if (l.empty()) {
                if (t.peek() == std::char_traits<char>::eof()) break;
                ok = false;
                break;
            }
            if (!text_lines.empty() && l.size() != text_lines[0].size()) ok = false;
            for (char c : l)
                if (!((c >= '0' && c <= '2') || c == '9')) ok = false;
            text_lines.push_back(l);
        }
        ascii = ok && !text_lines.empty() && !text_lines[0].empty();
    }
    // ANCESTRYMAP: the first non-blank line is "<snp> <ind> <genotype>"
    bool amap = false;
    std::string amap_base;
    if (!packed_hdr && !plink && !ascii) {
        std::ifstream t(geno, std::ios::binary);
        std::string l;
        std::vector<std::string> w;
        while (std::getline(t, l)) {
            w = split(l);
            if (!w.empty()) break;
        }
        const bool tok_ok = w.size() == 3 && ((w[2].size() == 1 && w[2][0] >= '0' && w[2][0] <= '2') || w[2] == "-1");
        amap_base = geno;
        if (amap_base.size() >= 5 && amap_base.compare(amap_base.size() - 5, 5, ".geno") == 0)
            amap_base.resize(amap_base.size() - 5);
        amap = tok_ok;
    }
    if (!packed_hdr && !plink && !ascii && !amap)
        throw std::runtime_error("not TGENO / GENO / EIGENSTRAT / PLINK .bed / ANCESTRYMAP");
    if (plink && static_cast<unsigned char>(head[2]) != 0x01) throw std::runtime_error("individual-major .bed");
    const bool tgeno = packed_hdr && tok[0] == "TGENO";
    std::size_t counts[2] = {0, 0};
    if (plink) {
        // geometry from the non-empty lines of the .bim and .fam next to the .bed
        std::string base(geno);
        if (base.size() >= 4 && base.compare(base.size() - 4, 4, ".bed") == 0) base.resize(base.size() - 4);
        auto lines = [](const std::string& p) {
            std::ifstream in(p, std::ios::binary);
            std::size_t n = 0;
            std::string l;
            while (std::getline(in, l)) {
                if (!l.empty() && l.back() == '\r') l.pop_back();
                if (l.empty()) {
                    if (in.peek() == std::char_traits<char>::eof()) break;
                    continue;
                }
                ++n;
            }
            return n;
        };
        counts[1] = lines(base + ".bim");
        counts[0] = lines(base + ".fam");
        if (counts[0] == 0 || counts[1] == 0) throw std::runtime_error(".bim / .fam missing or empty");
    } else if (ascii) {
        counts[0] = text_lines[0].size();
        counts[1] = text_lines.size();
    } else if (amap) {
        auto lines = [](const std::string& p) {
            std::ifstream in(p, std::ios::binary);
            std::size_t n = 0;
            std::string l;
            while (std::getline(in, l)) {
                if (!l.empty() && l.back() == '\r') l.pop_back();
                if (!l.empty()) ++n;
            }
            return n;
        };
        counts[1] = lines(amap_base + ".snp");
        counts[0] = lines(amap_base + ".ind");
        if (counts[0] == 0 || counts[1] == 0) throw std::runtime_error("ANCESTRYMAP .snp / .ind missing or empty");
    } else {
        int got = 0;
        std::size_t i = text.find(tok[0]) + tok[0].size();
        while (i < text.size() && got < 2) {
            while (i < text.size() && !std::isdigit(static_cast<unsigned char>(text[i]))) ++i;
            if (i >= text.size()) break;
            std::size_t v = 0;
            while (i < text.size() && std::isdigit(static_cast<unsigned char>(text[i]))) v = v * 10 + (text[i++] - '0');
            counts[got++] = v;
        }
        if (got < 2) throw std::runtime_error("header counts");
    }
    const std::size_t n_ind = counts[0], n_snp = counts[1];
    const std::size_t bpr = plink ? (n_ind + 3) / 4 : tgeno ? (n_snp + 3) / 4 : std::max<std::size_t>(48, (n_ind + 3) / 4);
    const std::size_t hdr = plink ? 3 : tgeno ? 48 : bpr;
    const std::size_t n_records = tgeno ? n_ind : n_snp;
    g.seekg(0, std::ios::end);
    const std::size_t fsize = static_cast<std::size_t>(g.tellg());
    const std::size_t present = (ascii || amap) ? n_records : std::min((fsize - hdr) / bpr, n_records);

    // ---- .ind ----
    std::ifstream fi(ind);
    if (!fi) throw std::runtime_error("cannot open .ind");
    struct Grp {
        std::string label;
        std::size_t first;
        std::vector<std::size_t> rows;
    };
    std::vector<Grp> groups;
    std::map<std::string, std::size_t> at;
    std::size_t row = 0;
    std::string line;
    while (std::getline(fi, line)) {
        const auto t = split(line);
        std::string label;
        if (plink) {
            // <fid> <iid> <father> <mother> <sex> <phenotype>: the phenotype names the group
            if (t.size() < 6) continue;
            if (row >= present) {
                ++row;
                continue;
            }
            label = t[5];
            if (label == "9" || label == "-9" || label == "0") {
                ++row;
                continue;
            }
            if (label == "1") label = "Control";
            else if (label == "2") label = "Case";
        } else {
            if (t.size() < 3 || row >= present) continue;
            label = t[2];
        }
        auto it = at.find(label);
        if (it == at.end()) {
            at[label] = groups.size();
            groups.push_back({label, groups.size(), {row}});
        } else {
            groups[it->second].rows.push_back(row);
        }
        ++row;
    }
    std::vector<const Grp*> pick;
    if (pop_mode == 0) {
        std::unordered_set<std::string> want;
        for (int i = 0; i < n_pops; ++i) want.insert(pops[i]);
        for (const auto& gr : groups)
            if (want.count(gr.label)) pick.push_back(&gr);
        for (int i = 0; i < n_pops; ++i)
            if (!at.count(pops[i])) throw std::runtime_error("unknown population");
    } else if (pop_mode == 1) {
        for (const auto& gr : groups) pick.push_back(&gr);
        std::stable_sort(pick.begin(), pick.end(), [](const Grp* a, const Grp* b) {
            return a->rows.size() != b->rows.size() ? a->rows.size() > b->rows.size() : a->first < b->first;
        });
        pick.resize(std::min<std::size_t>(static_cast<std::size_t>(top_k), pick.size()));
    } else {
        for (const auto& gr : groups)
            if (gr.rows.size() >= static_cast<std::size_t>(min_n)) pick.push_back(&gr);
    }
    if (pick.empty()) throw std::runtime_error("empty selection");
    std::sort(pick.begin(), pick.end(), [](const Grp* a, const Grp* b) { return a->label < b->label; });

    // ---- .snp ----
    std::ifstream fs(snp);
    if (!fs) throw std::runtime_error("cannot open .snp");
    std::vector<std::int32_t> chrom;
    std::vector<double> gpos, ppos;
    std::vector<char> ra, aa;
    std::map<std::string, int> other;
    int next_other = -1;
    while (std::getline(fs, line)) {
        auto f = split(line);
        if (f.empty()) break;
        if (plink) {
            // <chrom> <id> <genpos> <physpos> <A1> <A2>; A1 (the counted allele) is ref
            if (f.size() < 6) throw std::runtime_error("bad .bim line");
            for (int k : {4, 5})
                if (f[k][0] == '0') f[k] = "N";
            f = {f[1], f[0], f[2], f[3], f[4], f[5]};
        } else if (f.size() < 3) {
            throw std::runtime_error("bad .snp line");
        }
        int c = 0;
        const bool digits = std::all_of(f[1].begin(), f[1].end(), [](char ch) { return std::isdigit((unsigned char)ch); });
        if (!(digits && full(f[1], c))) {
            if (f[1] == "X" || f[1] == "x") c = 23;
            else if (f[1] == "Y" || f[1] == "y") c = 24;
            else if (f[1] == "MT" || f[1] == "mt" || f[1] == "M") c = 90;
            else if (other.count(f[1])) c = other[f[1]];
            else c = other[f[1]] = next_other--;
        }
        double gp = 0.0, pp = 0.0;
        if (!full(f[2], gp) || !std::isfinite(gp)) throw std::runtime_error("bad genpos");
        if (f.size() > 3 && !(full(f[3], pp) && std::isfinite(pp))) pp = 0.0;
        chrom.push_back(c);
        gpos.push_back(gp);
        ppos.push_back(pp);
        ra.push_back(f.size() >= 6 ? f[4][0] : 'N');
        aa.push_back(f.size() >= 6 ? f[5][0] : 'N');
    }
    const std::size_t M0 = std::min(n_snp, chrom.size());
    if (M0 == 0) throw std::runtime_error("no SNPs");
    if (!tgeno && M0 > present) throw std::runtime_error("fewer SNP records on disk than requested");

    // ---- canonical individual-major tile ----
    std::vector<std::size_t> sel;
    std::vector<std::uint64_t> offs{0};
    for (const Grp* gr : pick) {
        for (std::size_t r : gr->rows) {
            if (!tgeno && r >= n_ind) throw std::runtime_error("individual row beyond the header");
            sel.push_back(r);
        }
        offs.push_back(sel.size());
    }
    const std::size_t cbpr = (M0 + 3) / 4;
    std::vector<std::uint8_t> tile(sel.size() * cbpr, 0);
    if (tgeno) {
        for (std::size_t k = 0; k < sel.size(); ++k) {
            g.seekg(static_cast<std::streamoff>(hdr + sel[k] * bpr));
            g.read(reinterpret_cast<char*>(tile.data() + k * cbpr), static_cast<std::streamsize>(cbpr));
        }
    } else if (amap) {
        std::ifstream t(geno, std::ios::binary);
        std::string l;
        std::vector<std::uint8_t> rec(n_ind);
        for (std::size_t s = 0; s < M0; ++s) {
            for (std::size_t i = 0; i < n_ind; ++i) {
                if (!std::getline(t, l)) throw std::runtime_error("ANCESTRYMAP: unexpected end of file");
                const auto w = split(l);
                if (w.size() != 3) throw std::runtime_error("ANCESTRYMAP: malformed line");
                if (w[2] == "-1") rec[i] = 3;
                else if (w[2].size() == 1 && w[2][0] >= '0' && w[2][0] <= '2') rec[i] = static_cast<std::uint8_t>(w[2][0] - '0');
                else throw std::runtime_error("ANCESTRYMAP: illegal genotype");
            }
            for (std::size_t k = 0; k < sel.size(); ++k)
                tile[k * cbpr + s / 4] |= static_cast<std::uint8_t>(rec[sel[k]] << (6 - 2 * (s % 4)));
        }
    } else if (ascii) {
        for (std::size_t s = 0; s < M0; ++s)
            for (std::size_t k = 0; k < sel.size(); ++k) {
                const char ch = text_lines[s][sel[k]];
                const std::uint8_t code = ch == '9' ? 3 : static_cast<std::uint8_t>(ch - '0');
                tile[k * cbpr + s / 4] |= static_cast<std::uint8_t>(code << (6 - 2 * (s % 4)));
            }
    } else {
        std::vector<std::uint8_t> rec(bpr);
        for (std::size_t s = 0; s < M0; ++s) {
            g.seekg(static_cast<std::streamoff>(hdr + s * bpr));
            g.read(reinterpret_cast<char*>(rec.data()), static_cast<std::streamsize>(bpr));
            for (std::size_t k = 0; k < sel.size(); ++k) {
                std::uint8_t code;
                if (plink) {
                    static const std::uint8_t bed_to_canon[4] = {2, 3, 1, 0};
                    code = bed_to_canon[(rec[sel[k] / 4] >> (2 * (sel[k] % 4))) & 3];
                } else {
                    code = (rec[sel[k] / 4] >> (6 - 2 * (sel[k] % 4))) & 3;
                }
                tile[k * cbpr + s / 4] |= static_cast<std::uint8_t>(code << (6 - 2 * (s % 4)));
            }
        }
    }

    // ---- ploidy, decode ----
    const std::size_t N = sel.size();
    const int P = static_cast<int>(pick.size());
    std::vector<std::int32_t> pl(N, ploidy == 0 ? 2 : ploidy);
    if (ploidy == 0) cref_detect_ploidy(tile.data(), cbpr, M0, N, pl.data());
    auto* out = new Extract;
    for (auto v : pl) (v == 1 ? out->n_ph : out->n_dip)++;
    std::vector<double> Q(static_cast<std::size_t>(P) * M0), V(Q.size()), Nn(Q.size());
    cref_decode_af(tile.data(), cbpr, M0, N, offs.data(), P, pl.data(), 2, 0, Q.data(), V.data(), Nn.data());

    // ---- keep mask (pooled summary in population order), population-axis missingness ----
    std::vector<std::size_t> keep;
    for (std::size_t s = 0; s < M0; ++s) {
        double rc = 0.0, ac = 0.0, nm = 0.0;
        int missing_pops = 0;
        for (int p = 0; p < P; ++p) {
            const std::size_t o = static_cast<std::size_t>(p) + static_cast<std::size_t>(P) * s;
            const double prod = Q[o] * Nn[o];
            rc = rc + prod;
            ac += Nn[o];
            nm += Nn[o] / 2.0;
            missing_pops += Nn[o] <= 0.0;
        }
        const double raf = ac > 0.0 ? rc / ac : 0.0;
        const double minor = raf < 1.0 - raf ? raf : 1.0 - raf;
        const double miss = N > 0 ? 1.0 - nm / static_cast<double>(N) : 1.0;
        bool k = keep_snp(raf, minor, miss, ra[s], aa[s], chrom[s], maf_min, strand_mode == 0, drop_monomorphic != 0,
                          transversions_only != 0, autosomes_only != 0);
        if (k && maxmiss < 1.0 && static_cast<double>(missing_pops) / static_cast<double>(P) > maxmiss) k = false;
        if (k) keep.push_back(s);
    }
    if (keep.empty()) throw std::runtime_error("every SNP filtered out");
    const std::size_t Mk = keep.size();
    std::vector<double> Qk(static_cast<std::size_t>(P) * Mk), Vk(Qk.size()), Nk(Qk.size());
    std::vector<std::int32_t> ck(Mk);
    std::vector<double> gk(Mk), pk(Mk);
    for (std::size_t d = 0; d < Mk; ++d) {
        const std::size_t s = keep[d];
        for (int p = 0; p < P; ++p) {
            Qk[p + P * d] = Q[p + P * s];
            Vk[p + P * d] = V[p + P * s];
            Nk[p + P * d] = Nn[p + P * s];
        }
        ck[d] = chrom[s];
        gk[d] = gpos[s];
        pk[d] = ppos[s];
    }
    std::vector<std::int32_t> bid(Mk);
    const int nb = cref_assign_blocks(ck.data(), gk.data(), static_cast<std::int64_t>(Mk), blgsize, pk.data(), 2.0e6,
                                      bid.data());
    if (nb <= 0) throw std::runtime_error("no blocks");
    out->P = P;
    out->nb = nb;
    out->total = static_cast<long long>(M0);
    out->kept = static_cast<long long>(Mk);
    for (const Grp* gr : pick) out->labels.push_back(gr->label);
    const std::size_t tot = static_cast<std::size_t>(P) * P * nb;
    out->f2.assign(tot, 0.0);
    out->vpair.assign(tot, 0.0);
    out->bsz.assign(nb, 0);
    cref_f2_blocks(Qk.data(), Vk.data(), Nk.data(), P, static_cast<std::int64_t>(Mk), bid.data(), nb, out->f2.data(),
                   out->vpair.data(), out->bsz.data(), 8);
    return out;
}

}  // namespace

extern "C" {

void* cref_extract_open(const char* geno, const char* snp, const char* ind, int pop_mode, const char* const* pops,
                        int n_pops, long long top_k, long long min_n, double maf_min, double maxmiss,
                        int autosomes_only, int drop_monomorphic, int transversions_only, int strand_mode,
                        double blgsize, int ploidy, char* err, int err_len) {
    try {
        return run(geno, snp, ind, pop_mode, pops, n_pops, top_k, min_n, maf_min, maxmiss, autosomes_only,
                   drop_monomorphic, transversions_only, strand_mode, blgsize, ploidy);
    } catch (const std::exception& e) {
        if (err && err_len > 0) {
            std::strncpy(err, e.what(), static_cast<std::size_t>(err_len - 1));
            err[err_len - 1] = 0;
        }
        return nullptr;
    }
}

int cref_extract_dims(void* h, int* P, int* nb, long long* info) {
    const auto* x = static_cast<Extract*>(h);
    *P = x->P;
    *nb = x->nb;
    info[0] = x->total;
    info[1] = x->kept;
    info[2] = x->n_ph;
    info[3] = x->n_dip;
    return 0;
}

int cref_extract_fetch(void* h, double* f2, double* vpair, std::int32_t* bsz) {
    const auto* x = static_cast<Extract*>(h);
    std::copy(x->f2.begin(), x->f2.end(), f2);
    std::copy(x->vpair.begin(), x->vpair.end(), vpair);
    std::copy(x->bsz.begin(), x->bsz.end(), bsz);
    return 0;
}

const char* cref_extract_label(void* h, int i) { return static_cast<Extract*>(h)->labels[i].c_str(); }

void cref_extract_free(void* h) { delete static_cast<Extract*>(h); }

}  // extern "C"
