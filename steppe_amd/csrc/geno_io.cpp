// steppe_amd/csrc/geno_io.cpp — host readers of the packed EIGENSTRAT triple (see geno_io.hpp).
#include "geno_io.hpp"

#include <algorithm>
#include <cctype>
#include <charconv>
#include <cmath>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <unordered_map>
#include <unordered_set>

#include "common.hpp"

namespace stp {

namespace {

constexpr std::size_t kHeaderBytes = 48;

std::size_t packed_bytes(std::size_t n) { return (n + 3) / 4; }

[[noreturn]] void io_fail(const std::string& msg) { throw ArgError(msg); }

template <class T>
bool parse_all(const std::string& tok, T& out) {
    const char* b = tok.data();
    const char* e = b + tok.size();
    const auto [p, ec] = std::from_chars(b, e, out);
    return ec == std::errc{} && p == e;
}

std::vector<std::string> words(const std::string& line) {
    std::vector<std::string> t;
    std::istringstream ls(line);
    std::string w;
    while (ls >> w) t.push_back(w);
    return t;
}

constexpr unsigned char kBedMagic0 = 0x6c, kBedMagic1 = 0x1b, kBedSnpMajor = 0x01;

// The .bim / .fam beside a .bed: "x.bed" -> "x.bim", any other name -> name + ext.
std::string plink_sibling(const std::string& bed, const char* ext) {
    if (bed.size() >= 4 && bed.compare(bed.size() - 4, 4, ".bed") == 0) return bed.substr(0, bed.size() - 4) + ext;
    return bed + ext;
}

// Non-empty lines (a trailing CR removed); 0 when the file cannot be opened.
std::size_t count_lines(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return 0;
    std::size_t n = 0;
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty()) ++n;
    }
    return n;
}

bool genotype_char(char c, std::uint8_t& code) {
    if (c >= '0' && c <= '2') {
        code = static_cast<std::uint8_t>(c - '0');
        return true;
    }
    if (c == '9') {
        code = 3;
        return true;
    }
    return false;
}

// EIGENSTRAT text geometry; Unknown unless every line has the first line's length and
// only genotype characters (an empty line ends the file only at its end).
GenoHeader eigenstrat_geometry(const std::string& path) {
    GenoHeader h;
    std::ifstream in(path, std::ios::binary);
    if (!in) return h;
    std::string line;
    std::size_t n_ind = 0, n_snp = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) {
            if (in.peek() == std::char_traits<char>::eof()) break;
            return GenoHeader{};
        }
        if (n_snp == 0) n_ind = line.size();
        else if (line.size() != n_ind) return GenoHeader{};
        std::uint8_t code;
        for (char c : line)
            if (!genotype_char(c, code)) return GenoHeader{};
        ++n_snp;
    }
    if (n_snp == 0 || n_ind == 0) return GenoHeader{};
    h.format = GenoFormat::Eigenstrat;
    h.n_ind = n_ind;
    h.n_snp = n_snp;
    h.n_records = n_snp;
    h.bytes_per_record = packed_bytes(n_ind);
    h.header_bytes = 0;
    return h;
}

bool ancestrymap_code(const std::string& tok, std::uint8_t& code) {
    if (tok.size() == 1 && tok[0] >= '0' && tok[0] <= '2') {
        code = static_cast<std::uint8_t>(tok[0] - '0');
        return true;
    }
    if (tok == "-1") {
        code = 3;
        return true;
    }
    return false;
}

// ANCESTRYMAP: the first non-blank line has three tokens ending in a genotype, and the
// .snp / .ind beside the .geno give the geometry.
GenoHeader ancestrymap_geometry(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return GenoHeader{};
    std::string line;
    std::vector<std::string> t;
    while (std::getline(in, line)) {
        t = words(line);
        if (!t.empty()) break;
    }
    std::uint8_t code;
    if (t.size() != 3 || !ancestrymap_code(t[2], code)) return GenoHeader{};
    std::string base = path;
    if (base.size() >= 5 && base.compare(base.size() - 5, 5, ".geno") == 0) base.resize(base.size() - 5);
    const std::size_t n_snp = count_lines(base + ".snp"), n_ind = count_lines(base + ".ind");
    if (n_snp == 0 || n_ind == 0) return GenoHeader{};
    GenoHeader h;
    h.format = GenoFormat::Ancestrymap;
    h.n_ind = n_ind;
    h.n_snp = n_snp;
    h.n_records = n_snp;
    h.bytes_per_record = packed_bytes(n_ind);
    h.header_bytes = 0;
    return h;
}

struct RawGroup {
    std::string label;
    std::size_t first;
    std::vector<std::size_t> rows;
};

// The population selection over groups in order of first appearance, sorted by label.
std::vector<PopGroup> select_groups(const std::vector<RawGroup>& groups, const PopSelect& sel, const std::string& path) {
    std::vector<const RawGroup*> chosen;
    if (sel.mode == PopSelect::kExplicit) {
        const std::unordered_set<std::string> want(sel.labels.begin(), sel.labels.end());
        for (const auto& g : groups)
            if (want.count(g.label)) chosen.push_back(&g);
    } else if (sel.mode == PopSelect::kTopK) {
        for (const auto& g : groups) chosen.push_back(&g);
        std::stable_sort(chosen.begin(), chosen.end(), [](const RawGroup* a, const RawGroup* b) {
            if (a->rows.size() != b->rows.size()) return a->rows.size() > b->rows.size();
            return a->first < b->first;
        });
        chosen.resize(std::min(sel.k, chosen.size()));
    } else {
        for (const auto& g : groups)
            if (g.rows.size() >= sel.min_n) chosen.push_back(&g);
    }
    if (chosen.empty()) io_fail("ind: population selection is empty for " + path);
    std::sort(chosen.begin(), chosen.end(), [](const RawGroup* a, const RawGroup* b) { return a->label < b->label; });
    std::vector<PopGroup> out;
    out.reserve(chosen.size());
    for (const RawGroup* g : chosen) out.push_back(PopGroup{g->label, g->rows});
    return out;
}

void add_row(std::vector<RawGroup>& groups, std::unordered_map<std::string, std::size_t>& at, const std::string& label,
             std::size_t row) {
    const auto it = at.find(label);
    if (it == at.end()) {
        at.emplace(label, groups.size());
        groups.push_back(RawGroup{label, groups.size(), {row}});
    } else {
        groups[it->second].rows.push_back(row);
    }
}

// Chromosome codes: 1..22 (any digit string) as numbers, X 23, Y 24, MT/M 90, other
// names -1, -2, ... in order of first appearance.
int chrom_code(const std::string& tok, std::map<std::string, int>& other, int& next_other) {
    int chrom = 0;
    const bool numeric = !tok.empty() && std::all_of(tok.begin(), tok.end(), [](char c) {
        return std::isdigit(static_cast<unsigned char>(c));
    });
    if (numeric && parse_all(tok, chrom)) return chrom;
    if (tok == "X" || tok == "x") return 23;
    if (tok == "Y" || tok == "y") return 24;
    if (tok == "MT" || tok == "mt" || tok == "M") return 90;
    const auto it = other.find(tok);
    if (it != other.end()) return it->second;
    chrom = next_other--;
    other.emplace(tok, chrom);
    return chrom;
}

}  // namespace

GenoHeader parse_geno_header(const char* head, std::size_t len) {
    GenoHeader h;
    std::string text;
    for (std::size_t i = 0; i < len && i < kHeaderBytes && head[i] != '\0'; ++i) text.push_back(head[i]);
    const std::size_t b = text.find_first_not_of(" \t");
    if (b == std::string::npos) return h;
    const std::size_t e = text.find_first_of(" \t", b);
    const std::string magic = text.substr(b, e == std::string::npos ? std::string::npos : e - b);
    GenoFormat fmt;
    if (magic == "TGENO") fmt = GenoFormat::Tgeno;
    else if (magic == "GENO") fmt = GenoFormat::Geno;
    else return h;
    std::size_t counts[2] = {0, 0};
    int got = 0;
    std::size_t i = e == std::string::npos ? text.size() : e;
    while (i < text.size() && got < 2) {
        while (i < text.size() && !std::isdigit(static_cast<unsigned char>(text[i]))) ++i;
        if (i >= text.size()) break;
        std::size_t v = 0;
        while (i < text.size() && std::isdigit(static_cast<unsigned char>(text[i]))) {
            const std::size_t d = static_cast<std::size_t>(text[i] - '0');
            if (v > (std::numeric_limits<std::size_t>::max() - d) / 10) return h;   // overflow
            v = v * 10 + d;
            ++i;
        }
        counts[got++] = v;
    }
    if (got < 2) return h;
    h.format = fmt;
    h.n_ind = counts[0];
    h.n_snp = counts[1];
    if (fmt == GenoFormat::Tgeno) {
        h.header_bytes = kHeaderBytes;
        h.n_records = h.n_ind;
        h.bytes_per_record = packed_bytes(h.n_snp);
    } else {
        h.bytes_per_record = std::max(kHeaderBytes, packed_bytes(h.n_ind));
        h.header_bytes = h.bytes_per_record;
        h.n_records = h.n_snp;
    }
    return h;
}

GenoFile open_geno(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) io_fail("geno: cannot open " + path);
    char head[kHeaderBytes] = {};
    in.read(head, kHeaderBytes);
    const std::streamsize got = in.gcount();
    GenoFile f;
    f.path = path;
    if (got == static_cast<std::streamsize>(kHeaderBytes)) f.header = parse_geno_header(head, kHeaderBytes);
    const auto* m = reinterpret_cast<const unsigned char*>(head);
    const bool bed_magic = got >= 3 && m[0] == kBedMagic0 && m[1] == kBedMagic1;
    // (a file opening with the .bed magic cannot be EIGENSTRAT text: 'l' is not a genotype)
    if (f.header.format == GenoFormat::Unknown && !bed_magic) f.header = eigenstrat_geometry(path);
    if (f.header.format == GenoFormat::Unknown && !bed_magic) f.header = ancestrymap_geometry(path);
    if (f.header.format == GenoFormat::Eigenstrat || f.header.format == GenoFormat::Ancestrymap) {
        f.records_present = f.header.n_records;
        return f;
    }
    if (f.header.format == GenoFormat::Unknown && bed_magic) {
        if (m[2] != kBedSnpMajor)
            io_fail("geno: PLINK .bed is individual-major (mode byte 0x00); only SNP-major (0x01) is read: " + path);
        const std::string bim = plink_sibling(path, ".bim"), fam = plink_sibling(path, ".fam");
        const std::size_t n_snp = count_lines(bim), n_ind = count_lines(fam);
        if (n_snp == 0 || n_ind == 0)
            io_fail("geno: PLINK .bed found but its .bim / .fam (" + bim + ", " + fam + ") are missing or empty");
        f.header.format = GenoFormat::Plink;
        f.header.n_ind = n_ind;
        f.header.n_snp = n_snp;
        f.header.n_records = n_snp;
        f.header.bytes_per_record = packed_bytes(n_ind);
        f.header.header_bytes = 3;
    }
    if (f.header.format == GenoFormat::Unknown)
        io_fail("geno: not a packed TGENO / GENO file (48-byte header with magic TGENO or GENO) or a PLINK .bed: " +
                path);
    if (f.header.bytes_per_record == 0 || f.header.n_records == 0)
        io_fail("geno: degenerate header (zero records or stride): " + path);
    in.seekg(0, std::ios::end);
    const std::streamoff size = in.tellg();
    if (size < static_cast<std::streamoff>(f.header.header_bytes)) io_fail("geno: file smaller than its header: " + path);
    f.records_present = (static_cast<std::size_t>(size) - f.header.header_bytes) / f.header.bytes_per_record;
    if (f.records_present == 0) io_fail("geno: no complete record on disk: " + path);
    f.records_present = std::min(f.records_present, f.header.n_records);
    return f;
}

std::vector<PopGroup> read_ind(const std::string& path, const PopSelect& sel, std::size_t n_records_present,
                               std::size_t* n_individuals_total) {
    std::ifstream in(path);
    if (!in) io_fail("ind: cannot open " + path);
    std::vector<RawGroup> groups;
    std::unordered_map<std::string, std::size_t> at;
    std::size_t row = 0;
    std::string line;
    while (std::getline(in, line)) {
        const auto t = words(line);
        if (t.size() < 3) continue;
        if (row >= n_records_present) continue;
        add_row(groups, at, t[2], row);
        ++row;
    }
    if (n_individuals_total) *n_individuals_total = row;
    if (groups.empty()) io_fail("ind: no individuals parsed from " + path);
    return select_groups(groups, sel, path);
}

std::vector<PopGroup> read_fam(const std::string& path, const PopSelect& sel, std::size_t n_records_present) {
    std::ifstream in(path);
    if (!in) io_fail("fam: cannot open " + path);
    std::vector<RawGroup> groups;
    std::unordered_map<std::string, std::size_t> at;
    std::size_t row = 0;
    std::string line;
    while (std::getline(in, line)) {
        const auto t = words(line);
        if (t.size() < 6) continue;
        const std::string& pheno = t[5];
        if (row < n_records_present && pheno != "9" && pheno != "-9" && pheno != "0")
            add_row(groups, at, pheno == "1" ? "Control" : pheno == "2" ? "Case" : pheno, row);
        ++row;
    }
    if (groups.empty()) io_fail("fam: no individuals parsed from " + path);
    return select_groups(groups, sel, path);
}

SnpTable read_snp(const std::string& path) {
    std::ifstream in(path);
    if (!in) io_fail("snp: cannot open " + path);
    SnpTable t;
    std::map<std::string, int> other;
    int next_other = -1;
    std::string line;
    std::size_t line_no = 0;
    while (std::getline(in, line)) {
        ++line_no;
        const auto f = words(line);
        if (f.empty()) {
            if (in.peek() == std::char_traits<char>::eof()) break;
            io_fail("snp: blank line at line " + std::to_string(line_no) + " of " + path);
        }
        if (f.size() < 3)
            io_fail("snp: expected <id> <chrom> <genpos> at line " + std::to_string(line_no) + " of " + path);
        const int chrom = chrom_code(f[1], other, next_other);
        double g = 0.0;
        if (!parse_all(f[2], g) || !std::isfinite(g))
            io_fail("snp: malformed genetic position \"" + f[2] + "\" at line " + std::to_string(line_no));
        double ph = 0.0;
        if (f.size() > 3 && !(parse_all(f[3], ph) && std::isfinite(ph))) ph = 0.0;
        const bool alleles = f.size() >= 6;
        t.id.push_back(f[0]);
        t.chrom.push_back(chrom);
        t.genpos.push_back(g);
        t.physpos.push_back(ph);
        t.ref.push_back(alleles && !f[4].empty() ? f[4][0] : 'N');
        t.alt.push_back(alleles && !f[5].empty() ? f[5][0] : 'N');
    }
    return t;
}

SnpTable read_bim(const std::string& path) {
    std::ifstream in(path);
    if (!in) io_fail("bim: cannot open " + path);
    SnpTable t;
    std::map<std::string, int> other;
    int next_other = -1;
    std::string line;
    std::size_t line_no = 0;
    auto allele = [](const std::string& a) { return a.empty() || a[0] == '0' ? 'N' : a[0]; };
    while (std::getline(in, line)) {
        ++line_no;
        const auto f = words(line);
        if (f.empty()) {
            if (in.peek() == std::char_traits<char>::eof()) break;
            io_fail("bim: blank line at line " + std::to_string(line_no) + " of " + path);
        }
        if (f.size() < 6)
            io_fail("bim: expected <chrom> <id> <genpos> <physpos> <A1> <A2> at line " + std::to_string(line_no) +
                    " of " + path);
        const int chrom = chrom_code(f[0], other, next_other);
        double g = 0.0;
        if (!parse_all(f[2], g) || !std::isfinite(g))
            io_fail("bim: malformed genetic position \"" + f[2] + "\" at line " + std::to_string(line_no));
        double ph = 0.0;
        if (!(parse_all(f[3], ph) && std::isfinite(ph))) ph = 0.0;
        t.id.push_back(f[1]);
        t.chrom.push_back(chrom);
        t.genpos.push_back(g);
        t.physpos.push_back(ph);
        t.ref.push_back(allele(f[4]));
        t.alt.push_back(allele(f[5]));
    }
    return t;
}

void read_tgeno_selected(const GenoFile& f, const std::vector<PopGroup>& groups, std::size_t n_snp,
                         std::vector<std::uint8_t>& out) {
    STP_REQUIRE(f.header.format == GenoFormat::Tgeno, "geno: not a TGENO file");
    STP_REQUIRE(n_snp >= 1 && n_snp <= f.header.n_snp, "geno: SNP range out of bounds");
    STP_REQUIRE(!groups.empty(), "geno: empty population selection");
    const std::size_t bpr = packed_bytes(n_snp);
    std::size_t n = 0;
    for (const auto& g : groups)
        for (std::size_t r : g.rows) {
            STP_REQUIRE(r < f.records_present, "geno: individual row " + std::to_string(r) + " beyond the records on disk");
            ++n;
        }
    out.assign(n * bpr, 0);
    std::ifstream in(f.path, std::ios::binary);
    if (!in) io_fail("geno: cannot reopen " + f.path);
    std::size_t k = 0;
    for (const auto& g : groups)
        for (std::size_t r : g.rows) {
            in.seekg(static_cast<std::streamoff>(f.header.header_bytes + r * f.header.bytes_per_record));
            in.read(reinterpret_cast<char*>(out.data() + k * bpr), static_cast<std::streamsize>(bpr));
            if (in.gcount() != static_cast<std::streamsize>(bpr)) io_fail("geno: short read in " + f.path);
            ++k;
        }
}

AsciiGenoReader::AsciiGenoReader(const GenoFile& f) : f_(f), in_(f.path, std::ios::binary) {
    STP_REQUIRE(f.header.format == GenoFormat::Eigenstrat || f.header.format == GenoFormat::Ancestrymap,
                "geno: not an EIGENSTRAT or ANCESTRYMAP text file");
    if (!in_) io_fail("geno: cannot reopen " + f.path);
}

void AsciiGenoReader::next(std::size_t n, std::vector<std::uint8_t>& out) {
    const std::size_t bpr = f_.header.bytes_per_record, n_ind = f_.header.n_ind;
    out.assign(n * bpr, 0);
    std::string line;
    if (f_.header.format == GenoFormat::Ancestrymap) {
        for (std::size_t s = 0; s < n; ++s) {
            std::uint8_t* rec = out.data() + s * bpr;
            for (std::size_t i = 0; i < n_ind; ++i) {
                if (!std::getline(in_, line)) io_fail("geno: unexpected end of " + f_.path);
                ++line_;
                const auto t = words(line);
                std::uint8_t code = 0;
                if (t.size() != 3 || !ancestrymap_code(t[2], code))
                    io_fail("geno: malformed genotype line " + std::to_string(line_) + " of " + f_.path);
                rec[i >> 2] |= static_cast<std::uint8_t>(code << (6 - 2 * (i & 3)));
            }
        }
        return;
    }
    for (std::size_t s = 0; s < n; ++s) {
        if (!std::getline(in_, line)) io_fail("geno: unexpected end of " + f_.path + " at SNP line " + std::to_string(line_ + 1));
        ++line_;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.size() != n_ind) io_fail("geno: ragged SNP line " + std::to_string(line_) + " in " + f_.path);
        std::uint8_t* rec = out.data() + s * bpr;
        for (std::size_t i = 0; i < n_ind; ++i) {
            std::uint8_t code = 0;
            if (!genotype_char(line[i], code))
                io_fail("geno: illegal genotype character at SNP line " + std::to_string(line_) + " of " + f_.path);
            rec[i >> 2] |= static_cast<std::uint8_t>(code << (6 - 2 * (i & 3)));
        }
    }
}

void read_geno_records(const GenoFile& f, std::size_t s0, std::size_t s1, std::vector<std::uint8_t>& out) {
    STP_REQUIRE(f.header.format == GenoFormat::Geno || f.header.format == GenoFormat::Plink,
                "geno: not a SNP-major (GENO or PLINK .bed) file");
    STP_REQUIRE(s0 < s1 && s1 <= f.records_present, "geno: SNP records beyond the file");
    const std::size_t bpr = f.header.bytes_per_record;
    out.resize((s1 - s0) * bpr);
    std::ifstream in(f.path, std::ios::binary);
    if (!in) io_fail("geno: cannot reopen " + f.path);
    in.seekg(static_cast<std::streamoff>(f.header.header_bytes + s0 * bpr));
    in.read(reinterpret_cast<char*>(out.data()), static_cast<std::streamsize>(out.size()));
    if (in.gcount() != static_cast<std::streamsize>(out.size())) io_fail("geno: short read in " + f.path);
}

}  // namespace stp
