// steppe_amd/csrc/geno_io.hpp — readers for the packed EIGENSTRAT genotype triple
// (.geno TGENO / GENO, .snp, .ind) that feed the device extract.
//
// Formats (the reference's src/io/eigenstrat_format.hpp:18-128, eigenstrat_format.cpp:17-88):
//   .geno header: 48 bytes of text "<magic> <n_ind> <n_snp> ..." (NUL padded), magic
//   TGENO (individual-major: one record of ceil(n_snp/4) bytes per individual after the
//   48-byte header) or GENO (SNP-major PACKEDANCESTRYMAP: one record of
//   max(48, ceil(n_ind/4)) bytes per SNP, the header occupying the first record).
//   2-bit codes, first position in the two most significant bits, 3 = missing.
//   .ind: "<id> <sex> <population>" per individual, in record order.
//   .snp: "<id> <chrom> <genpos (Morgans)> [<physpos> [<ref> <alt>]]" per SNP.
// EIGENSTRAT ASCII (src/io/geno_reader.cpp:83-120, 516-583): one line per SNP of n_ind
//   characters 0 / 1 / 2 (copies of ref) or 9 (missing), every line the same length.
// ANCESTRYMAP (src/io/geno_reader.cpp:122-177, 648-744): "<snp_id> <sample_id> <genotype>"
//   lines, genotype 0 / 1 / 2 or -1 (missing), SNP-major by position (line L is SNP
//   L / n_ind, individual L % n_ind; the ids are not looked up), n_snp and n_ind counted
//   from the .snp and .ind beside the .geno.
// PLINK (src/io/geno_reader.cpp:208-238, 585-634; src/io/plink_reader.cpp): a SNP-major
//   .bed (magic 6c 1b 01, one record of ceil(n_ind/4) bytes per SNP, individual i in bits
//   2(i%4) of byte i/4, codes 00 two copies of A1, 01 missing, 10 heterozygous, 11 none),
//   its geometry counted from the non-empty lines of the .bim and .fam beside it; the
//   .bim "<chrom> <id> <genpos> <physpos> <A1> <A2>" (A1, the counted allele, is ref;
//   allele '0' is 'N'); the .fam phenotype column names the population (1 Control,
//   2 Case, 9 / -9 / 0 not assigned).
#pragma once

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

namespace stp {

enum class GenoFormat { Unknown = 0, Tgeno = 1, Geno = 2, Plink = 3, Eigenstrat = 4, Ancestrymap = 5 };

struct GenoHeader {
    GenoFormat format = GenoFormat::Unknown;
    std::size_t n_ind = 0, n_snp = 0, n_records = 0, bytes_per_record = 0, header_bytes = 48;
};

// Header text -> geometry (src/io/eigenstrat_format.cpp:17-88): the magic is the first
// whitespace-delimited word, the counts the first two digit runs after it.
GenoHeader parse_geno_header(const char* head, std::size_t len);

// An opened packed .geno: header plus the number of whole records present on disk
// (src/io/geno_reader.cpp:195-271).
struct GenoFile {
    std::string path;
    GenoHeader header;
    std::size_t records_present = 0;
};
GenoFile open_geno(const std::string& path);

// Population selection (src/io/ind_reader.hpp:20-31, detail/pop_select.hpp:25-76):
// explicit labels, the k largest populations (ties by first appearance), or every
// population with at least min_n individuals; the result is sorted by label and each
// population's rows ascend.
struct PopSelect {
    enum Mode { kExplicit = 0, kTopK = 1, kMinN = 2 } mode = kExplicit;
    std::vector<std::string> labels;
    std::size_t k = 0, min_n = 1;
};
struct PopGroup {
    std::string label;
    std::vector<std::size_t> rows;
};
// .ind rows at or beyond n_records_present are ignored (src/io/ind_reader.cpp:31-33).
std::vector<PopGroup> read_ind(const std::string& path, const PopSelect& sel, std::size_t n_records_present,
                               std::size_t* n_individuals_total = nullptr);

// .fam groups (src/io/plink_reader.cpp:104-151), selected as read_ind selects.
std::vector<PopGroup> read_fam(const std::string& path, const PopSelect& sel, std::size_t n_records_present);

// .snp table (src/io/snp_reader.cpp:20-63, detail/snp_text_parse.hpp): chromosome codes
// 1..22 as numbers, X 23, Y 24, MT/M 90, other names -1, -2, ... in order of appearance;
// genetic position strict, physical position 0 when unparseable, alleles 'N' when absent.
struct SnpTable {
    std::vector<std::string> id;
    std::vector<std::int32_t> chrom;
    std::vector<double> genpos, physpos;
    std::vector<char> ref, alt;
    std::size_t size() const { return chrom.size(); }
};
SnpTable read_snp(const std::string& path);
// .bim table (src/io/plink_reader.cpp:53-100), in the .snp table's terms.
SnpTable read_bim(const std::string& path);

// TGENO: the first ceil(n_snp/4) bytes of each selected individual's record, in the
// groups' order, into out (n_sel x that many bytes) (src/io/geno_reader.cpp:274-372).
void read_tgeno_selected(const GenoFile& f, const std::vector<PopGroup>& groups, std::size_t n_snp,
                         std::vector<std::uint8_t>& out);
// EIGENSTRAT or ANCESTRYMAP text, read in order: the next n SNPs packed into canonical
// SNP-major records of ceil(n_ind/4) bytes (first individual in the top bits).
class AsciiGenoReader {
public:
    explicit AsciiGenoReader(const GenoFile& f);
    void next(std::size_t n, std::vector<std::uint8_t>& out);

private:
    const GenoFile& f_;
    std::ifstream in_;
    std::size_t line_ = 0;
};

// GENO or PLINK .bed: SNP records [s0, s1) verbatim (s1 - s0 records of
// bytes_per_record bytes) (src/io/geno_reader.cpp:465-505, 585-634).
void read_geno_records(const GenoFile& f, std::size_t s0, std::size_t s1, std::vector<std::uint8_t>& out);

}  // namespace stp
