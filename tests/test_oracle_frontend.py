"""Pin the restated genotype-file extract (oracle/cpu_ref_frontend.cpp).

* bit-for-bit against the reference's own chain (oracle/_ref: GenoReader, read_ind,
  read_snp, transpose_to_canonical, detect_sample_ploidy_device, decode_af,
  build_snp_keep_mask, assign_blocks, compute_f2_blocks, glued as the CPU branch of
  run_extract_f2) on synthetic TGENO and GENO triples, for explicit / top-K / min-N
  selections, every filter, explicit ploidy, an all-zero genetic map (the 2 Mb physical
  fallback) and a file shorter than its header;
* the TGENO and GENO encodings of one panel give bit-identical tensors (the transpose);
* the header geometry and the error cases the reference raises on.
"""
import os

import numpy as np
import pytest

from genofiles import header, synthetic_triple

CASES = [
    dict(pops=["Pop_A", "Pop_C", "Pop_F", "Pop_H"]),
    dict(top_k=5, maxmiss=1.0, autosomes_only=False),
    dict(min_n=6, maf=0.05, drop_monomorphic=True, transversions_only=True, strand_mode=1),
    dict(pops=["Pop_B", "Pop_I", "Pop_E"], maxmiss=0.34, strand_mode=2, ploidy=2, blgsize=0.02),
    dict(pops=["Pop_D", "Pop_G"], ploidy=1, maxmiss=1.0),
]


def _paths(pre):
    return pre + ".geno", pre + ".snp", pre + ".ind"


def _bits(a):
    return np.ascontiguousarray(a).view(np.int64)


def _same(a, b):
    assert a["labels"] == b["labels"] and a["info"] == b["info"]
    assert np.array_equal(a["block_sizes"], b["block_sizes"])
    assert np.array_equal(_bits(a["f2"]), _bits(b["f2"])) and np.array_equal(_bits(a["vpair"]), _bits(b["vpair"]))


@pytest.mark.parametrize("fmt", ["tgeno", "geno"])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_bit_identical_to_reference(cref, ref, tmp_path, fmt, case):
    pre = str(tmp_path / f"p_{fmt}")
    synthetic_triple(pre, 11, n_snp=2400, fmt=fmt)
    a = cref.extract(*_paths(pre), **CASES[case])
    b = ref.extract(*_paths(pre), **CASES[case])
    _same(a, b)
    assert a["info"]["n_snp_kept"] > 100


def test_zero_map_and_short_file(cref, ref, tmp_path):
    pre = str(tmp_path / "z")
    synthetic_triple(pre, 12, n_snp=1600, fmt="geno", zero_map=True)
    a = cref.extract(*_paths(pre), pops=["Pop_A", "Pop_E"], maxmiss=1.0)
    b = ref.extract(*_paths(pre), pops=["Pop_A", "Pop_E"], maxmiss=1.0)
    _same(a, b)
    assert a["info"]["n_snp_total"] == 1600 and len(a["block_sizes"]) > 20
    # a GENO file with fewer SNP records than its header and .snp is an error in both
    synthetic_triple(pre, 12, n_snp=1600, fmt="geno", records_missing=37)
    for lib in (cref, ref):
        with pytest.raises(ValueError):
            lib.extract(*_paths(pre), pops=["Pop_A", "Pop_E"], maxmiss=1.0)
    pre = str(tmp_path / "s")
    synthetic_triple(pre, 13, n_snp=1200, fmt="tgeno", sizes=[3] * 9, records_missing=4)
    a = cref.extract(*_paths(pre), top_k=9, maxmiss=1.0)
    b = ref.extract(*_paths(pre), top_k=9, maxmiss=1.0)
    _same(a, b)
    assert a["info"]["n_pseudo_haploid"] + a["info"]["n_diploid"] == 27 - 4


def test_transpose_invariance(cref, tmp_path):
    t, g = str(tmp_path / "t"), str(tmp_path / "g")
    synthetic_triple(t, 21, n_snp=1800, fmt="tgeno")
    synthetic_triple(g, 21, n_snp=1800, fmt="geno")
    for kw in CASES[:3]:
        _same(cref.extract(*_paths(t), **kw), cref.extract(*_paths(g), **kw))


def test_header_geometry(tmp_path):
    """GENO records are max(48, ceil(n_ind/4)) bytes with the header in the first record."""
    from genofiles import write_triple
    for n_ind, want in ((10, 48), (300, 75)):
        codes = np.zeros((n_ind, 5), np.uint8)
        snp = {"id": list("abcde"), "chrom": ["1"] * 5, "genpos": [0.0, 0.1, 0.2, 0.3, 0.4],
               "physpos": [1.0] * 5, "ref": ["A"] * 5, "alt": ["C"] * 5}
        pre = str(tmp_path / f"h{n_ind}")
        write_triple(pre, codes, ["P"] * n_ind, snp, "geno")
        assert os.path.getsize(pre + ".geno") == want * 6
    assert header("TGENO", 3, 7)[:11] == b"TGENO 3 7 1" and len(header("TGENO", 3, 7)) == 48


def test_errors(cref, ref, tmp_path):
    pre = str(tmp_path / "e")
    synthetic_triple(pre, 31, n_snp=800)
    for lib in (cref, ref):
        with pytest.raises(ValueError):
            lib.extract(*_paths(pre), pops=["Pop_A", "Nowhere"])
        with pytest.raises(ValueError):
            lib.extract(*_paths(pre), pops=["Pop_A"], maf=0.6)        # folded MAF <= 0.5: nothing kept


# PLINK: the .fam phenotype names the group (Pop_A is written as 1 -> "Control", Pop_B as
# 2 -> "Case", Pop_I as the skipped -9)
PLINK_CASES = [
    dict(pops=["Control", "Pop_C", "Pop_F", "Pop_H"]),
    dict(top_k=5, maxmiss=1.0, autosomes_only=False),
    dict(min_n=6, maf=0.05, drop_monomorphic=True, transversions_only=True, strand_mode=1),
    dict(pops=["Case", "Pop_E", "Control"], maxmiss=0.34, strand_mode=2, ploidy=2, blgsize=0.02),
]


def _plink(pre):
    return pre + ".bed", pre + ".bim", pre + ".fam"


@pytest.mark.parametrize("case", range(len(PLINK_CASES)))
def test_plink_oracle_bit_identical_to_reference(cref, ref, tmp_path, case):
    pre = str(tmp_path / "b")
    synthetic_triple(pre, 14, n_snp=2400, fmt="plink")
    a = cref.extract(*_plink(pre), **PLINK_CASES[case])
    b = ref.extract(*_plink(pre), **PLINK_CASES[case])
    _same(a, b)
    assert a["info"]["n_snp_kept"] > 100
    assert "Pop_I" not in a["labels"] and "Pop_A" not in a["labels"]


def test_plink_equals_geno_and_short_bed(cref, ref, tmp_path):
    # the same panel as PLINK and as GENO: identical tensors over populations named alike
    b, g = str(tmp_path / "b"), str(tmp_path / "g")
    synthetic_triple(b, 15, n_snp=1800, fmt="plink")
    synthetic_triple(g, 15, n_snp=1800, fmt="geno")
    for kw in (dict(pops=["Pop_C", "Pop_D", "Pop_H"]), dict(pops=["Pop_E", "Pop_G"], ploidy=1, maxmiss=1.0)):
        _same(cref.extract(*_plink(b), **kw), cref.extract(*_paths(g), **kw))
    # fewer SNP records in the .bed than the .bim lists: an error in both
    synthetic_triple(b, 15, n_snp=1800, fmt="plink", records_missing=9)
    for lib in (cref, ref):
        with pytest.raises(ValueError):
            lib.extract(*_plink(b), pops=["Pop_C", "Pop_D"])


@pytest.mark.parametrize("case", [0, 1, 3])
def test_eigenstrat_text_bit_identical_to_reference(cref, ref, tmp_path, case):
    pre = str(tmp_path / "e")
    synthetic_triple(pre, 16, n_snp=1500, fmt="ascii")
    a = cref.extract(*_paths(pre), **CASES[case])
    b = ref.extract(*_paths(pre), **CASES[case])
    _same(a, b)
    # the same panel packed as GENO gives the same tensor
    g = str(tmp_path / "g")
    synthetic_triple(g, 16, n_snp=1500, fmt="geno")
    _same(a, cref.extract(*_paths(g), **CASES[case]))


def test_eigenstrat_text_errors(cref, ref, tmp_path):
    pre = str(tmp_path / "e")
    synthetic_triple(pre, 17, n_snp=400, fmt="ascii")
    raw = open(pre + ".geno", "rb").read().split(b"\n")
    raw[7] = raw[7][:-1]                                  # one ragged line: not a genotype file
    open(pre + ".geno", "wb").write(b"\n".join(raw))
    for lib in (cref, ref):
        with pytest.raises(ValueError):
            lib.extract(*_paths(pre), pops=["Pop_C", "Pop_D"])


@pytest.mark.parametrize("case", [0, 2])
def test_ancestrymap_bit_identical_to_reference(cref, ref, tmp_path, case):
    """ANCESTRYMAP "<snp> <ind> <genotype>" triples: the oracle equals the reference's
    reader chain, and the packed GENO of the same panel."""
    pre = str(tmp_path / "a")
    synthetic_triple(pre, 18, n_snp=600, fmt="ancestrymap")
    a = cref.extract(*_paths(pre), **CASES[case])
    b = ref.extract(*_paths(pre), **CASES[case])
    _same(a, b)
    g = str(tmp_path / "g")
    synthetic_triple(g, 18, n_snp=600, fmt="geno")
    _same(a, cref.extract(*_paths(g), **CASES[case]))


def test_individual_major_bed_is_rejected(cref, ref, tmp_path):
    pre = str(tmp_path / "b")
    synthetic_triple(pre, 19, n_snp=400, fmt="plink")
    raw = bytearray(open(pre + ".bed", "rb").read())
    raw[2] = 0x00                                          # mode byte: individual-major
    open(pre + ".bed", "wb").write(bytes(raw))
    for lib in (cref, ref):
        with pytest.raises(ValueError):
            lib.extract(pre + ".bed", pre + ".bim", pre + ".fam", pops=["Pop_C", "Pop_D"])
