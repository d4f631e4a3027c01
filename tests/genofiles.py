"""Write synthetic packed EIGENSTRAT triples (.geno TGENO / GENO, .snp, .ind) for tests.

Layout (the reference's src/io/eigenstrat_format.hpp:18-128): a 48-byte NUL-padded text
header "<magic> <n_ind> <n_snp> ..."; TGENO stores one record of ceil(n_snp/4) bytes per
individual after the header, GENO one record of max(48, ceil(n_ind/4)) bytes per SNP
with the header in the first record; 2-bit codes, first position in the top bits.
"""
from __future__ import annotations

import numpy as np


def pack_rows(codes: np.ndarray) -> np.ndarray:
    """(rows, n) codes 0..3 -> (rows, ceil(n/4)) bytes, position 0 in bits 7:6."""
    rows, n = codes.shape
    w = (n + 3) // 4
    pad = np.zeros((rows, 4 * w), np.uint8)
    pad[:, :n] = codes
    q = pad.reshape(rows, w, 4).astype(np.uint8)
    return (q[:, :, 0] << 6) | (q[:, :, 1] << 4) | (q[:, :, 2] << 2) | q[:, :, 3]


def header(magic: str, n_ind: int, n_snp: int, width: int = 48) -> bytes:
    text = f"{magic} {n_ind} {n_snp} 1a2b3c 4d5e6f".encode()
    return text + b"\0" * (width - len(text))


def write_triple(prefix: str, codes: np.ndarray, ind_pops, snp, fmt: str = "tgeno", records_missing: int = 0):
    """codes: (n_ind, n_snp) genotype codes 0/1/2 (reference allele copies) or 3 (missing).
    ind_pops: population label per individual.  snp: dict with id, chrom (str), genpos,
    physpos, ref, alt lists.  records_missing drops that many trailing records."""
    n_ind, n_snp = codes.shape
    if fmt == "tgeno":
        recs = pack_rows(codes)
        body = header("TGENO", n_ind, n_snp) + recs[: n_ind - records_missing].tobytes()
    else:
        bpr = max(48, (n_ind + 3) // 4)
        recs = np.zeros((n_snp, bpr), np.uint8)
        recs[:, : (n_ind + 3) // 4] = pack_rows(np.ascontiguousarray(codes.T))
        body = header("GENO", n_ind, n_snp, bpr) + recs[: n_snp - records_missing].tobytes()
    with open(prefix + ".geno", "wb") as f:
        f.write(body)
    with open(prefix + ".ind", "w") as f:
        for i, p in enumerate(ind_pops):
            f.write(f"I{i:05d} {'MF'[i % 2]} {p}\n")
    with open(prefix + ".snp", "w") as f:
        for k in range(len(snp["id"])):
            f.write(f"{snp['id'][k]} {snp['chrom'][k]} {snp['genpos'][k]!r} {snp['physpos'][k]:.0f} "
                    f"{snp['ref'][k]} {snp['alt'][k]}\n")


def synthetic_triple(prefix: str, seed: int, n_snp: int = 3000, fmt: str = "tgeno", sizes=None,
                     records_missing: int = 0, zero_map: bool = False):
    """A panel of 9 populations (sizes drawn or given), a third of the individuals
    pseudo-haploid, per-population missingness, 22 autosomes plus X and an unplaced contig,
    and allele columns that exercise the filters: strand-ambiguous pairs, 'N' alleles,
    equal alleles, monomorphic and low-frequency SNPs.  Returns the labels per individual."""
    rng = np.random.default_rng(seed)
    labels = ["Pop_" + c for c in "HCAFIBGDE"]
    if sizes is None:
        sizes = rng.integers(2, 14, len(labels))
    ind_pops = np.repeat(labels, sizes)
    rng.shuffle(ind_pops)
    n_ind = len(ind_pops)
    anc = rng.uniform(0.02, 0.98, n_snp)
    anc[rng.random(n_snp) < 0.04] = 0.0            # monomorphic in the whole panel
    anc[rng.random(n_snp) < 0.04] = 0.01
    freq = {}
    for lab in labels:
        fst = rng.uniform(0.02, 0.2)
        a, b = anc * (1 - fst) / fst, (1 - anc) * (1 - fst) / fst
        freq[lab] = np.where(anc > 0, rng.beta(np.maximum(a, 1e-3), np.maximum(b, 1e-3)), 0.0)
    codes = np.zeros((n_ind, n_snp), np.uint8)
    haploid = rng.random(n_ind) < 0.33
    for i, lab in enumerate(ind_pops):
        p = freq[lab]
        if haploid[i]:
            codes[i] = 2 * (rng.random(n_snp) < p)
        else:
            codes[i] = (rng.random(n_snp) < p).astype(np.uint8) + (rng.random(n_snp) < p)
        miss = rng.random(n_snp) < rng.uniform(0.0, 0.3)
        codes[i, miss] = 3
    chrom, genpos, physpos = [], [], []
    per = n_snp // 24
    for k in range(n_snp):
        c = min(k // per, 23)
        chrom.append(str(c + 1) if c < 22 else ("X" if c == 22 else "chrUn"))
    pos = np.sort(rng.uniform(0, 2.0, n_snp))
    for k in range(n_snp):
        genpos.append(0.0 if zero_map else float(np.round(pos[k], 6)))
        physpos.append(float(1000 + 400000 * (pos[k])))
    bases = np.array(list("ACGT"))
    ref = bases[rng.integers(0, 4, n_snp)]
    alt = bases[rng.integers(0, 4, n_snp)]
    odd = rng.random(n_snp)
    alt[odd < 0.03] = "N"
    ref[(odd >= 0.03) & (odd < 0.05)] = "x"
    low = (odd >= 0.05) & (odd < 0.08)
    ref[low] = np.char.lower(ref[low])
    snp = {"id": [f"rs{k}" for k in range(n_snp)], "chrom": chrom, "genpos": genpos, "physpos": physpos,
           "ref": list(ref), "alt": list(alt)}
    if fmt == "ascii":
        write_eigenstrat_text(prefix, codes, ind_pops, snp)
    elif fmt == "ancestrymap":
        write_eigenstrat_text(prefix, codes, ind_pops, snp)
        tok = np.array(["0", "1", "2", "-1"])
        n_ind, n_snp = codes.shape
        with open(prefix + ".geno", "w") as f:
            for s in range(n_snp):
                t = tok[codes[:, s]]
                f.write("".join(f"rs{s} I{i:05d} {t[i]}\n" for i in range(n_ind)))
    elif fmt == "plink":
        # group labels as .fam phenotypes: two of them become the reserved codes 1 / 2
        # (Control / Case) and one is the skipped missing phenotype -9
        pheno = {"Pop_A": "1", "Pop_B": "2", "Pop_I": "-9"}
        write_plink(prefix, codes, [pheno.get(p, p) for p in ind_pops], snp, records_missing)
    else:
        write_triple(prefix, codes, ind_pops, snp, fmt, records_missing)
    return list(ind_pops)


def write_plink(prefix: str, codes: np.ndarray, ind_pops, snp, records_missing: int = 0):
    """PLINK SNP-major .bed (magic 6c 1b 01; per SNP ceil(n_ind/4) bytes, individual i in
    bits 2(i%4) of byte i/4: 00 two copies of A1, 01 missing, 10 heterozygous, 11 no A1),
    .bim (<chrom> <id> <genpos> <physpos> <A1> <A2>, A1 = the counted allele) and .fam
    (<fid> <iid> <father> <mother> <sex> <phenotype>, the phenotype naming the group).
    codes: (n_ind, n_snp) copies of A1 0/1/2, or 3 missing."""
    n_ind, n_snp = codes.shape
    bed_of = np.array([3, 2, 0, 1], np.uint8)          # A1 copies 0, 1, 2, missing -> .bed code
    c = bed_of[codes.T]                                 # (n_snp, n_ind)
    w = (n_ind + 3) // 4
    pad = np.zeros((n_snp, 4 * w), np.uint8)
    pad[:, :n_ind] = c
    q = pad.reshape(n_snp, w, 4)
    rec = q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)
    with open(prefix + ".bed", "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]) + rec[: n_snp - records_missing].tobytes())
    with open(prefix + ".fam", "w") as f:
        for i, p in enumerate(ind_pops):
            f.write(f"F{i} I{i:05d} 0 0 {1 + i % 2} {p}\n")
    with open(prefix + ".bim", "w") as f:
        for k in range(len(snp["id"])):
            a1 = "0" if snp["ref"][k] == "N" else snp["ref"][k]
            f.write(f"{snp['chrom'][k]}\t{snp['id'][k]}\t{snp['genpos'][k]!r}\t{snp['physpos'][k]:.0f}\t"
                    f"{a1}\t{snp['alt'][k]}\n")


def write_eigenstrat_text(prefix: str, codes: np.ndarray, ind_pops, snp):
    """EIGENSTRAT text .geno (one line per SNP, one character 0 / 1 / 2 / 9 per
    individual) with the same .ind and .snp as the packed triple."""
    n_ind, n_snp = codes.shape
    chars = np.array([ord("0"), ord("1"), ord("2"), ord("9")], np.uint8)[codes.T]   # (n_snp, n_ind)
    body = np.concatenate([chars, np.full((n_snp, 1), ord("\n"), np.uint8)], axis=1)
    with open(prefix + ".geno", "wb") as f:
        f.write(body.tobytes())
    with open(prefix + ".ind", "w") as f:
        for i, p in enumerate(ind_pops):
            f.write(f"I{i:05d} {'MF'[i % 2]} {p}\n")
    with open(prefix + ".snp", "w") as f:
        for k in range(len(snp["id"])):
            f.write(f"{snp['id'][k]} {snp['chrom'][k]} {snp['genpos'][k]!r} {snp['physpos'][k]:.0f} "
                    f"{snp['ref'][k]} {snp['alt'][k]}\n")
