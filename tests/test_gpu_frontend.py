"""Genotype files -> f2 on the device (stp_extract_f2_files) against the oracle.

Synthetic TGENO and GENO triples (tests/genofiles.py) through the device front end
(host readers, the on-device GENO transpose + gather, ploidy detection, decode, the
device keep mask and compaction, blocks, the f2 contraction), compared with the restated
oracle (oracle/cpu_ref_frontend.cpp, itself bit-identical to the reference's own chain):
labels, SNP counts, ploidy counts, block sizes and vpair exactly; f2 within 1e-12
relative (floor 1e-6 of the tensor's largest value) in the FP64 form and 1e-9 in the
split-BF16 form.  Also the reference-style prefix call, the STPF2BK1 output directory,
GENO streamed in several transpose blocks, and the errors the reference raises.
"""
import json
import os

import numpy as np
import pytest

import steppe_amd as st
from steppe_amd import _core
from genofiles import synthetic_triple

pytestmark = pytest.mark.gpu

CASES = [
    dict(pops=["Pop_A", "Pop_C", "Pop_F", "Pop_H"]),
    dict(top_k=5, maxmiss=1.0, autosomes_only=False),
    dict(min_n=6, maf=0.05, drop_monomorphic=True, transversions_only=True, strand_mode="keep"),
    dict(pops=["Pop_B", "Pop_I", "Pop_E"], maxmiss=0.34, strand_mode="flip", ploidy="2", blgsize=0.02),
    dict(pops=["Pop_D", "Pop_G"], ploidy="1", maxmiss=1.0),
]
STRAND = {"drop": 0, "keep": 1, "flip": 2}


def _oracle_kw(kw):
    o = dict(kw)
    o["strand_mode"] = STRAND[o.get("strand_mode", "drop")]
    o["ploidy"] = {"auto": 0, "1": 1, "2": 2}[str(o.get("ploidy", "auto"))]
    return o


def _paths(pre):
    return pre + ".geno", pre + ".snp", pre + ".ind"


def check(h, info, o, rtol, what):
    assert h.pops == o["labels"], what
    for k, v in o["info"].items():
        assert info[k] == v, (what, k, info[k], v)
    assert np.array_equal(np.asarray(h.block_sizes), o["block_sizes"]), what
    f2, vp = h._f2_numpy(), h._vpair_numpy()
    assert np.array_equal(vp, o["vpair"]), what
    floor = 1e-6 * np.abs(o["f2"]).max()
    rel = np.abs(f2 - o["f2"]) / np.maximum(np.abs(o["f2"]), floor)
    assert rel.max() <= rtol, (what, float(rel.max()))
    print(f"{what}: P={h.P} kept {info['n_snp_kept']}/{info['n_snp_total']} blocks {h.n_block} "
          f"max rel f2 {rel.max():.2e}")


@pytest.mark.parametrize("fmt", ["tgeno", "geno"])
def test_extract_matches_oracle(cref, tmp_path, fmt):
    pre = str(tmp_path / f"p_{fmt}")
    synthetic_triple(pre, 11, n_snp=2400, fmt=fmt)
    for i, kw in enumerate(CASES):
        h, info = _core.extract_f2_files(*_paths(pre), precision="fp64", **kw)
        o = cref.extract(*_paths(pre), **_oracle_kw(kw))
        check(h, info, o, 1e-12, f"{fmt} case {i} fp64")
    h, info = _core.extract_f2_files(*_paths(pre), precision="split_bf16", **CASES[0])
    check(h, info, cref.extract(*_paths(pre), **_oracle_kw(CASES[0])), 1e-9, f"{fmt} split")


def test_geno_streamed_in_blocks(cref, tmp_path, monkeypatch):
    """A GENO file transposed in 64-SNP-multiple blocks (STEPPE_GENO_BLOCK_BYTES) gives the
    same tensor as one block, with a zero genetic map (2 Mb physical blocks)."""
    pre = str(tmp_path / "g")
    synthetic_triple(pre, 12, n_snp=2000, fmt="geno", zero_map=True)
    kw = dict(pops=["Pop_A", "Pop_E", "Pop_H"], maxmiss=1.0)
    one, _ = _core.extract_f2_files(*_paths(pre), precision="fp64", **kw)
    monkeypatch.setenv("STEPPE_GENO_BLOCK_BYTES", str(48 * 64 * 3))
    many, info = _core.extract_f2_files(*_paths(pre), precision="fp64", **kw)
    assert np.array_equal(one._f2_numpy(), many._f2_numpy())
    check(many, info, cref.extract(*_paths(pre), **kw), 1e-12, "geno in blocks")


def test_prefix_call_and_output_dir(cref, tmp_path):
    pre = str(tmp_path / "x")
    synthetic_triple(pre, 13, n_snp=1500, fmt="tgeno")
    pops = ["Pop_C", "Pop_A", "Pop_I"]
    h = _core.run_extract_f2(pre, pops, "", 0, 0.05, 0.0, 1.0, True, False, False, "auto", "drop", "fp64")
    assert h.pops == sorted(pops)
    # the package facade takes the reference's keyword form
    hf = st.extract_f2(pre, pops=pops, maxmiss=1.0, precision="fp64")
    assert hf.pops == h.pops and np.array_equal(hf._f2_numpy(), h._f2_numpy())
    out = str(tmp_path / "f2dir")
    path, info = _core.extract_f2_files(*_paths(pre), pops, out=out, maxmiss=1.0, precision="fp64")
    assert path == out
    f2, vp, bs, names = _core.read_f2_host(out)
    assert names == sorted(pops) and np.array_equal(f2, h._f2_numpy()) and np.array_equal(bs, h.block_sizes)
    meta = json.load(open(os.path.join(out, "meta.json")))
    assert meta["format"] == "STPF2BK1" and meta["n_snp_kept"] == info["n_snp_kept"]
    assert meta["filters"]["maxmiss"] == 1.0 and meta["blgsize_cm"] == 5.0
    # the handle feeds the fits directly
    d = st.f4(h, [("Pop_A", "Pop_C", "Pop_I", "Pop_A")])
    assert len(d) == 1


def test_errors(tmp_path):
    pre = str(tmp_path / "e")
    synthetic_triple(pre, 31, n_snp=800, fmt="geno")
    with pytest.raises(ValueError):
        _core.extract_f2_files(*_paths(pre), ["Pop_A", "Nowhere"])
    with pytest.raises(ValueError):
        _core.extract_f2_files(*_paths(pre), ["Pop_A"], maf=0.6)
    synthetic_triple(pre, 31, n_snp=800, fmt="geno", records_missing=5)
    with pytest.raises(ValueError):
        _core.extract_f2_files(*_paths(pre), ["Pop_A"])
    # an individual-major PLINK .bed (mode byte 0) is refused, as by the reference
    synthetic_triple(pre, 31, n_snp=400, fmt="plink")
    raw = bytearray(open(pre + ".bed", "rb").read())
    raw[2] = 0x00
    open(pre + ".bed", "wb").write(bytes(raw))
    with pytest.raises(ValueError, match="individual-major"):
        _core.extract_f2_files(pre + ".bed", pre + ".bim", pre + ".fam", ["Pop_C"])


PLINK_CASES = [
    dict(pops=["Control", "Pop_C", "Pop_F", "Pop_H"]),
    dict(top_k=5, maxmiss=1.0, autosomes_only=False),
    dict(pops=["Case", "Pop_E", "Control"], maxmiss=0.34, strand_mode="flip", ploidy="2", blgsize=0.02),
]


def test_plink_matches_oracle(cref, tmp_path, monkeypatch):
    """PLINK .bed/.bim/.fam through the device (recoded in the transpose), against the
    oracle, also when the .bed is transposed in several blocks; the prefix call picks the
    .bed when no .geno exists."""
    pre = str(tmp_path / "pl")
    synthetic_triple(pre, 14, n_snp=2400, fmt="plink")
    paths = (pre + ".bed", pre + ".bim", pre + ".fam")
    for i, kw in enumerate(PLINK_CASES):
        h, info = _core.extract_f2_files(*paths, precision="fp64", **kw)
        check(h, info, cref.extract(*paths, **_oracle_kw(kw)), 1e-12, f"plink case {i}")
    monkeypatch.setenv("STEPPE_GENO_BLOCK_BYTES", str(64 * 3 * 30))
    h, info = _core.extract_f2_files(*paths, precision="fp64", **PLINK_CASES[0])
    check(h, info, cref.extract(*paths, **_oracle_kw(PLINK_CASES[0])), 1e-12, "plink in blocks")
    monkeypatch.delenv("STEPPE_GENO_BLOCK_BYTES")
    hp = _core.run_extract_f2(pre, ["Pop_C", "Pop_D", "Pop_H"], precision="fp64")
    assert hp.pops == ["Pop_C", "Pop_D", "Pop_H"]
    # the same panel as GENO gives the same tensor over populations named alike
    g = str(tmp_path / "g")
    synthetic_triple(g, 14, n_snp=2400, fmt="geno")
    hg = _core.run_extract_f2(g, ["Pop_C", "Pop_D", "Pop_H"], precision="fp64")
    assert np.array_equal(hp._f2_numpy(), hg._f2_numpy()) and np.array_equal(hp._vpair_numpy(), hg._vpair_numpy())


def test_eigenstrat_text_matches_oracle(cref, tmp_path, monkeypatch):
    """EIGENSTRAT text .geno (one character per genotype) through the device, against the
    oracle, in one and in several transpose blocks, and equal to the packed GENO of the
    same panel."""
    pre = str(tmp_path / "t")
    synthetic_triple(pre, 16, n_snp=1500, fmt="ascii")
    for i, kw in enumerate(CASES[:2]):
        h, info = _core.extract_f2_files(*_paths(pre), precision="fp64", **kw)
        check(h, info, cref.extract(*_paths(pre), **_oracle_kw(kw)), 1e-12, f"text case {i}")
    monkeypatch.setenv("STEPPE_GENO_BLOCK_BYTES", str(64 * 30))
    many, _ = _core.extract_f2_files(*_paths(pre), precision="fp64", **CASES[0])
    monkeypatch.delenv("STEPPE_GENO_BLOCK_BYTES")
    g = str(tmp_path / "g")
    synthetic_triple(g, 16, n_snp=1500, fmt="geno")
    packed, _ = _core.extract_f2_files(*_paths(g), precision="fp64", **CASES[0])
    assert np.array_equal(many._f2_numpy(), packed._f2_numpy())
    assert np.array_equal(many._vpair_numpy(), packed._vpair_numpy())


def test_ancestrymap_matches_oracle(cref, tmp_path):
    pre = str(tmp_path / "a")
    synthetic_triple(pre, 11, n_snp=2400, fmt="ancestrymap")
    h, info = _core.extract_f2_files(*_paths(pre), precision="fp64", **CASES[0])
    check(h, info, cref.extract(*_paths(pre), **_oracle_kw(CASES[0])), 1e-12, "ancestrymap")
