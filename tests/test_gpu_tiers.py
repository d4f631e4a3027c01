"""The host-memory and disk tiers of the f2 result against the resident one.

The streamed tiers contract the same columns with the same kernels, chunk by chunk of
blocks, so every tier must hold a bit-identical tensor: f2, vpair and block sizes equal
to the resident handle's (and the resident handle is itself within 1e-12 of the oracle,
tests/test_gpu_decode_f2.py).  Chunks of a few blocks, chunks made only of empty blocks,
a last partial chunk, the f2.bin the disk tier writes (read back with the STPF2BK1
reader), the genotype-file path, the automatic choice and the errors are covered.
"""
import os
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (first import on a fresh box is slow; keep it out of test timeouts)

import steppe_amd as st
from steppe_amd import _core, synth
from genofiles import synthetic_triple

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _same(a, b, what):
    fa, fb = a._f2_numpy(), b._f2_numpy()
    assert np.array_equal(fa.view(np.int64), fb.view(np.int64)), what
    assert np.array_equal(a._vpair_numpy(), b._vpair_numpy()), what
    assert a.block_sizes == b.block_sizes, what


@pytest.fixture(scope="module")
def panel():
    # ragged blocks with two runs of empty blocks (ids 3..6 and 11), so some chunks hold
    # no SNP at all
    p = synth.scaled_config(P=23, M=6_001, seed=21, block_lo=60, block_hi=900)
    sizes = np.bincount(p.block_id, minlength=p.n_block).tolist()
    sizes = sizes[:3] + [0, 0, 0, 0] + sizes[3:7] + [0] + sizes[7:]
    bid = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    assert len(bid) == p.n_snp
    return p, bid, len(sizes)


def _extract(p, bid, nb, **kw):
    return _core.run_extract_f2(p.packed, p.n_snp, p.pop_offsets, bid, nb, pops=p.pop_names,
                                ploidy="auto", precision=kw.pop("precision", "fp64"), **kw)


@pytest.mark.parametrize("precision", ["fp64", "split_bf16"])
def test_tile_tiers_bit_identical(panel, tmp_path, precision):
    p, bid, nb = panel
    res = _extract(p, bid, nb, precision=precision)
    assert res.tier == "resident"
    for chunk in (1, 2, 3, 5, nb - 1, nb, 10 * nb):
        _say(f"[{precision}] host tier, chunk_blocks={chunk} of {nb}")
        h = _extract(p, bid, nb, precision=precision, tier="host", chunk_blocks=chunk)
        assert h.tier == "host" and h.device_pointers() == (0, 0)
        _same(res, h, f"host chunk {chunk}")
    for chunk in (2, 4, nb):
        path = str(tmp_path / f"c{chunk}.bin")
        _say(f"[{precision}] disk tier, chunk_blocks={chunk} of {nb}")
        d = _extract(p, bid, nb, precision=precision, tier="disk", disk_path=path, chunk_blocks=chunk)
        assert d.tier == "disk"
        _same(res, d, f"disk chunk {chunk}")
        # the file is an STPF2BK1 f2.bin the reader takes as it is
        f2, vp, bs, _ = _core.read_f2_host(_as_dir(path, tmp_path / f"dir{chunk}"))
        assert np.array_equal(f2, res._f2_numpy()) and np.array_equal(vp, res._vpair_numpy())
        assert bs.tolist() == res.block_sizes
    assert np.all(res._vpair_numpy()[:, :, 3:7] == 0) and np.all(res._f2_numpy()[:, :, 11] == 0)


def _as_dir(bin_path, d):
    os.makedirs(d, exist_ok=True)
    os.replace(bin_path, os.path.join(d, "f2.bin"))
    return str(d)


def test_auto_and_env_chunk(panel, monkeypatch):
    p, bid, nb = panel
    res = _extract(p, bid, nb)
    auto = _extract(p, bid, nb, tier="auto")
    assert auto.tier == "resident"
    _same(res, auto, "auto")
    monkeypatch.setenv("STEPPE_F2_CHUNK_BLOCKS", "4")
    h = _extract(p, bid, nb, tier="host")
    _same(res, h, "env chunk")


def test_streamed_tiers_refuse_fits(panel, tmp_path):
    p, _, _ = panel
    h = _extract(p, p.block_id, p.n_block, tier="host", chunk_blocks=3)
    names = p.pop_names
    with pytest.raises(ValueError, match="resident"):
        st.f4(h, [(names[0], names[1], names[2], names[3])])
    with pytest.raises(ValueError, match="resident"):
        st.qpadm(h, names[0], names[1:3], names[3:8])
    # the downloaded tensor uploads into a resident handle that fits as usual
    r = _core.upload_f2(h._f2_numpy(), h._vpair_numpy(), h.block_sizes, h.pops)
    assert r.tier == "resident"
    assert len(st.f4(r, [(names[0], names[1], names[2], names[3])])) == 1


def test_tier_errors(panel):
    p, bid, nb = panel
    with pytest.raises(ValueError, match="disk_path"):
        _extract(p, bid, nb, tier="disk")
    with pytest.raises(ValueError, match="tier"):
        _extract(p, bid, nb, tier="nvme")
    with pytest.raises(ValueError, match="cannot open"):
        _extract(p, bid, nb, tier="disk", disk_path="/nonexistent_dir/x/f2.bin")


def test_file_path_tiers(tmp_path):
    pre = str(tmp_path / "t")
    synthetic_triple(pre, 41, n_snp=3000, fmt="geno")
    paths = (pre + ".geno", pre + ".snp", pre + ".ind")
    kw = dict(pops=["Pop_A", "Pop_C", "Pop_E", "Pop_H", "Pop_I"], maxmiss=1.0, precision="fp64")
    res, info = _core.extract_f2_files(*paths, tier="resident", **kw)
    _say(f"file path: P={res.P} blocks {res.n_block} kept {info['n_snp_kept']}")
    host, _ = _core.extract_f2_files(*paths, tier="host", chunk_blocks=2, **kw)
    disk, _ = _core.extract_f2_files(*paths, tier="disk", disk_path=str(tmp_path / "f.bin"), chunk_blocks=3, **kw)
    assert host.tier == "host" and disk.tier == "disk" and host.pops == res.pops == disk.pops
    _same(res, host, "file host")
    _same(res, disk, "file disk")
