"""GPU parity: the HIP decode and f2 paths against the oracle, through the C ABI.

Bar: integer counts, ploidy, q/v/n and vpair bit-exact; f2 (FP64 form) within
1e-12 relative of the oracle with an absolute floor of 1e-6 f2 units on |r|
(|c - r| <= 1e-12 * max(|r|, 1e-6)).
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (first import on a fresh box is slow; keep it out of test timeouts)

from steppe_amd import _core, synth

pytestmark = pytest.mark.gpu

F2_RTOL = 1e-12
F2_FLOOR = 1e-6
# Blocks of a handful of SNPs make single-SNP terms near zero, where any FP64
# evaluation order (the oracle's (q_i-q_j)^2 - h_i - h_j or the expanded form)
# differs by a few ulps of the summands: those stress tests use a floor of 1e-4
# f2 units (about 1% of a typical per-block f2).
F2_FLOOR_TINY_BLOCKS = 1e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel_err(c, r, floor=F2_FLOOR):
    return float(np.max(np.abs(c - r) / np.maximum(np.abs(r), floor))) if c.size else 0.0


def _decode_counts_gpu(engine, tile, P, M):
    import torch
    ld = ((M + 63) // 64) * 64
    a2 = torch.zeros(P * ld, dtype=torch.int16, device="cuda")
    n = torch.zeros(P * ld, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    _core.check(_core.lib().stp_decode_counts(engine.h, C.byref(tile), C.c_void_p(a2.data_ptr()),
                                              C.c_void_p(n.data_ptr()), ld))
    a2h = a2.cpu().numpy().view(np.uint16).reshape(P, ld)[:, :M].astype(np.int64)
    nh = n.cpu().numpy().view(np.uint16).reshape(P, ld)[:, :M].astype(np.int64)
    return a2h, nh


def _decode_af_gpu(engine, tile, P, M):
    q = np.empty(P * M)
    v = np.empty(P * M)
    n = np.empty(P * M)
    _core.check(_core.lib().stp_decode_af(engine.h, C.byref(tile), _core._ptr(q), _core._ptr(v),
                                          _core._ptr(n)))
    return q.reshape(M, P).T, v.reshape(M, P).T, n.reshape(M, P).T


def _f2_qvn_gpu(engine, q, v, n, block_id, nb, precision=_core.STP_PREC_FP64):
    P, M = q.shape
    Q = np.ascontiguousarray(q.T)
    V = np.ascontiguousarray(v.T)
    N = np.ascontiguousarray(n.T)
    bid = np.ascontiguousarray(block_id, np.int32)
    f2 = np.empty(P * P * nb)
    vp = np.empty(P * P * nb)
    bs = np.empty(nb, np.int32)
    _core.check(_core.lib().stp_compute_f2_blocks(engine.h, _core._ptr(Q), _core._ptr(V), _core._ptr(N),
                                                  P, M, _core._ptr(bid), nb, precision, _core._ptr(f2),
                                                  _core._ptr(vp), _core._ptr(bs)))
    return (f2.reshape((P, P, nb), order="F"), vp.reshape((P, P, nb), order="F"), bs)


PANELS = {
    "config1": lambda: synth.config1(),
    "ragged": lambda: synth.scaled_config(P=37, M=20_011, seed=3, block_lo=40, block_hi=2500),
    "unaligned_bpr": lambda: synth.scaled_config(P=9, M=4_001, seed=5, block_lo=100, block_hi=700,
                                                 bytes_per_record=1_003),
}


@pytest.fixture(scope="module", params=sorted(PANELS))
def panel(request):
    _say(f"[panel {request.param}]")
    return request.param, PANELS[request.param]()


def test_decode_counts_bit_exact(engine, cref, panel):
    name, p = panel
    for mode in ("sample", "auto", 2):
        tile, keep = _core.tile_from_panel(p, mode)
        _say(f"decode_counts {name} ploidy={mode}")
        a2, n = _decode_counts_gpu(engine, tile, p.n_pop, p.n_snp)
        qc, vc, nc = cref.decode_af(p, "detect" if mode == "auto" else mode)
        assert np.array_equal(n, nc.astype(np.int64))
        with np.errstate(invalid="ignore", divide="ignore"):
            q = np.where(n > 0, a2 / (2.0 * n), 0.0)
        assert np.array_equal(q.view(np.int64), np.ascontiguousarray(qc).view(np.int64))


def test_decode_af_bit_exact(engine, cref, panel):
    name, p = panel
    tile, keep = _core.tile_from_panel(p, "sample")
    q, v, n = _decode_af_gpu(engine, tile, p.n_pop, p.n_snp)
    qc, vc, nc = cref.decode_af(p, "sample")
    for a, b in ((q, qc), (v, vc), (n, nc)):
        assert np.array_equal(np.ascontiguousarray(a).view(np.int64), np.ascontiguousarray(b).view(np.int64))


def test_detect_ploidy_bit_exact(engine, cref, panel):
    name, p = panel
    tile, keep = _core.tile_from_panel(p, "auto")
    out = np.empty(p.n_ind, np.int32)
    _core.check(_core.lib().stp_detect_ploidy(engine.h, C.byref(tile), _core._ptr(out)))
    assert np.array_equal(out, cref.detect_ploidy(p))


def test_f2_fp64_matches_oracle(engine, cref, panel):
    name, p = panel
    q, v, n = cref.decode_af(p, "sample")
    _say(f"f2 fp64 {name} P={p.n_pop} M={p.n_snp} nb={p.n_block}")
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block)
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block, n_threads=8)
    assert np.array_equal(bs, br)
    assert np.array_equal(vp, vr)
    err = _rel_err(f2, fr)
    _say(f"  max rel err (floor {F2_FLOOR}) = {err:.3e}")
    assert err <= F2_RTOL
    assert np.array_equal(f2, f2.transpose(1, 0, 2))


def test_extract_end_to_end_and_f2_dir(engine, cref, panel, tmp_path):
    name, p = panel
    h = _core.run_extract_f2(p.packed, p.n_snp, p.pop_offsets, p.block_id, p.n_block,
                             pops=p.pop_names, ploidy="auto", precision="fp64")
    assert h.P == p.n_pop and h.n_block == p.n_block
    q, v, n = cref.decode_af(p, "detect")
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block, n_threads=8)
    f2, vp = h._f2_numpy(), h._vpair_numpy()
    assert np.array_equal(vp, vr) and h.block_sizes == br.tolist()
    assert _rel_err(f2, fr) <= F2_RTOL
    out = _core.write_f2_dir(str(tmp_path / name), f2, vp, h.block_sizes, h.pops)
    h2 = _core.read_f2(out)
    assert np.array_equal(h2._f2_numpy(), f2) and np.array_equal(h2._vpair_numpy(), vp)
    assert h2.pops == p.pop_names


@pytest.mark.parametrize("P", [1, 2, 16, 17, 63, 64, 65, 130])
def test_f2_population_counts_edge_cases(engine, cref, P):
    p = synth.scaled_config(P=P, M=700, seed=100 + P, block_lo=1, block_hi=200)
    q, v, n = cref.decode_af(p, "sample")
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block)
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block)
    assert np.array_equal(vp, vr) and np.array_equal(bs, br)
    assert _rel_err(f2, fr, F2_FLOOR_TINY_BLOCKS) <= F2_RTOL


def test_f2_empty_and_tiny_blocks(engine, cref):
    # block 1 and block 4 are empty (ids skipped), blocks of 1 SNP, a last block of 64+1
    p = synth.scaled_config(P=20, M=300, seed=9)
    sizes = [5, 0, 1, 64, 0, 65, 165]
    bid = np.repeat(np.arange(len(sizes), dtype=np.int32), sizes)
    q, v, n = cref.decode_af(p, "sample")
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, bid, len(sizes))
    fr, vr, br = cref.f2_blocks(q, v, n, bid, len(sizes))
    assert bs.tolist() == sizes and np.array_equal(bs, br)
    assert np.array_equal(vp, vr)
    assert np.all(f2[:, :, 1] == 0.0) and np.all(vp[:, :, 4] == 0.0)
    assert _rel_err(f2, fr, F2_FLOOR_TINY_BLOCKS) <= F2_RTOL


def test_invalid_block_partition_is_rejected(engine, cref):
    p = synth.scaled_config(P=4, M=50, seed=1)
    q, v, n = cref.decode_af(p, "sample")
    bad = np.zeros(50, np.int32)
    bad[10:] = 2
    bad[30:] = 1  # decreasing -> not contiguous
    with pytest.raises(ValueError, match="non-decreasing"):
        _f2_qvn_gpu(engine, q, v, n, bad, 3)
    with pytest.raises(ValueError, match="out of range"):
        _f2_qvn_gpu(engine, q, v, n, np.full(50, 7, np.int32), 3)
