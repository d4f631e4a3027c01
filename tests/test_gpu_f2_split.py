"""GPU parity of the split-BF16 f2 form (fixed-point BF16 slices on BF16 MFMA).

Bar: vpair bit-exact; f2 within 1e-9 relative of the FP64 oracle with an absolute
floor of 1e-6 f2 units on |r| (|c - r| <= 1e-9 * max(|r|, 1e-6)) — the tolerance
the split form is specified to; the measured error is printed and is typically
orders of magnitude below it.
"""
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (keep the first, slow import out of test timeouts)

from steppe_amd import _core, synth
from test_gpu_decode_f2 import F2_FLOOR_TINY_BLOCKS, PANELS, _f2_qvn_gpu, _rel_err

pytestmark = pytest.mark.gpu

SPLIT_RTOL = 1e-9
FLOOR = 1e-6


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=sorted(PANELS))
def panel(request):
    _say(f"[panel {request.param}]")
    return request.param, PANELS[request.param]()


def test_split_matches_oracle(engine, cref, panel):
    name, p = panel
    q, v, n = cref.decode_af(p, "sample")
    _say(f"f2 split {name} P={p.n_pop} M={p.n_snp} nb={p.n_block}")
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block, _core.STP_PREC_SPLIT_BF16)
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block, n_threads=8)
    assert np.array_equal(bs, br)
    assert np.array_equal(vp, vr)
    err = _rel_err(f2, fr)
    _say(f"  split vs oracle: max rel err (floor {FLOOR}) = {err:.3e}")
    assert err <= SPLIT_RTOL
    assert np.array_equal(f2, f2.transpose(1, 0, 2))
    f64, _, _ = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block, _core.STP_PREC_FP64)
    _say(f"  split vs fp64 gpu: {_rel_err(f2, f64):.3e}")
    assert _rel_err(f2, f64) <= SPLIT_RTOL


@pytest.mark.parametrize("P", [1, 2, 31, 32, 33, 127, 128, 129, 260])
def test_split_population_counts_edge_cases(engine, cref, P):
    p = synth.scaled_config(P=P, M=900, seed=200 + P, block_lo=1, block_hi=300)
    q, v, n = cref.decode_af(p, "sample")
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block, _core.STP_PREC_SPLIT_BF16)
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block)
    assert np.array_equal(vp, vr) and np.array_equal(bs, br)
    assert _rel_err(f2, fr, F2_FLOOR_TINY_BLOCKS) <= SPLIT_RTOL


def test_split_long_blocks_and_extreme_frequencies(engine, cref):
    # 8000-SNP blocks (the largest the workloads allow) with frequencies pinned at
    # 0 and 1 in half the populations: the carry chain and the level bounds at
    # their worst case.
    rng = np.random.default_rng(5)
    P, M = 40, 16000
    q = rng.random((P, M))
    q[: P // 2] = (q[: P // 2] > 0.5).astype(np.float64)
    n = rng.integers(1, 120, (P, M)).astype(np.float64)
    q = np.round(q * 2 * n) / (2 * n)          # frequencies of the form a2 / (2n)
    v = (rng.random((P, M)) > 0.1).astype(np.float64)
    q, n = q * v, n * v
    bid = np.repeat(np.arange(2, dtype=np.int32), 8000)
    f2, vp, bs = _f2_qvn_gpu(engine, q, v, n, bid, 2, _core.STP_PREC_SPLIT_BF16)
    fr, vr, br = cref.f2_blocks(q, v, n, bid, 2, n_threads=8)
    assert np.array_equal(vp, vr)
    err = _rel_err(f2, fr)
    _say(f"  extreme: max rel err = {err:.3e}")
    assert err <= SPLIT_RTOL


def test_split_extract_end_to_end(engine, cref):
    p = synth.config1()
    h = _core.run_extract_f2(p.packed, p.n_snp, p.pop_offsets, p.block_id, p.n_block,
                             pops=p.pop_names, ploidy="auto", precision="split_bf16")
    q, v, n = cref.decode_af(p, "detect")
    fr, vr, br = cref.f2_blocks(q, v, n, p.block_id, p.n_block)
    assert np.array_equal(h._vpair_numpy(), vr)
    assert _rel_err(h._f2_numpy(), fr) <= SPLIT_RTOL


def _near_zero_panel(seed=11):
    """Q/V/N with exactly-zero and near-zero block f2 entries: populations 1..5 copy the
    single pseudo-haploid individual of population 0 (h = 0, so f2 = 0 on every block),
    population 6 differs from it at one SNP per block, the rest are random."""
    rng = np.random.default_rng(seed)
    P, M = 48, 4096
    n = rng.integers(1, 120, (P, M)).astype(np.float64)
    q = np.round(rng.random((P, M)) * 2 * n) / (2 * n)
    n[:7] = 1.0
    q[0] = (rng.random(M) > 0.5).astype(np.float64)
    q[1:7] = q[0]
    q[6, ::512] = 1.0 - q[0, ::512]
    v = (rng.random((P, M)) > 0.05).astype(np.float64)
    v[:7] = 1.0
    q, n = q * v, n * v
    bid = np.repeat(np.arange(8, dtype=np.int32), 512)
    return q, v, n, bid


@pytest.mark.parametrize("prec", [_core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16])
def test_cancellation_prone_entries_recomputed(engine, cref, prec):
    # Entries whose terms cancel to (near) zero are recomputed in FP64 by the fix-up
    # kernel: they match the reference's long-double sum to its rounding, far inside the
    # FP64 path's 1e-12 bar even where the exact value is 0.
    q, v, n, bid = _near_zero_panel()
    f2, vp, _ = _f2_qvn_gpu(engine, q, v, n, bid, 8, prec)
    fr, vr, _ = cref.f2_blocks(q, v, n, bid, 8)
    assert np.array_equal(vp, vr)
    assert np.all(fr[0, 1:6, :] == 0.0) and np.all(f2[0, 1:6, :] == 0.0)
    small = np.abs(fr) < 1e-4
    assert small.sum() >= 5 * 8
    err = float(np.max(np.abs(f2 - fr)[small] / np.maximum(np.abs(fr[small]), 1e-6)))
    _say(f"  {_core.PRECISION_NAMES[prec]}: {int(small.sum())} entries below 1e-4, max rel err {err:.2e}")
    assert err <= 1e-12
    assert _rel_err(f2, fr) <= (1e-12 if prec == _core.STP_PREC_FP64 else SPLIT_RTOL)


@pytest.mark.parametrize("prec", [_core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16])
def test_fixup_list_overflow_scan(engine, monkeypatch, prec):
    # With room for a single flagged entry the list overflows and the scan kernel
    # recomputes every entry below the threshold: the same tensor as the list path.
    q, v, n, bid = _near_zero_panel(seed=12)
    f2_list, vp_list, _ = _f2_qvn_gpu(engine, q, v, n, bid, 8, prec)
    monkeypatch.setenv("STEPPE_F2_FIXUP_CAP", "1")
    f2_scan, vp_scan, _ = _f2_qvn_gpu(engine, q, v, n, bid, 8, prec)
    assert np.array_equal(vp_list, vp_scan)
    assert np.array_equal(f2_list, f2_scan)


@pytest.mark.parametrize("shape", [(1, 900, 1, 300), (33, 900, 1, 300), (129, 900, 1, 300), (260, 900, 1, 300),
                                   (70, 20_000, 2_000, 5_000)])
def test_split_ring_layouts_bit_identical(engine, cref, monkeypatch, shape):
    # The contraction's two LDS ring layouts (a tuner dimension) on uneven population
    # counts, 1-300-SNP blocks (odd numbers of 16-SNP steps) and long blocks (many carry
    # intervals): the same tensors bit for bit.
    P, M, lo, hi = shape
    p = synth.scaled_config(P=P, M=M, seed=300 + P, block_lo=lo, block_hi=hi)
    q, v, n = cref.decode_af(p, "sample")
    out = {}
    for ring in (0, 1):
        monkeypatch.setenv("STEPPE_F2_TUNE",
                           f"sched=0,fp64_rows=64,fp64_prefetch=1,fp64_wg=2,count_wg=2,split_ring={ring}")
        out[ring] = _f2_qvn_gpu(engine, q, v, n, p.block_id, p.n_block, _core.STP_PREC_SPLIT_BF16)
    monkeypatch.delenv("STEPPE_F2_TUNE")
    _say(f"  P={P} M={M} blocks {lo}-{hi}: ring layouts compared over {out[0][0].size} entries")
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    fr, _, _ = cref.f2_blocks(q, v, n, p.block_id, p.n_block)
    assert _rel_err(out[1][0], fr, F2_FLOOR_TINY_BLOCKS) <= SPLIT_RTOL
