"""The f2 autotuner (stp_f2_autotune) and the launch configurations it chooses between.

Every configuration must write the same tensors (the same per-entry sums in the same
order, whatever the schedule, wavefront tile, prefetch or grid): each candidate, forced
with STEPPE_F2_TUNE, is compared bit for bit with the default configuration.  The tuner
returns a table with one timing per candidate, leaves the winner's tensors in the outputs,
writes its choice to the tune cache file, and later calls of the same workload (also a
fresh context reading the file) use it.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest
import torch

from steppe_amd import _core, synth

pytestmark = pytest.mark.gpu

SPLIT_CANDIDATES = [f"sched={s},fp64_rows=64,fp64_prefetch=1,fp64_wg=2,count_wg={c},split_ring={g}"
                    for g in (0, 1) for s in (0, 1) for c in (1, 2, 4)]
FP64_CANDIDATES = [f"sched={s},fp64_rows={r},fp64_prefetch={p},fp64_wg={w},count_wg=2,split_ring=0"
                   for s in (0, 1) for r in (64, 32) for p in (1, 0) for w in (1, 2, 3) if not (r == 64 and w == 3)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def counts(engine):
    panel = synth.scaled_config(P=150, M=40_000, seed=303, block_lo=50, block_hi=3000)
    P, M = panel.n_pop, panel.n_snp
    ld = (M + 63) // 64 * 64
    a2 = torch.empty(P * ld, dtype=torch.int16, device="cuda")
    n = torch.empty_like(a2)
    tile, keep = _core.tile_from_panel(panel)
    torch.cuda.synchronize()
    _core.check(_core.lib().stp_decode_counts(engine.h, C.byref(tile), C.c_void_p(a2.data_ptr()),
                                              C.c_void_p(n.data_ptr()), ld))
    return panel, a2, n, ld


def _f2(engine, counts, prec):
    panel, a2, n, ld = counts
    P, nb = panel.n_pop, panel.n_block
    f2 = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    vp = torch.empty_like(f2)
    torch.cuda.synchronize()
    _core.check(_core.lib().stp_f2_blocks_from_counts(
        engine.h, C.c_void_p(a2.data_ptr()), C.c_void_p(n.data_ptr()), ld, P, panel.n_snp,
        _core._ptr(np.ascontiguousarray(panel.block_id, np.int32)), nb, prec, C.c_void_p(f2.data_ptr()),
        C.c_void_p(vp.data_ptr())))
    return f2.cpu(), vp.cpu()


@pytest.mark.parametrize("prec", [_core.STP_PREC_SPLIT_BF16, _core.STP_PREC_FP64])
def test_every_configuration_same_tensor(engine, counts, monkeypatch, prec):
    monkeypatch.setenv("STEPPE_TUNE_CACHE", "/nonexistent-dir-for-tests/f2_tune.tsv")
    ref = _f2(engine, counts, prec)
    cands = SPLIT_CANDIDATES if prec == _core.STP_PREC_SPLIT_BF16 else FP64_CANDIDATES
    for c in cands:
        monkeypatch.setenv("STEPPE_F2_TUNE", c)
        got = _f2(engine, counts, prec)
        same = torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
        _say(f"  {c}: {'same' if same else 'DIFFERENT'}")
        assert same, c


@pytest.mark.parametrize("prec", [_core.STP_PREC_SPLIT_BF16, _core.STP_PREC_FP64])
def test_autotune_chooses_persists_and_is_used(engine, counts, tmp_path, monkeypatch, prec):
    cache = tmp_path / "f2_tune.tsv"
    monkeypatch.setenv("STEPPE_TUNE_CACHE", str(cache))
    monkeypatch.delenv("STEPPE_F2_TUNE", raising=False)
    panel, a2, n, ld = counts
    ref = _f2(engine, counts, prec)
    P, nb = panel.n_pop, panel.n_block
    f2 = torch.full((P * P * nb,), -7.0, dtype=torch.float64, device="cuda")
    vp = torch.full_like(f2, -7.0)
    torch.cuda.synchronize()
    rep = _core.f2_autotune(engine, a2.data_ptr(), n.data_ptr(), ld, P, panel.n_snp, panel.block_id, nb, prec,
                            f2.data_ptr(), vp.data_ptr())
    cands = SPLIT_CANDIDATES if prec == _core.STP_PREC_SPLIT_BF16 else FP64_CANDIDATES
    table = ", ".join(f"{c['config']}: {c['ms']:.3f} ms" for c in rep["candidates"])
    _say(f"  key {rep['key']}; chosen {rep['chosen']}; {table}")
    assert sorted(c["config"] for c in rep["candidates"]) == sorted(cands)
    assert all(c["ms"] > 0 for c in rep["candidates"])
    win = rep["candidates"][rep["chosen"]]
    assert win["ms"] == min(c["ms"] for c in rep["candidates"])
    # the winner's tensors are left in the outputs
    assert torch.equal(f2.cpu(), ref[0]) and torch.equal(vp.cpu(), ref[1])
    # the cache file holds the choice under the workload's key
    lines = [l.split("\t") for l in cache.read_text().splitlines()]
    assert [rep["key"], win["config"]] in lines
    # a fresh context reads it; its f2 of the same workload is unchanged
    ctx2 = _core.Context(0)
    try:
        got = _f2(ctx2, counts, prec)
    finally:
        _core.lib().stp_close(ctx2.h)
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
