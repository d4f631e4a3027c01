"""The captured f2 extraction (stp_f2_graph_*) against the eager calls.

A graph replays ploidy detection, decode, prep, pair counts, contraction and fix-up.
Every replay must rewrite exactly the tensors that stp_decode_counts +
stp_f2_blocks_from_counts produce on the same inputs (bit-identical: the same kernels
on the same data), also after the outputs were cleared, after eager calls of another
shape ran in between (which reuse and grow the backend's own workspace), and with
stage timing on.  Replays are checked one by one, with eager work and synchronisation
between them.
"""
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from steppe_amd import _core, synth

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _device_tile(panel, detect):
    packed = torch.from_numpy(np.ascontiguousarray(panel.packed)).cuda()
    tile, keep = _core.make_tile(None, panel.n_snp, panel.pop_offsets, None if detect else panel.sample_ploidy,
                                 detect=detect, bytes_per_record=panel.packed.shape[1],
                                 n_individuals=panel.packed.shape[0], device_ptr=packed.data_ptr())
    keep.append(packed)
    return tile, keep


def _eager(ctx, tile, panel, prec):
    P, M, nb = panel.n_pop, panel.n_snp, panel.n_block
    ld = (M + 63) // 64 * 64
    a2 = torch.empty(P * ld, dtype=torch.int16, device="cuda")
    n = torch.empty_like(a2)
    f2 = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    vp = torch.empty_like(f2)
    torch.cuda.synchronize()
    L = _core.lib()
    _core.check(L.stp_decode_counts(ctx.h, C.byref(tile), C.c_void_p(a2.data_ptr()), C.c_void_p(n.data_ptr()), ld))
    _core.check(L.stp_f2_blocks_from_counts(ctx.h, C.c_void_p(a2.data_ptr()), C.c_void_p(n.data_ptr()), ld, P, M,
                                            _core._ptr(np.ascontiguousarray(panel.block_id, np.int32)), nb, prec,
                                            C.c_void_p(f2.data_ptr()), C.c_void_p(vp.data_ptr())))
    return f2.cpu(), vp.cpu(), a2.cpu(), n.cpu()


@pytest.mark.parametrize("prec", [_core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16])
@pytest.mark.parametrize("detect", [True, False])
def test_graph_replays_equal_eager(engine, prec, detect):
    panel = synth.scaled_config(P=37, M=20_000, seed=91, block_lo=50, block_hi=3000)
    tile, keep = _device_tile(panel, detect)
    f2_ref, vp_ref, _, _ = _eager(engine, tile, panel, prec)
    P, nb = panel.n_pop, panel.n_block
    f2 = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    vp = torch.empty_like(f2)
    torch.cuda.synchronize()
    g = _core.F2Graph(engine, tile, panel.block_id, nb, prec, f2.data_ptr(), vp.data_ptr())
    assert torch.equal(f2.cpu(), f2_ref) and torch.equal(vp.cpu(), vp_ref), "eager pass of the graph"
    other = synth.scaled_config(P=50, M=30_000, seed=92, block_lo=50, block_hi=3000)
    other_tile, other_keep = _device_tile(other, True)
    for r in range(4):
        f2.zero_()
        vp.fill_(-1.0)
        torch.cuda.synchronize()
        g.launch()
        same = torch.equal(f2.cpu(), f2_ref) and torch.equal(vp.cpu(), vp_ref)
        _say(f"  replay {r}: {'equal' if same else 'DIFFERENT'}")
        assert same, f"replay {r} differs from the eager extraction"
        # eager work of another shape between replays (grows the backend's workspace)
        _eager(engine, other_tile, other, prec)
    g.close()


def test_graph_own_counts_and_timing(engine):
    panel = synth.config1()
    tile, keep = _device_tile(panel, True)
    prec = _core.STP_PREC_SPLIT_BF16
    f2_ref, vp_ref, _, _ = _eager(engine, tile, panel, prec)
    P, nb = panel.n_pop, panel.n_block
    f2 = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    vp = torch.empty_like(f2)
    torch.cuda.synchronize()
    engine.timing(True)
    try:
        g = _core.F2Graph(engine, tile, panel.block_id, nb, prec, f2.data_ptr(), vp.data_ptr())
        engine.read_timing(reset=True)
        for _ in range(3):
            g.launch()
        t = engine.read_timing(reset=True)
        _say(f"  stage timing over 3 replays: {t}")
        assert torch.equal(f2.cpu(), f2_ref) and torch.equal(vp.cpu(), vp_ref)
        # either the stage events were captured (3 timed replays) or none were
        assert t["f2"][1] in (0, 3)
    finally:
        engine.timing(False)


def test_graph_rejects_host_tile(engine):
    panel = synth.config1()
    tile, keep = _core.make_tile(panel.packed, panel.n_snp, panel.pop_offsets, panel.sample_ploidy)
    f2 = torch.empty(8 * 8 * panel.n_block, dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="resident"):
        _core.F2Graph(engine, tile, panel.block_id, panel.n_block, "fp64", f2.data_ptr(), f2.data_ptr())
