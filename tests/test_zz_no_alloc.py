"""stp_decode_counts_noalloc (the decode entry point that allocates no device memory)
against stp_decode_counts, through the C ABI.

Two calls in a row on one GPU, the second with a smaller panel (fewer populations,
individuals and SNPs), in each ploidy mode (detected on the device, given per sample,
fixed): counts bit-identical to the allocating entry point's.  Host-side rejections:
a host-resident tile and inputs beyond the scratch limits.
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


def _decode(ctx, tile, P, M, noalloc):
    ld = (M + 63) // 64 * 64
    a2 = torch.full((P * ld,), -1, dtype=torch.int16, device="cuda")
    n = torch.full((P * ld,), -1, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    L = _core.lib()
    fn = L.stp_decode_counts_noalloc if noalloc else L.stp_decode_counts
    _core.check(fn(ctx.h, C.byref(tile), C.c_void_p(a2.data_ptr()), C.c_void_p(n.data_ptr()), ld),
                "decode (no allocation)" if noalloc else "decode")
    v = lambda t: t.view(P, ld)[:, :M].cpu()
    return v(a2), v(n)


def _tile(panel, mode):
    packed = torch.from_numpy(np.ascontiguousarray(panel.packed)).cuda()
    kw = dict(bytes_per_record=panel.packed.shape[1], n_individuals=panel.packed.shape[0],
              device_ptr=packed.data_ptr())
    if mode == "detect":
        tile, keep = _core.make_tile(None, panel.n_snp, panel.pop_offsets, None, detect=True, **kw)
    elif mode == "sample":
        tile, keep = _core.make_tile(None, panel.n_snp, panel.pop_offsets, panel.sample_ploidy, **kw)
    else:
        tile, keep = _core.make_tile(None, panel.n_snp, panel.pop_offsets, None, ploidy=2, **kw)
    keep.append(packed)
    return tile, keep


@pytest.mark.parametrize("mode", ["detect", "sample", "fixed"])
def test_noalloc_decode_equals_decode(engine, mode):
    big = synth.scaled_config(P=41, M=30_000, seed=71, block_lo=200, block_hi=3000)
    small = synth.scaled_config(P=7, M=9_000, seed=72, block_lo=200, block_hi=3000)
    for name, panel in (("first call", big), ("second call, smaller", small)):
        tile, keep = _tile(panel, mode)
        _say(f"[{mode}] {name}: P={panel.n_pop} M={panel.n_snp} individuals={panel.packed.shape[0]}")
        a2_na, n_na = _decode(engine, tile, panel.n_pop, panel.n_snp, True)
        a2, n = _decode(engine, tile, panel.n_pop, panel.n_snp, False)
        assert torch.equal(a2_na, a2) and torch.equal(n_na, n), f"{mode}: {name} differs from stp_decode_counts"


def test_noalloc_decode_rejections(engine):
    panel = synth.config1()
    L = _core.lib()
    ld = (panel.n_snp + 63) // 64 * 64
    buf = torch.empty(panel.n_pop * ld, dtype=torch.int16, device="cuda")
    host_tile, keep = _core.make_tile(panel.packed, panel.n_snp, panel.pop_offsets, panel.sample_ploidy)
    rc = L.stp_decode_counts_noalloc(engine.h, C.byref(host_tile), C.c_void_p(buf.data_ptr()),
                                     C.c_void_p(buf.data_ptr()), ld)
    assert rc != 0 and b"device" in L.stp_last_error()
    # beyond the scratch: more individuals than it holds (checked before any device work)
    tile, keep2 = _tile(panel, "detect")
    tile.n_individuals = (1 << 20) + 1
    rc = L.stp_decode_counts_noalloc(engine.h, C.byref(tile), C.c_void_p(buf.data_ptr()),
                                     C.c_void_p(buf.data_ptr()), ld)
    assert rc != 0 and b"individuals" in L.stp_last_error()
