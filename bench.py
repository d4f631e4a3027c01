"""Benchmark: the f2 block-tensor extraction on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c3] [--precision fp64|split_bf16]

One step = one pass of the hot path over the synthetic panel already resident in
HBM: decode the packed 2-bit genotypes to per-population counts, then build the
per-jackknife-block f2 and vpair tensors [n_block x P x P] in HBM.  Prints one
JSON line.  The metric is f2 pop-pair.SNPs/s: W / t with W = P(P+1)/2 * M
(unordered population pairs incl. the diagonal, times SNPs).

Workloads (synthetic, seeded; population frequencies Balding-Nichols, genotypes
Binomial(2,p), 40% pseudo-haploid populations with Beta(2,2) missingness):
  c2: 200 populations (sizes U{1..60}), 1.2M SNPs, blocks U{200..3000} SNPs
  c3: 1000 populations (sizes U{1..60}), 1.2M SNPs, blocks LogNormal(ln 1500, 0.6)
      clipped to [50, 8000]

Extra JSON fields: ``roofline`` for the dominant kernel (the f2 contraction),
timed with HIP events on the engine's own stream; ``stages`` with every stage's
average device time; ``cpu_baseline``: the reference CpuBackend (oracle/_ref,
compiled from the reference sources) or the restated oracle, single-threaded, on
a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 20261202
MI355X_PEAK = {"hbm_GBs": 8000.0, "fp64_TFLOPs": 78.6, "bf16_TFLOPs": 2500.0}
METRIC = "f2 pop-pair·SNPs/sec at ≤1e-9 rel err (split-BF16 vs FP64); qpAdm fits/sec"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_params(name: str):
    rng = np.random.default_rng(SEED)
    M = 1_200_000
    if name == "c2":
        P = 200
    elif name == "c3":
        P = 1000
    else:
        raise SystemExit(f"unknown workload {name}")
    sizes = rng.integers(1, 61, P)
    haploid = rng.random(P) < 0.4
    missing = np.where(haploid, rng.beta(2.0, 2.0, P), 0.05).astype(np.float32)
    fst = rng.uniform(0.02, 0.15, P).astype(np.float32)
    from steppe_amd import synth
    if name == "c2":
        bs = synth.uniform_block_sizes(M, 200, 3000, rng)
    else:
        bs = synth.lognormal_block_sizes(M, 1500.0, 0.6, 50, 8000, rng)
    return dict(P=P, M=M, sizes=sizes, haploid=haploid, missing=missing, fst=fst, block_sizes=bs)


def build_panel(ctx, wp):
    """Generate the packed panel directly in HBM (returns the torch buffer)."""
    import torch
    from steppe_amd import _core
    P, M = wp["P"], wp["M"]
    sizes = wp["sizes"].astype(np.int64)
    n_ind = int(sizes.sum())
    bpr = ((M + 3) // 4 + 255) // 256 * 256
    ind_pop = np.repeat(np.arange(P, dtype=np.int32), sizes)
    ind_hap = np.repeat(wp["haploid"].astype(np.int8), sizes)
    packed = torch.empty(n_ind * bpr, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    _core.check(_core.lib().stpx_synth_panel(
        ctx.h, SEED, P, M, n_ind, bpr, _core._ptr(np.ascontiguousarray(wp["fst"])),
        _core._ptr(np.ascontiguousarray(wp["missing"])), _core._ptr(ind_pop), _core._ptr(ind_hap),
        C.c_void_p(packed.data_ptr())), "synth")
    offs = np.zeros(P + 1, np.uint64)
    offs[1:] = np.cumsum(sizes)
    return packed, n_ind, bpr, offs


def cpu_baseline(packed, n_ind, bpr, offs, wp, target_s: float = 15.0):
    """Single-thread reference CpuBackend (or the restated oracle) on a sample:
    all populations x the leading blocks totalling ~target_s seconds of work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    from steppe_amd import synth
    lib = oracle_lib.ref()
    kind = "reference"
    if lib is None:
        lib, kind = oracle_lib.cref(), "port"
    P = wp["P"]
    bs = wp["block_sizes"]
    host = packed.view(n_ind, bpr)

    def sample(n_blocks):
        m = int(bs[:n_blocks].sum())
        nbytes = (m + 3) // 4
        sub = host[:, :nbytes].cpu().numpy()
        panel = synth.Panel(packed=np.ascontiguousarray(sub), n_snp=m, pop_offsets=offs,
                            sample_ploidy=np.zeros(n_ind, np.int32), block_id=synth.blocks_from_sizes(bs[:n_blocks]),
                            n_block=n_blocks)
        q, v, n = oracle_lib.cref().decode_af(panel, "detect")
        return m, q, v, n, panel.block_id

    m, q, v, n, bid = sample(1)
    t0 = time.perf_counter()
    lib.f2_blocks(q, v, n, bid, 1)
    t1 = time.perf_counter() - t0
    rate = P * (P + 1) / 2 * m / max(t1, 1e-6)
    need = target_s * rate / (P * (P + 1) / 2)
    k = int(np.searchsorted(np.cumsum(bs), need)) + 1
    k = max(1, min(k, len(bs)))
    m, q, v, n, bid = sample(k)
    t0 = time.perf_counter()
    lib.f2_blocks(q, v, n, bid, k)
    dt = time.perf_counter() - t0
    value = P * (P + 1) / 2 * m / dt
    return {"value": value, "unit": "pop-pair·SNPs/s", "cores": 1, "kind": kind,
            "sample": f"all {P} populations x first {k} blocks ({m} SNPs), f2 blocks from decoded "
                      f"Q/V/N, {dt:.1f} s single-threaded ({os.path.basename(lib.path)})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c2", "c3"])
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split_bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding is "
                                   "design-only (block-range shards), --gpus must be 1"}))
        sys.exit(2)

    import torch
    from steppe_amd import _core

    ctx = _core.Context.get(0)
    arch, n_cu = ctx.info()
    wp = workload_params(args.workload)
    P, M = wp["P"], wp["M"]
    bs = wp["block_sizes"]
    nb = len(bs)
    block_id = np.ascontiguousarray(np.repeat(np.arange(nb, dtype=np.int32), bs))
    log(f"[bench] {args.workload}: P={P} M={M} n_block={nb} on {arch} ({n_cu} CUs)")
    t0 = time.perf_counter()
    packed, n_ind, bpr, offs = build_panel(ctx, wp)
    log(f"[bench] panel in HBM: {n_ind} individuals x {bpr} B ({packed.numel()/1e9:.2f} GB) "
        f"in {time.perf_counter()-t0:.1f} s")

    prec = _core.parse_precision(args.precision)
    ld = (M + 63) // 64 * 64
    a2 = torch.empty(P * ld, dtype=torch.int16, device="cuda")
    nn = torch.empty(P * ld, dtype=torch.int16, device="cuda")
    f2 = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    vp = torch.empty(P * P * nb, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    tile, keep = _core.make_tile(None, M, offs, None, detect=True, bytes_per_record=bpr,
                                 n_individuals=n_ind, device_ptr=packed.data_ptr())
    L = _core.lib()
    bid_p = _core._ptr(block_id)

    def step():
        _core.check(L.stp_decode_counts(ctx.h, C.byref(tile), C.c_void_p(a2.data_ptr()),
                                        C.c_void_p(nn.data_ptr()), ld), "decode")
        _core.check(L.stp_f2_blocks_from_counts(ctx.h, C.c_void_p(a2.data_ptr()), C.c_void_p(nn.data_ptr()),
                                                ld, P, M, bid_p, nb, prec, C.c_void_p(f2.data_ptr()),
                                                C.c_void_p(vp.data_ptr())), "f2")

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i+1}/{args.warmup} done")
    ctx.timing(True)
    ctx.read_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = ctx.read_timing(reset=True)
    ctx.timing(False)
    log(f"[bench] {args.steps} steps in {dt:.3f} s; stages {tm}")

    W = P * (P + 1) / 2 * M
    value = W * args.steps / dt
    ms_step = dt / args.steps * 1e3
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in tm.items()}
    # Algorithmic work of the contraction: 4 MACs per (pair, SNP) = 8 W FLOP.
    f_alg = 8.0 * W
    f2_s = avg["f2"] / 1e3
    peak = MI355X_PEAK["fp64_TFLOPs"] if prec == _core.STP_PREC_FP64 else MI355X_PEAK["bf16_TFLOPs"]
    achieved = f_alg / f2_s / 1e12 if f2_s > 0 else 0.0
    roofline = {"kernel": "f2_fp64_kernel" if prec == _core.STP_PREC_FP64 else "f2_split_kernel",
                "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "work": "8*W FLOP per launch, W = P(P+1)/2*M"}
    dec_bytes = n_ind * bpr + 4 * P * M
    stages = {k: {"avg_ms": avg[k]} for k in avg}
    stages["decode"]["hbm_GBs"] = dec_bytes / (avg["decode"] / 1e3) / 1e9 if avg["decode"] else None
    stages["decode"]["algorithmic_bytes"] = dec_bytes

    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(packed, n_ind, bpr, offs, wp)
        log(f"[bench] cpu baseline {cpu}")

    out = {
        "metric": METRIC, "value": value, "unit": "pop-pair·SNPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == _core.STP_PREC_FP64 else "bf16x(split)->f64",
        "data": "synthetic (Balding-Nichols panel generated in HBM)",
        "config": {"workload": f"{args.workload}: {P} populations x {M} SNPs, {n_ind} individuals, "
                               f"{nb} jackknife blocks, decode + f2 ({args.precision})",
                   "P": P, "M": M, "n_block": nb, "n_individuals": n_ind,
                   "precision": args.precision, "parallelism": "single GPU"},
        "roofline": roofline, "stages": stages, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
