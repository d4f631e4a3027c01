"""Benchmark: the f2 block-tensor extraction on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c2] [--precision split_bf16|fp64]
                    [--qpadm-models N] [--no-qpadm] [--qpgraph-graphs N] [--no-qpgraph]
                    [--fstat-quartets N] [--no-fstats] [--frontend-snps N] [--no-frontend]
                    [--no-tiers] [--no-qpgraph-search] [--no-cpu-baseline]
    python bench.py --calibrate-copy      # a plain 1 GiB device copy, for counter calibration

One step = one pass of the hot path over the synthetic panel already resident in
HBM: decode the packed 2-bit genotypes to per-population counts, then build the
per-jackknife-block f2 and vpair tensors [n_block x P x P] in HBM.  The timed steps
replay one captured HIP graph of the whole step (stp_f2_graph_*; --eager times the
separate calls instead).  Prints one JSON line.  The metric is f2 pop-pair.SNPs/s: W / t with W = P(P+1)/2 * M
(unordered population pairs incl. the diagonal, times SNPs).

Workloads (synthetic, seeded; population frequencies Balding-Nichols, genotypes
Binomial(2,p), 40% pseudo-haploid populations with Beta(2,2) missingness):
  c2: 200 populations (sizes U{1..60}), 1.2M SNPs, blocks U{200..3000} SNPs
  c3: 1000 populations (sizes U{1..60}), 1.2M SNPs, blocks LogNormal(ln 1500, 0.6)
      clipped to [50, 8000]

The default workload is c3 in the split-BF16 form (the metric's "split-BF16 vs
FP64" configuration).

Extra JSON fields: ``roofline`` for the dominant kernel (the f2 contraction),
timed with HIP events on the engine's own stream: achieved = k * F_alg / t, with
F_alg = 8 W FLOP (4 MACs per pair-SNP) and k = 1 for the FP64 form, k = 7.25 for the
split form (29 BF16/FP16 MFMA products per pair-SNP over 4 MACs); ``traffic`` is the HBM bytes
per launch from the committed rocprofv3 counter summary of the same workload
(profiles/), or null; ``stages`` with every stage's average device time;
``cpu_baseline``: the reference CpuBackend (oracle/_ref, compiled from the reference
sources) or the restated oracle, single-threaded, on a bounded sample of the same
workload.  ``qpadm``: the batched qpAdm metric (fits/s) for 10,000 random models
(1 target + 2-5 sources, 6-12 right populations, drawn without overlap) on the
resident c2 tensor, with its own single-threaded reference CPU rate on a sample of
the same models.  ``qpgraph``: qpGraph fits/s for 5,000 random graphs (8-12 leaves, 1-4
admixture events, each with its own simulated f2 over 300 blocks), 10 restarts each to
convergence by batched L-BFGS on the device score and adjoint, with the reference CPU
rate and the score agreement on a sample of the same graphs.  ``fstats``: standalone
f4 statistics/s for 1,000,000 random quartets on the resident c2 tensor, the
all-combination f3 sweep (1.3M triples) and f4 sweep (64.7M quartets, Gram form on
FP64 MFMA) at |z| >= 3, with the reference CPU run_f4 rate on a sample.  ``frontend``:
genotype files -> resident f2 for GENO and TGENO triples of the c2 shape (random codes,
1.2M SNPs), with the reference's own reader / filter / f2 chain on a 40k-SNP sample.
``tiers``: the same decode + f2 of the default workload with the result kept in HBM,
streamed into host memory, and streamed into an f2.bin file (wall time and GB/s of
result).  ``qpgraph_search``: the topology search on the AT2 qpGraph fixture (all 1590
trees and one-admixture graphs on 5 of its populations, one device fleet), candidate
topologies fitted/s with the reference's published rate beside it and the reference CPU
fit rate on a sample of the candidates.
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
SPLIT_K = 29 * 2 / 8.0   # executed BF16/FP16 MFMA FLOP per algorithmic FLOP in the split form
PROFILES = {("c3", "split_bf16"): ("profiles/f2_split_c3.md", "f2_split_kernel", "f2_split.hip"),
            ("c3", "fp64"): ("profiles/f2_fp64_c3.md", "f2_fp64_kernel", "f2_fp64.hip")}
METRIC = "f2 pop-pair·SNPs/sec at ≤1e-9 rel err (split-BF16 vs FP64); qpAdm fits/sec"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_params(name: str):
    from steppe_amd import synth
    return synth.baseline_workload(name, SEED)


def build_panel(ctx, wp):
    """Generate the packed panel directly in HBM (returns the torch buffer)."""
    from steppe_amd import synth
    dp = synth.device_panel(ctx, wp)
    return dp.packed, dp.n_ind, dp.bytes_per_record, dp.pop_offsets


def host_cpu_info() -> dict:
    """The host CPU model and the oracle build's compiler and flags."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    info = {"cpu_model": model, "compiler": None, "flags": None}
    for path in (os.path.join(ROOT, "oracle", "_ref", "build_info.txt"),):
        try:
            for line in open(path):
                k, _, v = line.partition(":")
                if k.strip() in ("compiler", "flags"):
                    info[k.strip()] = v.strip()
        except OSError:
            pass
    return info


def cpu_baseline(packed, n_ind, bpr, offs, wp, dev_f2=None, dev_vp=None, target_s: float = 15.0):
    """Single-thread reference CpuBackend (or the restated oracle) on a sample:
    all populations x the leading blocks totalling ~target_s seconds of work.  When the
    device tensors are given ([nb, P, P] views of f2 and vpair), the reference's blocks
    are compared with them: max relative f2 error (floor 1e-6) and vpair equality."""
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
    fr, vr, _ = lib.f2_blocks(q, v, n, bid, k)
    dt = time.perf_counter() - t0
    value = P * (P + 1) / 2 * m / dt
    out = {"value": value, "unit": "pop-pair·SNPs/s", "cores": 1, "kind": kind,
           "sample": f"all {P} populations x first {k} blocks ({m} SNPs), f2 blocks from decoded "
                     f"Q/V/N, {dt:.1f} s single-threaded ({os.path.basename(lib.path)})"}
    out.update(host_cpu_info())
    if dev_f2 is not None:
        import torch
        worst, vp_equal = 0.0, True
        for b in range(k):
            want = torch.from_numpy(np.ascontiguousarray(fr[:, :, b].T))     # [j, i]
            got = dev_f2[b].cpu()
            worst = max(worst, ((got - want).abs() / want.abs().clamp_min(1e-6)).max().item())
            vp_equal &= bool(torch.equal(dev_vp[b].cpu(), torch.from_numpy(np.ascontiguousarray(vr[:, :, b].T))))
        out["compared_blocks"] = k
        out["max_rel_err_device_vs_reference"] = worst
        out["vpair_equal"] = vp_equal
    return out


def source_sha(name: str) -> str:
    import hashlib
    with open(os.path.join(ROOT, "steppe_amd", "csrc", name), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()[:16]


def profiled_traffic(workload: str, precision: str):
    """HBM bytes per launch of the f2 kernel from the committed counter summary (the
    'traffic GB' column of its row), used only when the summary names the same kernel
    source hash as the file in the tree (a summary of an older kernel gives None)."""
    key = PROFILES.get((workload, precision))
    if not key:
        return None, None
    path, kernel, src = key
    want = source_sha(src)
    traffic, sha = None, None
    try:
        for line in open(os.path.join(ROOT, path)):
            if line.startswith("Kernel source:"):
                sha = line.split("sha256")[-1].strip().strip("`. ")
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6 and cells[0] == f"`{kernel}`":
                traffic = float(cells[5]) * 1e9
    except OSError:
        return None, None
    if sha != want:
        return None, f"{path} (kernel source {sha}, tree {want}: stale, not used)"
    return traffic, f"{path} (kernel source sha256 {want})"


_C2 = {}


def c2_tensor(ctx):
    """The resident c2 tensor (200 populations, FP64 f2 from the c2 panel), built once."""
    if "h" in _C2:
        return _C2["h"]
    import torch
    from steppe_amd import _core
    wp = workload_params("c2")
    P, M = wp["P"], wp["M"]
    bs = wp["block_sizes"]
    nb = len(bs)
    block_id = np.ascontiguousarray(np.repeat(np.arange(nb, dtype=np.int32), bs))
    packed, n_ind, bpr, offs = build_panel(ctx, wp)
    tile, keep = _core.make_tile(None, M, offs, None, detect=True, bytes_per_record=bpr,
                                 n_individuals=n_ind, device_ptr=packed.data_ptr())
    raw = C.c_void_p()
    _core.check(_core.lib().stp_extract_f2(ctx.h, C.byref(tile), _core._ptr(block_id), nb,
                                           _core.STP_PREC_FP64, C.byref(raw)), "extract")
    h = _core.F2Handle(raw, [f"p{i}" for i in range(P)], 0)
    del packed
    torch.cuda.empty_cache()
    _C2["h"] = h
    return h


def qpadm_bench(ctx, n_models: int, cpu: bool):
    """C4: n_models random models on the resident c2 tensor in one batched call."""
    from steppe_amd import _core
    h = c2_tensor(ctx)
    P, nb = h.P, h.n_block
    from steppe_amd import synth
    models = synth.random_qpadm_models(np.random.default_rng(SEED + 4), P, n_models)
    _core.qpadm_models(h, models[:64], raw=True)          # warm-up (code load, buffers)
    ctx.timing(True)
    ctx.read_timing(reset=True)
    t0 = time.perf_counter()
    recs = _core.qpadm_models(h, models, raw=True)
    dt = time.perf_counter() - t0
    tm = ctx.read_timing(reset=True)
    ctx.timing(False)
    ok = sum(1 for r in recs if r.status in (0, 4))
    log(f"[bench] qpadm: {n_models} models in {dt:.3f} s ({ok} ok); kernel {tm['qpadm']}")
    out = {"metric": "qpAdm fits/sec", "value": n_models / dt, "unit": "fits/s", "models": n_models,
           "seconds": dt, "kernel_ms": tm["qpadm"][0],
           "workload": f"c4: {n_models} random models (1 target + 2-5 sources, 6-12 right) on the "
                       f"resident c2 tensor ({P} populations, {nb} blocks, FP64), weights + p + "
                       f"rank sweep + popdrop + LOO SEs per model",
           "cpu_baseline": None}
    # A rotation of the published size: one target, every 2- and 3-source subset of a
    # 40-population pool (780 + 9880 = 10,660 models), 8 shared right populations.
    import itertools
    pool = list(range(1, 41))
    right = list(range(41, 49))
    rot = [(0, list(c), right) for k in (2, 3) for c in itertools.combinations(pool, k)]
    ctx.timing(True)
    ctx.read_timing(reset=True)
    t0 = time.perf_counter()
    rrecs = _core.qpadm_models(h, rot, raw=True)
    rdt = time.perf_counter() - t0
    rtm = ctx.read_timing(reset=True)
    ctx.timing(False)
    out["rotation"] = {"models": len(rot), "seconds": rdt, "value": len(rot) / rdt, "unit": "fits/s",
                       "kernel_ms": rtm["qpadm"][0],
                       "ok": sum(1 for r in rrecs if r.status in (0, 4)),
                       "workload": "target 0, every 2- and 3-source subset of populations 1-40, right 41-48, "
                                   "jackknife all, on the resident c2 tensor",
                       "baseline": {"value": 2340.0, "unit": "fits/s",
                                    "source": "reference README: 10,660-model rotation in 4.55 s on an RTX 5090 "
                                              "(AADR)"}}
    log(f"[bench] qpadm rotation: {len(rot)} models in {rdt:.3f} s ({len(rot) / rdt:.0f} fits/s)")
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        lib = oracle_lib.ref()
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib.cref(), "port"
        f2h = h._f2_numpy()
        vph = h._vpair_numpy()
        bsz = np.asarray(h.block_sizes, np.int32)
        t0 = time.perf_counter()
        n = 0
        while n < len(models) and time.perf_counter() - t0 < 10.0:
            t, l, r = models[n]
            lib.qpadm(f2h, vph, bsz, t, l, r)
            n += 1
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "fits/s", "cores": 1, "kind": kind,
                               "sample": f"the first {n} of the same models, {cdt:.1f} s single-threaded "
                                         f"({os.path.basename(lib.path)})"}
        log(f"[bench] qpadm cpu baseline {out['cpu_baseline']}")
    return out


def fstats_bench(ctx, n_quartets: int, cpu: bool):
    """Standalone f-statistics on the resident c2 tensor: run_f4 over n_quartets random
    quartets, the all-combination f3 sweep and the all-combination f4 sweep (Gram form)
    over its 200 populations at the default filter (|z| >= 3)."""
    from math import comb
    from steppe_amd import _core
    h = c2_tensor(ctx)
    P, nb = h.P, h.n_block
    rng = np.random.default_rng(SEED + 6)
    quartets = np.argsort(rng.random((n_quartets, P)), axis=1)[:, :4].astype(np.int32)
    _core.run_f4(h, quartets[:1000])                       # warm-up
    ctx.timing(True)
    ctx.read_timing(reset=True)
    t0 = time.perf_counter()
    d = _core.run_f4(h, quartets)
    t_f4 = time.perf_counter() - t0
    tm = ctx.read_timing(reset=True)
    f4_ms = tm["fstat"][0]
    # f4 kernel: 4 f2 series of nb doubles per statistic (algorithmic bytes)
    f4_bytes = 4.0 * nb * 8.0 * n_quartets
    t0 = time.perf_counter()
    s3 = _core.run_fstat_sweep(h, 3, None, min_z=3.0)
    t_s3 = time.perf_counter() - t0
    tm3 = ctx.read_timing(reset=True)
    _core.run_fstat_sweep(h, 4, list(range(40)), min_z=3.0, method="gram")   # warm-up
    ctx.read_timing(reset=True)
    t0 = time.perf_counter()
    s4 = _core.run_fstat_sweep(h, 4, None, min_z=3.0, method="gram")
    t_s4 = time.perf_counter() - t0
    tm4 = ctx.read_timing(reset=True)
    ctx.timing(False)
    # Gram work: per (c2, c3) the upper triangle of S (m = c2 rows) over nb blocks
    gram_flop = float(sum(c2 * (c2 + 1) * nb * (P - 1 - c2) for c2 in range(2, P - 1)))
    n4, n3 = comb(P, 4), comb(P, 3)
    log(f"[bench] fstats: run_f4 {n_quartets} in {t_f4:.3f} s (kernel {f4_ms:.1f} ms); f3 sweep {n3} in "
        f"{t_s3:.3f} s ({s3['survivors']} kept); f4 sweep {n4} in {t_s4:.3f} s (device {tm4['fstat'][0]:.1f} ms, "
        f"{s4['survivors']} kept)")
    out = {"metric": "f4 statistics/sec", "value": n_quartets / t_f4, "unit": "stats/s",
           "workload": f"{n_quartets} random f4 quartets on the resident c2 tensor ({P} populations, {nb} "
                       f"blocks, FP64): est / se / z / p by the block jackknife",
           "seconds": t_f4, "kernel_ms": f4_ms,
           "f4_kernel": {"algorithmic_bytes": f4_bytes, "GBs": f4_bytes / (f4_ms / 1e3) / 1e9 if f4_ms else None,
                         "work": "4 f2 series x n_block x 8 B per statistic"},
           "f3_sweep": {"combinations": n3, "seconds": t_s3, "device_ms": tm3["fstat"][0],
                        "survivors": s3["survivors"], "value": n3 / t_s3, "unit": "triples/s"},
           "f4_sweep": {"combinations": n4, "seconds": t_s4, "device_ms": tm4["fstat"][0],
                        "survivors": s4["survivors"], "value": n4 / t_s4, "unit": "quartets/s",
                        "method": "gram", "gram_TFLOPs": gram_flop / (tm4["fstat"][0] / 1e3) / 1e12
                        if tm4["fstat"][0] else None,
                        "gram_work": "sum over (c2, c3) of c2 (c2 + 1) n_block FLOP (upper triangle of S)"},
           "cpu_baseline": None}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        lib = oracle_lib.ref()
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib.cref(), "port"
        f2h = h._f2_numpy()
        vph = h._vpair_numpy()
        bsz = np.asarray(h.block_sizes, np.int32)
        n, chunk = 0, 2000
        t0 = time.perf_counter()
        while n < n_quartets and time.perf_counter() - t0 < 10.0:
            o = lib.f4(f2h, vph, bsz, quartets[n:n + chunk])
            m = np.abs(o["est"] - np.asarray(d["est"][n:n + chunk]))
            assert np.all(m <= 1e-12 * (np.abs(o["est"]) + np.nan_to_num(o["se"]))), "f4 bench row disagrees"
            n += chunk
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "stats/s", "cores": 1, "kind": kind,
                               "sample": f"the first {n} of the same quartets through run_f4, {cdt:.1f} s "
                                         f"single-threaded ({os.path.basename(lib.path)})"}
        log(f"[bench] fstats cpu baseline {out['cpu_baseline']}")
    return out


def write_random_triple(prefix: str, fmt: str, pop_of, n_snp: int, seed: int):
    """A packed triple of uniformly random 2-bit codes (1/4 missing) with random ACGT allele
    columns and 22 autosomes of 0.05-Morgan-spaced SNPs: input for timing the file path."""
    rng = np.random.default_rng(seed)
    n_ind = len(pop_of)
    chrom = np.minimum(np.arange(n_snp) * 22 // n_snp + 1, 22)
    gpos = np.arange(n_snp) * (0.05 * 22 * 30 / n_snp)
    if fmt == "plink":
        bpr = (n_ind + 3) // 4
        with open(prefix + ".bed", "wb") as f:
            f.write(bytes([0x6C, 0x1B, 0x01]))
            step = max(1, (64 << 20) // bpr)
            for r0 in range(0, n_snp, step):
                f.write(rng.integers(0, 256, (min(step, n_snp - r0), bpr), dtype=np.uint8).tobytes())
        with open(prefix + ".fam", "w") as f:
            f.write("".join(f"F{g} I{g} 0 0 1 p{pop_of[g]:03d}\n" for g in range(n_ind)))
        alle = np.array(list("ACGT"))
        ra, aa = alle[rng.integers(0, 4, n_snp)], alle[rng.integers(0, 4, n_snp)]
        with open(prefix + ".bim", "w") as f:
            f.write("".join(f"{chrom[k]} s{k} {gpos[k]:.6f} {k * 100} {ra[k]} {aa[k]}\n" for k in range(n_snp)))
        return
    if fmt == "tgeno":
        bpr = (n_snp + 3) // 4
        head = f"TGENO {n_ind} {n_snp} 0 0".encode().ljust(48, b"\0")
        n_rec = n_ind
    else:
        bpr = max(48, (n_ind + 3) // 4)
        head = f"GENO {n_ind} {n_snp} 0 0".encode().ljust(bpr, b"\0")
        n_rec = n_snp
    with open(prefix + ".geno", "wb") as f:
        f.write(head)
        step = max(1, (64 << 20) // bpr)
        for r0 in range(0, n_rec, step):
            f.write(rng.integers(0, 256, (min(step, n_rec - r0), bpr), dtype=np.uint8).tobytes())
    with open(prefix + ".ind", "w") as f:
        f.write("".join(f"I{g} U p{pop_of[g]:03d}\n" for g in range(n_ind)))
    alle = np.array(list("ACGT"))
    ra, aa = alle[rng.integers(0, 4, n_snp)], alle[rng.integers(0, 4, n_snp)]
    with open(prefix + ".snp", "w") as f:
        f.write("".join(f"s{k} {chrom[k]} {gpos[k]:.6f} {k * 100} {ra[k]} {aa[k]}\n" for k in range(n_snp)))


def frontend_bench(ctx, n_snp: int, cpu: bool):
    """Genotype files -> resident f2: GENO and PLINK .bed (SNP-major, transposed on the
    device) and TGENO triples of the c2 shape (200 populations, its individuals, n_snp SNPs), explicit
    selection of every population, autosomes only, strand-ambiguous and non-ACGT pairs
    dropped, no population-missingness cap (random codes leave some single-sample
    population without data at almost every SNP), FP64 f2."""
    import tempfile
    from steppe_amd import _core
    wp = workload_params("c2")
    sizes = wp["sizes"]
    pop_of = np.repeat(np.arange(len(sizes)), sizes)
    pops = [f"p{i:03d}" for i in range(len(sizes))]
    out = {"metric": "genotype-file extract SNPs/sec", "unit": "SNPs/s", "cpu_baseline": None}
    with tempfile.TemporaryDirectory() as d:
        for fmt in ("geno", "tgeno", "plink"):
            pre = os.path.join(d, fmt)
            t0 = time.perf_counter()
            write_random_triple(pre, fmt, pop_of, n_snp, SEED + 7)
            t_write = time.perf_counter() - t0
            files = _core.genotype_triple(pre)
            _core.extract_f2_files(*files, pops[:3], maxmiss=1.0, precision="fp64")   # warm-up (code load)
            ctx.timing(True)
            ctx.read_timing(reset=True)
            t0 = time.perf_counter()
            h, info = _core.extract_f2_files(*files, pops, maxmiss=1.0, precision="fp64")
            dt = time.perf_counter() - t0
            tm = ctx.read_timing(reset=True)
            ctx.timing(False)
            geno_bytes = os.path.getsize(files[0])
            log(f"[bench] frontend {fmt}: {n_snp} SNPs x {len(pop_of)} individuals ({geno_bytes / 1e9:.2f} GB) in "
                f"{dt:.3f} s (files written in {t_write:.1f} s); kept {info['n_snp_kept']}; stages {tm}")
            out[fmt] = {"seconds": dt, "value": n_snp / dt, "geno_bytes": geno_bytes, "n_snp_kept": info["n_snp_kept"],
                        "n_block": info["n_block"], "frontend_ms": tm["frontend"][0], "decode_ms": tm["decode"][0],
                        "f2_ms": tm["f2"][0] + tm["prep"][0] + tm["count"][0]}
            del h
        if cpu:
            sys.path.insert(0, os.path.join(ROOT, "tests"))
            import oracle_lib
            lib = oracle_lib.ref()
            kind = "reference"
            if lib is None:
                lib, kind = oracle_lib.cref(), "port"
            m = 40_000
            pre = os.path.join(d, "sample")
            write_random_triple(pre, "geno", pop_of, m, SEED + 8)
            t0 = time.perf_counter()
            lib.extract(pre + ".geno", pre + ".snp", pre + ".ind", pops=pops, maxmiss=1.0, autosomes_only=True)
            cdt = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": m / cdt, "unit": "SNPs/s", "cores": 1, "kind": kind,
                                   "sample": f"a GENO triple of the same shape with {m} SNPs, {cdt:.1f} s single-"
                                             f"threaded ({os.path.basename(lib.path)})"}
            log(f"[bench] frontend cpu baseline {out['cpu_baseline']}")
    out["value"] = out["geno"]["value"]
    out["workload"] = (f"GENO, TGENO and PLINK .bed triples of random codes, {len(sizes)} populations x {len(pop_of)} "
                       f"individuals x {n_snp} SNPs, all populations selected, autosomes, strand-ambiguous / non-ACGT "
                       f"pairs dropped, maxmiss 1, FP64 f2; "
                       f"value = SNPs/s of the GENO file (read, transpose, ploidy, decode, keep, compact, f2)")
    return out


def tier_bench(ctx, tile, block_id, nb: int, P: int, prec: int):
    """The same decode + f2 extract with the result kept in HBM, streamed into host memory,
    and streamed into an STPF2BK1 f2.bin file (in a temporary directory, removed after):
    wall time per tier and the rate at which the result leaves the device."""
    import shutil
    import tempfile
    from steppe_amd import _core
    L = _core.lib()
    result_bytes = 2 * P * P * nb * 8
    out = {"result_GB": result_bytes / 1e9, "chunk_MiB": 256}
    tmp = tempfile.mkdtemp(prefix="stp_tier_")
    try:
        for name, tier in (("resident", 0), ("host", 1), ("disk", 2)):
            path = os.path.join(tmp, "f2.bin")
            if tier == 2 and shutil.disk_usage(tmp).free < 3 * result_bytes:
                out[name] = {"skipped": "not enough free disk space"}
                continue
            raw = C.c_void_p()
            t0 = time.perf_counter()
            _core.check(L.stp_extract_f2_tiered(ctx.h, C.byref(tile), _core._ptr(block_id), nb, prec, tier,
                                                path.encode(), 0, C.byref(raw)), "tier")
            dt = time.perf_counter() - t0
            _core.check(L.stp_f2_handle_free(raw))
            out[name] = {"s": dt, "GBs": result_bytes / dt / 1e9}
            log(f"[bench] tier {name}: {dt:.3f} s ({result_bytes / dt / 1e9:.1f} GB/s of result)")
            if os.path.exists(path):
                os.remove(path)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def qpgraph_bench(ctx, n_graphs: int, cpu: bool, numstart: int = 10):
    """C5: n_graphs random graphs, each fitted from ``numstart`` restarts by the device
    L-BFGS on the score + adjoint (one workgroup per restart, to convergence); the f3
    basis and covariance of every graph are computed on the device inside the timed
    region."""
    from steppe_amd import _core, qpgraph
    t0 = time.perf_counter()
    from steppe_amd import synth
    graphs = synth.random_graph_cases(np.random.default_rng(SEED + 5), n_graphs)
    handles = [_core.upload_f2(f2, None, bs, names) for _, f2, bs, names in graphs]
    log(f"[bench] qpgraph: {n_graphs} graphs simulated and resident in {time.perf_counter() - t0:.1f} s")
    # warm-up: code load and staging buffers
    w = qpgraph.GraphSet([qpgraph.Topology(graphs[0][0], graphs[0][3])], handles[:1])
    w.fit_fleet(numstart=2, maxit=3)
    del w
    ctx.timing(True)
    ctx.read_timing(reset=True)
    t0 = time.perf_counter()
    topos = [qpgraph.Topology(edges, names) for edges, _, _, names in graphs]
    gset = qpgraph.GraphSet(topos, handles)
    t_basis = time.perf_counter() - t0
    fits = qpgraph.fit_sets(gset, numstart=numstart)
    dt = time.perf_counter() - t0
    tm = ctx.read_timing(reset=True)
    ctx.timing(False)
    ok = sum(1 for f in fits if f is not None and f["status"] == 0)
    kern_ms = tm["qpgraph"][0]
    iters = np.array([f["iterations"] for f in fits if f is not None])
    log(f"[bench] qpgraph: {n_graphs} fits in {dt:.3f} s (basis {t_basis:.3f} s, {ok} ok); device fit + final "
        f"evaluation kernels {kern_ms:.1f} ms; L-BFGS iterations per graph (most over restarts) mean "
        f"{iters.mean():.1f}, max {iters.max()}")
    out = {"metric": "qpGraph fits/sec", "value": n_graphs / dt, "unit": "fits/s", "graphs": n_graphs, "ok": ok,
           "seconds": dt, "basis_seconds": t_basis, "restarts": numstart, "kernel_ms": kern_ms,
           "iterations_mean": float(iters.mean()), "iterations_max": int(iters.max()),
           "workload": f"c5: {n_graphs} random graphs (8-12 leaves, 1-4 admixture events), each with its own "
                       f"simulated f2 over 300 blocks resident in HBM; device f3 basis + covariance, then "
                       f"{numstart} restarts per graph, each run to convergence by projected L-BFGS on the "
                       f"device score and adjoint gradient inside one kernel (FP64)",
           "cpu_baseline": None}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        lib = oracle_lib.ref()
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib.cref(), "port"
        t0 = time.perf_counter()
        n, worse = 0, 0
        excess = 0.0
        while n < n_graphs and time.perf_counter() - t0 < 10.0:
            edges, f2, bs, names = graphs[n]
            r = lib.qpgraph_fit(f2, None, bs, edges, names)
            mine = fits[n]["score"] if fits[n] is not None else float("inf")
            rel = (mine - r["score"]) / abs(r["score"])
            excess = max(excess, rel)
            worse += rel > 1e-7
            n += 1
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / cdt, "unit": "fits/s", "cores": 1, "kind": kind,
                               "sample": f"the first {n} of the same graphs, numstart 10, {cdt:.1f} s "
                                         f"single-threaded ({os.path.basename(lib.path)})"}
        out["vs_cpu_sample"] = {"graphs": n, "device_worse_by_1e-7_rel": int(worse),
                                "max_rel_score_excess": excess}
        log(f"[bench] qpgraph cpu baseline {out['cpu_baseline']} {out['vs_cpu_sample']}")
    return out


SEARCH_NAMES = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
                "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]
SEARCH_POPS = ["Mbuti", "Han", "Iran_GanjDareh_N", "Israel_Natufian", "Czechia_EBA_CordedWare"]


def qpgraph_search_bench(ctx, cpu: bool, reps: int = 3):
    """The topology search on the AT2 qpGraph fixture (9 populations, 710 blocks): every
    tree and one-admixture graph on 5 of them (1590 candidates), 10 restarts per admixture
    graph, fitted in one device fleet against one shared f3 basis; then the 8-seed
    hill-climb.  Reports the fleet's wall time (the reference's fit_all_wall_ms) and the
    whole search's."""
    from steppe_amd import _core, qpgraph
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from fixtures import read_raw_fixture
    f2, bs = read_raw_fixture(os.path.join(ROOT, "tests", "golden", "at2", "f2_qpgraph_9pop.bin"))
    h = _core.upload_f2(f2, None, bs, SEARCH_NAMES)
    qpgraph.run_qpgraph_search(h, SEARCH_POPS[:3], run_heuristic=False)   # warm-up: code load
    ctx.timing(True)
    ctx.read_timing(reset=True)
    fleet, wall = [], []
    d = None
    for _ in range(reps):
        t0 = time.perf_counter()
        d = qpgraph.run_qpgraph_search(h, SEARCH_POPS)
        wall.append(time.perf_counter() - t0)
        fleet.append(d["fit_all_wall_ms"])
    tm = ctx.read_timing(reset=True)
    ctx.timing(False)
    n = d["n_candidates"]
    fl = float(np.median(fleet))
    out = {"metric": "qpGraph topology search: candidate topologies fitted/sec", "value": n / (fl / 1e3),
           "unit": "topologies/s", "candidates": n, "fleet_ms": fl, "search_wall_s": float(np.median(wall)),
           "kernel_ms_per_search": tm["qpgraph"][0] / reps, "best_score": d["best_score"],
           "best_hash": d["best_hash"], "heuristic_recovered": d["heuristic_recovered"],
           "workload": "AT2 qpGraph fixture (9 populations, 710 blocks); all 105 trees and 1485 one-admixture "
                       "graphs on Mbuti, Han, Iran_GanjDareh_N, Israel_Natufian, Czechia_EBA_CordedWare; "
                       "10 restarts per admixture graph (plus grid seeds), device L-BFGS, FP64; then the "
                       "8-seed hill-climb (in search_wall_s)",
           "baseline": {"value": 1450.0, "unit": "topologies/s", "source": "reference README: 1,590 graphs in "
                        "1.1 s on an RTX 5090 (AADR)"},
           "cpu_baseline": None}
    log(f"[bench] qpgraph search: {n} candidates, fleet {fl:.1f} ms ({out['value']:.0f} topologies/s), "
        f"whole search {out['search_wall_s']:.3f} s, kernels {out['kernel_ms_per_search']:.1f} ms")
    if cpu:
        import oracle_lib
        lib = oracle_lib.ref()
        kind = "reference"
        if lib is None:
            lib, kind = oracle_lib.cref(), "port"
        cands = qpgraph.enumerate_graphs(SEARCH_POPS, 1)
        pick = cands[::8]
        t0 = time.perf_counter()
        k = 0
        while k < len(pick) and time.perf_counter() - t0 < 10.0:
            lib.qpgraph_fit(f2, None, bs, pick[k]["edges"], SEARCH_NAMES)
            k += 1
        cdt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / cdt, "unit": "topologies/s", "cores": 1, "kind": kind,
                               "sample": f"run_qpgraph on {k} of the candidates (every 8th, numstart 10), "
                                         f"{cdt:.1f} s single-threaded ({os.path.basename(lib.path)})"}
        log(f"[bench] qpgraph search cpu baseline {out['cpu_baseline']}")
    return out


def calibrate_copy():
    """A plain 1 GiB device-to-device copy, three times (counter calibration)."""
    import torch
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c3", choices=["c2", "c3"])
    ap.add_argument("--precision", default="split_bf16", choices=["fp64", "split_bf16"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--qpadm-models", type=int, default=10_000)
    ap.add_argument("--no-qpadm", action="store_true")
    ap.add_argument("--qpgraph-graphs", type=int, default=5_000)
    ap.add_argument("--no-qpgraph", action="store_true")
    ap.add_argument("--fstat-quartets", type=int, default=1_000_000)
    ap.add_argument("--no-fstats", action="store_true")
    ap.add_argument("--frontend-snps", type=int, default=1_200_000)
    ap.add_argument("--no-frontend", action="store_true")
    ap.add_argument("--no-tiers", action="store_true")
    ap.add_argument("--no-qpgraph-search", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--eager", action="store_true", help="time eager calls instead of the captured graph")
    ap.add_argument("--no-autotune", action="store_true", help="use the default launch configuration")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the split-vs-FP64 comparison leg (for profiling the split kernels alone)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
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
    # The contraction's launch configuration for this workload, chosen by timing every
    # candidate (stp_f2_autotune, outside the timed region; the graph below uses it).
    tune = None
    if not args.no_autotune:
        tune = {"split_bf16" if not prec == _core.STP_PREC_FP64 else "fp64":
                _core.f2_autotune(ctx, a2.data_ptr(), nn.data_ptr(), ld, P, M, block_id, nb, prec, f2.data_ptr(),
                                  vp.data_ptr())}
        for k, r in tune.items():
            log(f"[bench] autotune {k}: chosen {r['candidates'][r['chosen']]['config']} "
                f"({r['candidates'][r['chosen']]['ms']:.2f} ms; default {r['candidates'][0]['ms']:.2f} ms)")
    # Eager steps (one stp_decode_counts + one stp_f2_blocks_from_counts call each), timed
    # for comparison with the captured graph.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        step()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) / 3 * 1e3
    ctx.timing(True)
    graph = None
    run = step
    if not args.eager:
        # The whole step (ploidy detection, decode, prep, pair counts, contraction, fix-up)
        # as one captured HIP graph, with the stage events captured inside it.
        graph = _core.F2Graph(ctx, tile, block_id, nb, prec, f2.data_ptr(), vp.data_ptr(), a2.data_ptr(),
                              nn.data_ptr(), ld)
        graph.launch()
        run = graph.launch
    ctx.read_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = ctx.read_timing(reset=True)
    stage_source = "graph" if graph is not None else "eager"
    if tm["f2"][1] == 0:
        # the runtime did not capture the stage events: time the kernels on eager steps
        for _ in range(args.steps):
            step()
        tm = ctx.read_timing(reset=True)
        stage_source = "eager steps after the timed graph replays"
    ctx.timing(False)
    log(f"[bench] {args.steps} {'graph replays' if graph else 'eager steps'} in {dt:.3f} s "
        f"(eager step {eager_ms:.2f} ms); stages ({stage_source}) {tm}")

    W = P * (P + 1) / 2 * M
    value = W * args.steps / dt
    ms_step = dt / args.steps * 1e3
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in tm.items()}
    # Algorithmic work of the contraction: 4 MACs per (pair, SNP) = 8 W FLOP.
    f_alg = 8.0 * W
    f2_s = avg["f2"] / 1e3
    fp64 = prec == _core.STP_PREC_FP64
    peak = MI355X_PEAK["fp64_TFLOPs"] if fp64 else MI355X_PEAK["bf16_TFLOPs"]
    kf = 1.0 if fp64 else SPLIT_K
    achieved = kf * f_alg / f2_s / 1e12 if f2_s > 0 else 0.0
    traffic, traffic_src = profiled_traffic(args.workload, args.precision)
    roofline = {"kernel": "f2_fp64_kernel" if fp64 else "f2_split_kernel",
                "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                "k": kf, "algorithmic_TFLOPs": f_alg / f2_s / 1e12 if f2_s > 0 else 0.0,
                "work": ("8*W FLOP per launch, W = P(P+1)/2*M" if fp64 else
                         "k*8*W FLOP per launch (k = 7.25 BF16/FP16 MFMA FLOP per algorithmic FLOP: 21 BF16 "
                         "digit products + 8 FP16 w-digit x V products per pair-SNP), W = P(P+1)/2*M")}
    dec_bytes = n_ind * bpr + 4 * P * M
    stages = {k: {"avg_ms": avg[k]} for k in avg}
    stages["decode"]["hbm_GBs"] = dec_bytes / (avg["decode"] / 1e3) / 1e9 if avg["decode"] else None
    stages["decode"]["algorithmic_bytes"] = dec_bytes

    # The split form against the plain FP64 form on the same counts: the whole tensor's
    # max relative error (floor 1e-6), vpair equality, and the FP64 step time.
    accuracy = None
    if not fp64 and not args.no_accuracy:
        f2d = torch.empty_like(f2)
        vpd = torch.empty_like(vp)
        torch.cuda.synchronize()

        if tune is not None:
            tune["fp64"] = _core.f2_autotune(ctx, a2.data_ptr(), nn.data_ptr(), ld, P, M, block_id, nb,
                                             _core.STP_PREC_FP64, f2d.data_ptr(), vpd.data_ptr())
            r = tune["fp64"]
            log(f"[bench] autotune fp64: chosen {r['candidates'][r['chosen']]['config']} "
                f"({r['candidates'][r['chosen']]['ms']:.2f} ms; default {r['candidates'][0]['ms']:.2f} ms)")
        if graph is not None:
            g64 = _core.F2Graph(ctx, tile, block_id, nb, _core.STP_PREC_FP64, f2d.data_ptr(), vpd.data_ptr(),
                                a2.data_ptr(), nn.data_ptr(), ld)
            step64 = g64.launch
        else:
            def step64():
                _core.check(L.stp_decode_counts(ctx.h, C.byref(tile), C.c_void_p(a2.data_ptr()),
                                                C.c_void_p(nn.data_ptr()), ld), "decode")
                _core.check(L.stp_f2_blocks_from_counts(ctx.h, C.c_void_p(a2.data_ptr()), C.c_void_p(nn.data_ptr()),
                                                        ld, P, M, bid_p, nb, _core.STP_PREC_FP64,
                                                        C.c_void_p(f2d.data_ptr()), C.c_void_p(vpd.data_ptr())), "f2")
        step64()
        n64 = 3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n64):
            step64()
        torch.cuda.synchronize()
        ms64 = (time.perf_counter() - t0) / n64 * 1e3
        worst = 0.0
        fs_, fd_ = f2.view(nb, P, P), f2d.view(nb, P, P)
        for b0 in range(0, nb, 64):
            x, y = fs_[b0:b0 + 64], fd_[b0:b0 + 64]
            worst = max(worst, ((x - y).abs() / y.abs().clamp_min(1e-6)).max().item())
        accuracy = {"max_rel_err_split_vs_fp64": worst, "floor": 1e-6, "elements": P * P * nb,
                    "vpair_equal": bool(torch.equal(vp, vpd)), "fp64_ms_per_step": ms64,
                    "split_ms_per_step": ms_step, "fp64_over_split_step": ms64 / ms_step}
        log(f"[bench] split vs FP64: max rel err {worst:.3e}; FP64 step {ms64:.1f} ms ({ms64 / ms_step:.2f}x split)")
        if graph is not None:
            g64.close()
        del f2d, vpd
        torch.cuda.empty_cache()

    tiers = None
    if not args.no_tiers:
        tiers = tier_bench(ctx, tile, block_id, nb, P, prec)
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        cpu = cpu_baseline(packed, n_ind, bpr, offs, wp, f2.view(nb, P, P), vp.view(nb, P, P))
        log(f"[bench] cpu baseline {cpu}")
    if graph is not None:
        graph.close()
    qp = None
    if not args.no_qpadm:
        del packed, a2, nn, f2, vp
        torch.cuda.empty_cache()
        qp = qpadm_bench(ctx, args.qpadm_models, not args.no_cpu_baseline)
    else:
        del packed, a2, nn, f2, vp
        torch.cuda.empty_cache()
    fs = None
    if not args.no_fstats:
        fs = fstats_bench(ctx, args.fstat_quartets, not args.no_cpu_baseline)
    fe = None
    if not args.no_frontend:
        fe = frontend_bench(ctx, args.frontend_snps, not args.no_cpu_baseline)
    qg = None
    if not args.no_qpgraph:
        qg = qpgraph_bench(ctx, args.qpgraph_graphs, not args.no_cpu_baseline)
    qs = None
    if not args.no_qpgraph_search:
        qs = qpgraph_search_bench(ctx, not args.no_cpu_baseline)

    if cpu and accuracy is not None and "max_rel_err_device_vs_reference" in cpu:
        accuracy["max_rel_err_vs_reference"] = cpu["max_rel_err_device_vs_reference"]
        accuracy["reference_blocks_compared"] = cpu["compared_blocks"]
    out = {
        "metric": METRIC, "value": value, "unit": "pop-pair·SNPs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == _core.STP_PREC_FP64 else "bf16x(split)->f64",
        "data": "synthetic (Balding-Nichols panel generated in HBM)",
        "config": {"workload": f"{args.workload}: {P} populations x {M} SNPs, {n_ind} individuals, "
                               f"{nb} jackknife blocks, decode + f2 ({args.precision})",
                   "P": P, "M": M, "n_block": nb, "n_individuals": n_ind,
                   "precision": args.precision, "parallelism": "single GPU",
                   "graph": graph is not None, "eager_ms_per_step": eager_ms, "stage_timing": stage_source},
        "roofline": roofline, "accuracy": accuracy, "stages": stages, "cpu_baseline": cpu, "qpadm": qp, "qpgraph": qg, "fstats": fs, "frontend": fe,
        "tiers": tiers, "qpgraph_search": qs, "autotune": tune,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
