"""Parity at the full sizes of the large workloads, on the device.

The small-panel tests pin every kernel against the oracle; these run the same paths at
the sizes where 64-bit offsets, per-die work queues over hundreds of blocks, 1000-row
tiles and the qpAdm launch grouping come into play:

* c3 (1000 populations x 1.2M SNPs, LogNormal blocks): FP64 and split-BF16 tensors from
  the same device counts; split vs FP64 over the whole tensor within 1e-9 relative
  (floor 1e-6), vpair bit-equal between the forms and equal to V V^T of the counts on
  every block, symmetry on every block; the FP64 tensor against the oracle on all 1000
  populations for the first, a middle and the last block (1e-12, floor 1e-6, vpair
  exact); the device counts against the oracle's decode on those columns (bit-exact)
  and the device ploidy against the oracle's detection.
* c2 (200 populations x 1.2M SNPs, FP64): the oracle on 10 blocks spread over the
  tensor, symmetry and vpair = V V^T on all blocks.
* c4 (10,000 random models on the c2 tensor): property checks on every result (status,
  sum of weights 1, dof = nr - nl + 1, p in [0, 1], popdrop row 0 = the full model), the
  oracle on 8 spread models with leave-one-block-out SEs and 64 without (the qpAdm
  tolerances of test_gpu_qpadm).
* c5 (5,000 random graphs, 8-12 leaves, 1-4 admixture events, 300 blocks each): every
  fleet fit's theta in [0, 1], its score equal to the device score at its theta, fits
  from one random start per graph no worse than that start, and the oracle's
  finite-difference search on 9 spread graphs no better than the device L-BFGS (1e-7).
"""
import concurrent.futures as cf
import ctypes as C
import sys

import numpy as np
import pytest
import torch

from steppe_amd import _core, synth
from test_gpu_qpadm import DERIVED_KEYS, compare, check_fields, merge

pytestmark = pytest.mark.gpu

F2_RTOL = 1e-12       # FP64 path vs the oracle
SPLIT_RTOL = 1e-9     # split-BF16 path vs the FP64 path and the oracle
FLOOR = 1e-6          # absolute floor on |reference| for the relative error
ORACLE_THREADS = 16


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Extracted:
    """Decoded counts and f2 tensors of a workload, all in device memory."""

    def __init__(self, ctx, name, precisions):
        self.wp = synth.baseline_workload(name)
        self.dp = synth.device_panel(ctx, self.wp)
        P, M = self.wp["P"], self.wp["M"]
        self.P, self.M, self.nb = P, M, self.dp.n_block
        self.ld = (M + 63) // 64 * 64
        self.tile, self._keep = _core.make_tile(None, M, self.dp.pop_offsets, None, detect=True,
                                                bytes_per_record=self.dp.bytes_per_record,
                                                n_individuals=self.dp.n_ind, device_ptr=self.dp.packed.data_ptr())
        L = _core.lib()
        self.ploidy = np.zeros(self.dp.n_ind, np.int32)
        _core.check(L.stp_detect_ploidy(ctx.h, C.byref(self.tile), _core._ptr(self.ploidy)), "ploidy")
        self.a2 = torch.empty(P * self.ld, dtype=torch.int16, device="cuda")
        self.n = torch.empty(P * self.ld, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        _core.check(L.stp_decode_counts(ctx.h, C.byref(self.tile), C.c_void_p(self.a2.data_ptr()),
                                        C.c_void_p(self.n.data_ptr()), self.ld), "decode")
        self.f2, self.vp = {}, {}
        for prec in precisions:
            f2 = torch.empty(P * P * self.nb, dtype=torch.float64, device="cuda")
            vp = torch.empty_like(f2)
            torch.cuda.synchronize()
            _core.check(L.stp_f2_blocks_from_counts(
                ctx.h, C.c_void_p(self.a2.data_ptr()), C.c_void_p(self.n.data_ptr()), self.ld, P, M,
                _core._ptr(self.dp.block_id), self.nb, prec, C.c_void_p(f2.data_ptr()),
                C.c_void_p(vp.data_ptr())), "f2")
            # [nb, j, i]: element (i, j, b) at i + P j + P^2 b
            self.f2[prec] = f2.view(self.nb, P, P)
            self.vp[prec] = vp.view(self.nb, P, P)
        self.bounds = np.concatenate([[0], np.cumsum(self.wp["block_sizes"])]).astype(np.int64)

    def counts(self, s0, s1):
        """Device counts of columns [s0, s1) as host (a2, n) int arrays [P, s1 - s0]."""
        a2 = self.a2.view(self.P, self.ld)[:, s0:s1].to(torch.int32) & 0xFFFF
        n = self.n.view(self.P, self.ld)[:, s0:s1].to(torch.int32) & 0xFFFF
        return a2.cpu().numpy(), n.cpu().numpy()


def _max_rel_dev(a, r, floor=FLOOR, chunk=64):
    """max |a - r| / max(|r|, floor) over two device tensors [nb, P, P], blockwise."""
    worst = 0.0
    for b0 in range(0, a.shape[0], chunk):
        x, y = a[b0:b0 + chunk], r[b0:b0 + chunk]
        e = ((x - y).abs() / y.abs().clamp_min(floor)).max().item()
        worst = max(worst, e)
    return worst


def _check_vpair_and_symmetry(ex, prec):
    """vpair = V V^T of the device counts on every block (fp32 matmul of 0/1 masks is
    exact below 2^24) and both tensors symmetric on every block."""
    f2, vp = ex.f2[prec], ex.vp[prec]
    nmat = ex.n.view(ex.P, ex.ld)
    for b in range(ex.nb):
        s0, s1 = int(ex.bounds[b]), int(ex.bounds[b + 1])
        v = (nmat[:, s0:s1] != 0).to(torch.float32)
        want = (v @ v.T.contiguous()).to(torch.float64)
        assert torch.equal(vp[b], want), f"vpair block {b} differs from V V^T of the counts"
        assert torch.equal(f2[b], f2[b].T), f"f2 block {b} is not symmetric"
        assert torch.equal(vp[b], vp[b].T), f"vpair block {b} is not symmetric"


def _oracle_blocks(ex, cref, blocks, precisions):
    """The oracle's decode and f2 on the columns of each listed block, against the device
    counts (bit-exact) and the device tensors of each precision."""
    worst = {p: 0.0 for p in precisions}
    for b in blocks:
        s0, s1 = int(ex.bounds[b]), int(ex.bounds[b + 1])
        panel, lead = synth.host_subpanel(ex.dp, s0, s1, ex.ploidy)
        q, v, n = cref.decode_af(panel, "sample")
        a2d, nd = ex.counts(s0, s1)
        q_o, v_o, n_o = q[:, lead:], v[:, lead:], n[:, lead:]
        assert np.array_equal(nd.astype(np.float64), n_o), f"block {b}: n differs from the oracle's decode"
        assert np.array_equal(nd > 0, v_o > 0), f"block {b}: V differs"
        with np.errstate(invalid="ignore", divide="ignore"):
            qd = np.where(nd > 0, a2d / (2.0 * nd), 0.0)
        assert np.array_equal(qd, q_o), f"block {b}: q = a2 / 2n differs from the oracle's ac / n"
        fr, vr, br = cref.f2_blocks(q, v, n, panel.block_id, 2, n_threads=ORACLE_THREADS)
        assert br[1] == s1 - s0
        fr_b = torch.from_numpy(np.ascontiguousarray(fr[:, :, 1].T))   # [j, i]
        vr_b = torch.from_numpy(np.ascontiguousarray(vr[:, :, 1].T))
        for prec in precisions:
            fd = ex.f2[prec][b].cpu()
            vd = ex.vp[prec][b].cpu()
            assert torch.equal(vd, vr_b), f"block {b}: vpair differs from the oracle ({prec})"
            e = ((fd - fr_b).abs() / fr_b.abs().clamp_min(FLOOR)).max().item()
            worst[prec] = max(worst[prec], e)
        _say(f"  block {b} (SNPs {s0}..{s1}, {s1 - s0}): max rel err "
             + ", ".join(f"{_core.PRECISION_NAMES.get(p, p)} {worst[p]:.2e}" for p in precisions))
    return worst


# ---------------------------------------------------------------------------------------
# c3
# ---------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c3(engine):
    _say("[c3] 1000 populations x 1.2M SNPs: panel, decode, FP64 and split-BF16 f2")
    ex = Extracted(engine, "c3", (_core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16))
    _say(f"[c3] n_block={ex.nb}, n_individuals={ex.dp.n_ind}")
    yield ex
    del ex
    torch.cuda.empty_cache()


def test_c3_split_vs_fp64_whole_tensor(c3):
    F, S = _core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16
    assert torch.equal(c3.vp[F], c3.vp[S]), "vpair differs between the FP64 and split forms"
    err = _max_rel_dev(c3.f2[S], c3.f2[F])
    _say(f"[c3] split vs FP64 over {c3.nb} x {c3.P} x {c3.P}: max rel err (floor {FLOOR}) {err:.3e}")
    assert err <= SPLIT_RTOL


def test_c3_vpair_and_symmetry_every_block(c3):
    _check_vpair_and_symmetry(c3, _core.STP_PREC_SPLIT_BF16)
    _check_vpair_and_symmetry(c3, _core.STP_PREC_FP64)


def test_c3_ploidy_against_oracle(c3, cref):
    panel, _ = synth.host_subpanel(c3.dp, 0, min(1000, c3.M), c3.ploidy)
    want = cref.detect_ploidy(panel)
    assert np.array_equal(c3.ploidy, want)
    assert set(np.unique(c3.ploidy)) <= {1, 2}


def test_c3_first_middle_last_block_against_oracle(c3, cref):
    F, S = _core.STP_PREC_FP64, _core.STP_PREC_SPLIT_BF16
    worst = _oracle_blocks(c3, cref, [0, c3.nb // 2, c3.nb - 1], (F, S))
    assert worst[F] <= F2_RTOL
    assert worst[S] <= SPLIT_RTOL


# ---------------------------------------------------------------------------------------
# c2 and c4
# ---------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def c2(engine):
    _say("[c2] 200 populations x 1.2M SNPs: panel, decode, FP64 f2")
    ex = Extracted(engine, "c2", (_core.STP_PREC_FP64,))
    _say(f"[c2] n_block={ex.nb}")
    yield ex
    del ex
    torch.cuda.empty_cache()


def test_c2_blocks_against_oracle(c2, cref):
    F = _core.STP_PREC_FP64
    blocks = sorted(set(np.linspace(0, c2.nb - 1, 10).round().astype(int).tolist()))
    worst = _oracle_blocks(c2, cref, blocks, (F,))
    assert worst[F] <= F2_RTOL


def test_c2_vpair_and_symmetry_every_block(c2):
    _check_vpair_and_symmetry(c2, _core.STP_PREC_FP64)


@pytest.fixture(scope="module")
def c2_handle(c2):
    """The c2 tensor as a resident f2 handle (an upload of the device tensor's host copy)."""
    F = _core.STP_PREC_FP64
    f2 = c2.f2[F].permute(2, 1, 0).cpu().numpy()       # (i, j, b) view, F-contiguous
    vp = c2.vp[F].permute(2, 1, 0).cpu().numpy()
    bs = np.asarray(c2.wp["block_sizes"], np.int32)
    h = _core.upload_f2(f2, vp, bs, [f"p{i}" for i in range(c2.P)])
    return h, np.asfortranarray(f2), np.asfortranarray(vp), bs


def _qpadm_properties(models, recs):
    n_ok = 0
    for k, ((t, l, r), d) in enumerate(zip(models, recs)):
        nl, nr = len(l), len(r) - 1
        assert d["model_index"] == k
        assert d["status"] in ("ok", "rank_deficient", "non_spd_covariance", "chisq_undefined"), (k, d["status"])
        if d["status"] != "ok":
            continue
        n_ok += 1
        w = np.asarray(d["weight"])
        assert len(w) == nl and np.all(np.isfinite(w)), k
        assert abs(w.sum() - 1.0) <= 1e-9, (k, w.sum())
        assert d["dof"] == nr - nl + 1, (k, d["dof"], nl, nr)
        assert 0.0 <= d["p"] <= 1.0 and d["chisq"] >= 0.0, (k, d["p"], d["chisq"])
        assert len(d["popdrop_pat"]) == nl + 1
        assert d["popdrop_pat"][0] == "0" * nl
        assert d["popdrop_chisq"][0] == d["chisq"] and d["popdrop_p"][0] == d["p"], k
        assert d["popdrop_dof"][0] == d["dof"], k
        assert d["popdrop_feasible"][0] == bool(np.all((w >= 0.0) & (w <= 1.0))), k
        se = np.asarray(d["se"])
        assert np.all(np.isfinite(se) & (se > 0.0)), (k, se)
    return n_ok


def test_c4_ten_thousand_models(c2_handle, cref):
    h, f2, vp, bs = c2_handle
    models = synth.random_qpadm_models(np.random.default_rng(synth.SEED + 4), h.P, 10_000)
    _say(f"[c4] {len(models)} models on the c2 tensor ({h.P} populations, {h.n_block} blocks)")
    recs = _core.qpadm_models(h, models)
    n_ok = _qpadm_properties(models, recs)
    _say(f"[c4] {n_ok} of {len(models)} ok; properties hold on every result")
    assert n_ok >= 0.99 * len(models)

    # the oracle on spread models: 8 with leave-one-block-out SEs, 64 without
    full = np.linspace(0, len(models) - 1, 8).round().astype(int).tolist()
    worst, where = {}, {}
    with cf.ThreadPoolExecutor(8) as ex:
        outs = list(ex.map(lambda k: cref.qpadm(f2, vp, bs, *models[k]), full))
    for k, o in zip(full, outs):
        merge(worst, compare(recs[k], o, f"c4 model {k}"), k, where)
    _say(f"[c4] jackknife=all, worst model per field {where}")
    check_fields(worst, "c4 (8 models, jackknife all)")

    quick = np.linspace(0, len(models) - 1, 64).round().astype(int).tolist()
    sub = [models[k] for k in quick]
    recs_n = _core.qpadm_models(h, sub, jackknife="none")
    with cf.ThreadPoolExecutor(16) as ex:
        outs = list(ex.map(lambda m: cref.qpadm(f2, vp, bs, *m, jackknife=0), sub))
    worst = {}
    for k, d, o in zip(quick, recs_n, outs):
        merge(worst, compare(d, o, f"c4 model {k} (no jackknife)"))
        # the rank sweep and weights do not depend on the jackknife policy
        assert d["weight"] == recs[k]["weight"] and d["chisq"] == recs[k]["chisq"], k
    for kk in DERIVED_KEYS:
        worst.pop(kk, None)
    check_fields(worst, "c4 (64 models, jackknife none)")


# ---------------------------------------------------------------------------------------
# c5
# ---------------------------------------------------------------------------------------
def test_c5_five_thousand_graphs(engine, cref):
    from steppe_amd import qpgraph
    graphs = synth.random_graph_cases(np.random.default_rng(synth.SEED + 5), 5000)
    _say(f"[c5] {len(graphs)} graphs simulated")
    handles = [_core.upload_f2(f2, None, bs, names) for _, f2, bs, names in graphs]
    gset = qpgraph.GraphSet([qpgraph.Topology(e, names) for e, _, _, names in graphs], handles)
    assert np.all(np.asarray(gset.status) == 0)
    fits = qpgraph.fit_sets(gset, numstart=10)
    n = len(graphs)
    ig = np.arange(n, dtype=np.int32)
    theta = np.zeros((n, gset.dmax))
    for g, f in enumerate(fits):
        assert f is not None and f["status"] == 0, g
        d = len(f["theta"])
        assert np.all((f["theta"] >= 0.0) & (f["theta"] <= 1.0)), g
        assert np.all(f["theta_lo"] <= f["theta"] + 1e-15) and np.all(f["theta"] <= f["theta_hi"] + 1e-15), g
        theta[g, :d] = f["theta"]
    score = np.array([f["score"] for f in fits])
    again = gset.eval(ig, theta, mode=0)["score"]
    rel = np.abs(again - score) / np.abs(score)
    _say(f"[c5] fleet: best score = device score at its theta to {rel.max():.1e}; "
         f"iterations mean {np.mean([f['iterations'] for f in fits]):.1f}")
    assert rel.max() <= 1e-12

    # one random start per graph: the fit never ends above its start
    x0 = np.random.default_rng(55).uniform(0.0, 1.0, (n, gset.dmax))
    th, sc, it = gset.fit(ig, x0)
    start = gset.eval(ig, x0, mode=0)["score"]
    assert np.all(it >= 0) and np.all(sc <= start)
    # the fit kernel's score at its returned theta is the score kernel's, up to the
    # rounding of its own evaluation order (the inner NNLS can end on the other side of a
    # near-tie in its active set, which moves the score in the 12th digit)
    again = gset.eval(ig, th, mode=0)["score"]
    dev = np.abs(again - sc) / np.abs(sc)
    _say(f"[c5] fit-kernel score vs score kernel at the same theta: max rel {dev.max():.1e}, "
         f"{int(np.sum(dev > 0))} of {n} not bit-equal")
    assert dev.max() <= 1e-10

    # the oracle's finite-difference search on spread graphs
    pick = np.linspace(0, n - 1, 9).round().astype(int).tolist()
    with cf.ThreadPoolExecutor(9) as ex:
        outs = list(ex.map(lambda g: cref.qpgraph_fit(graphs[g][1], None, graphs[g][2], graphs[g][0],
                                                      graphs[g][3]), pick))
    for g, o in zip(pick, outs):
        _say(f"  graph {g}: device {score[g]:.10f} vs oracle search {o['score']:.10f}")
        assert score[g] <= o["score"] * (1 + 1e-7) + 1e-12, g
