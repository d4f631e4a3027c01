"""qpAdm models beyond the batched kernel (sources x (right - 1) > 64 or right - 1 > 32):
the large-model kernels (qpadm_large.hip) against the oracle's fit chain.

The reference fits these on its large path (src/core/qpadm/model_search.cpp:85-106) with
the same chain as every other model, so the bar is the batched kernel's: weights, chi^2
and p within 1e-9 relative, SE / z / nested p within 1e-7, ranks, dofs, statuses and
popdrop patterns exactly (test_gpu_qpadm's compare / check_fields).  Shapes:
(sources, right populations) = (3, 20) on the batched kernel, then (5, 14) (m = 65, the
first past its envelope), (3, 34) (right - 1 = 33), (5, 30) and (8, 40) on the large path,
in one mixed batch whose results come back in input order.
"""
import concurrent.futures as cf
import sys

import numpy as np
import pytest

from test_gpu_qpadm import DERIVED_KEYS, _upload, check_fields, compare, merge

pytestmark = pytest.mark.gpu

SHAPES = [(3, 20), (5, 14), (3, 34), (5, 30), (8, 40)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tensor(cref):
    """A 60-population tensor over 24 blocks from a synthetic panel (the oracle's decode and
    f2), small enough in blocks for the oracle's dense fits of the largest shapes."""
    from steppe_amd import synth
    p = synth.scaled_config(P=60, M=12_000, seed=505, block_lo=300, block_hi=700)
    q, v, n = cref.decode_af(p, "sample")
    f2, vp, bs = cref.f2_blocks(q, v, n, p.block_id, p.n_block, n_threads=8)
    _say(f"[tensor] P=60, {p.n_block} blocks")
    return np.asfortranarray(f2), np.asfortranarray(vp), bs


def _models(rng, P):
    out = []
    for ns, nright in SHAPES:
        pick = rng.choice(P, 1 + ns + nright, replace=False)
        out.append((int(pick[0]), [int(x) for x in pick[1:1 + ns]], [int(x) for x in pick[1 + ns:]]))
    return out


def test_large_models_against_oracle(cref, tensor):
    from steppe_amd import _core
    f2, vp, bs = tensor
    h = _upload(f2, bs, vp)
    models = _models(np.random.default_rng(17), f2.shape[0])
    res = _core.qpadm_models(h, models, jackknife="none")
    with cf.ThreadPoolExecutor(len(models)) as ex:
        outs = list(ex.map(lambda m: cref.qpadm(f2, vp, bs, *m, jackknife=0), models))
    worst = {}
    for (t, l, r), d, o in zip(models, res, outs):
        _say(f"  nl={len(l)} right={len(r)}: status {d['status']}, weights {np.round(d['weight'], 6)}")
        merge(worst, compare(d, o, f"nl={len(l)} right={len(r)}"))
    for k in DERIVED_KEYS:
        worst.pop(k, None)
    check_fields(worst, "large models (no jackknife)")


def test_large_models_jackknife_against_oracle(cref, tensor):
    # leave-one-block-out standard errors on the shapes whose dense oracle fits finish in
    # seconds per refit
    from steppe_amd import _core
    f2, vp, bs = tensor
    h = _upload(f2, bs, vp)
    models = [m for m, (ns, nright) in zip(_models(np.random.default_rng(19), f2.shape[0]), SHAPES)
              if ns * (nright - 1) <= 150]
    res = _core.qpadm_models(h, models)
    with cf.ThreadPoolExecutor(len(models)) as ex:
        outs = list(ex.map(lambda m: cref.qpadm(f2, vp, bs, *m), models))
    worst, where = {}, {}
    for (t, l, r), d, o in zip(models, res, outs):
        _say(f"  nl={len(l)} right={len(r)}: se {np.round(d['se'], 6)} (oracle {np.round(o['se'], 6)})")
        merge(worst, compare(d, o, f"nl={len(l)} right={len(r)} (jackknife)"), (len(l), len(r)), where)
    _say(f"  worst model per field: {where}")
    check_fields(worst, "large models (jackknife all)")


def test_large_model_policies_and_rank(cref, tensor):
    from steppe_amd import _core
    f2, vp, bs = tensor
    h = _upload(f2, bs, vp)
    model = _models(np.random.default_rng(23), f2.shape[0])[1]   # (5, 14)
    for pol, code in (("feasible", 1),):
        d = _core.qpadm_models(h, [model], jackknife=pol)[0]
        check_fields(compare(d, cref.qpadm(f2, vp, bs, *model, jackknife=code), pol), f"jackknife={pol}")
    d = _core.qpadm_models(h, [model], rank=1, jackknife="none")[0]
    o = cref.qpadm(f2, vp, bs, *model, rank=1, jackknife=0)
    w = compare(d, o, "rank 1")
    for k in DERIVED_KEYS:
        w.pop(k, None)
    check_fields(w, "rank 1")


def test_oversized_model_rejected(tensor):
    from steppe_amd import _core
    f2, vp, bs = tensor
    h = _upload(f2, bs, vp)
    with pytest.raises(ValueError, match="too large"):
        # 8 sources x 129 right beyond the base exceeds the large path's right-set limit
        big = np.zeros((140, 140, 2))
        hb = _upload(big + 0.01 * (1 - np.eye(140))[:, :, None], np.array([100, 100], np.int32))
        _core.run_qpadm(hb, 0, list(range(1, 9)), list(range(9, 139)))
    assert _core.run_qpadm(h, 0, [1, 2], [3, 4, 5, 6, 7, 8])["status"] in ("ok", "chisq_undefined")


def test_sources_one_more_than_right(cref, tensor):
    # sources = right populations beyond the base + 1: the reference's seed still finds its
    # nl - 1 right singular vectors, the full model has 0 degrees of freedom (status
    # chisq_undefined) and its weights are defined
    from steppe_amd import _core
    f2, vp, bs = tensor
    h = _upload(f2, bs, vp)
    rng = np.random.default_rng(29)
    models = []
    for ns in (2, 4, 6):
        pick = rng.choice(f2.shape[0], 1 + ns + ns, replace=False)
        models.append((int(pick[0]), [int(x) for x in pick[1:1 + ns]], [int(x) for x in pick[1 + ns:]]))
    res = _core.qpadm_models(h, models)
    worst = {}
    for m, d in zip(models, res):
        o = cref.qpadm(f2, vp, bs, *m)
        assert d["status"] == o["status"] == "chisq_undefined" and d["dof"] == 0
        merge(worst, compare(d, o, f"nl={len(m[1])} right={len(m[2])}"))
    check_fields(worst, "sources = right - 1 + 1")
