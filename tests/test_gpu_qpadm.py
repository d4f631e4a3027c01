"""Batched qpAdm on the device-resident f2 tensor against the oracle and the goldens.

* the ADMIXTOOLS 2 goldens through the device path, at the reference's own
  tolerances (tests/reference/test_qpadm_parity.cu:231-284);
* the device against the restated CpuBackend fit chain (oracle/cpu_ref_qpadm.cpp) on
  the same tensors: weights, chi^2 and p within 1e-9 relative (FP64 path), the
  cancellation-formed SE / z / nested p within 1e-7, dof, ranks, statuses, popdrop
  patterns and feasibility exactly;
* the 84-model rotation as one device batch, a missing-block tensor, a random mixed
  batch of models on a synthetic tensor, and the argument checks.
"""
import json
import os

import numpy as np
import pytest

from fixtures import read_raw_fixture

pytestmark = pytest.mark.gpu

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
FIT_RTOL = 1e-9      # weights, chi^2, p: device vs oracle, relative
FIT_ATOL = 1e-13     # absolute floor for values near zero
# Quantities formed by cancellation get their own stated bounds: the jackknife SE is the
# spread of nb refits around their mean and z = w / SE, so both carry the refits' relative
# error divided by the spread (1e-7); chi^2 differences between adjacent ranks are
# compared against the larger chi^2 (1e-9 of it), and their p-values against 1e-7.
DERIVED_RTOL = 1e-7

EXACT_KEYS = ("status", "dof", "est_rank", "f4rank", "rankdrop_f4rank", "rankdrop_dof",
              "rankdrop_dofdiff", "popdrop_pat", "popdrop_wt", "popdrop_dof", "popdrop_f4rank",
              "popdrop_feasible")
PRIMARY_KEYS = ("weight", "chisq", "p", "rankdrop_chisq", "rankdrop_p", "popdrop_chisq", "popdrop_p")
DERIVED_KEYS = ("se", "z", "rankdrop_p_nested")


def _rel(a, b, scale=None):
    a = np.atleast_1d(np.asarray(a, np.float64))
    b = np.atleast_1d(np.asarray(b, np.float64))
    assert a.shape == b.shape
    both_nan = np.isnan(a) & np.isnan(b)
    d = np.abs(a - b)
    ref = np.abs(b) if scale is None else np.asarray(scale, np.float64)
    return np.where(both_nan, 0.0, d / np.maximum(ref, FIT_ATOL / FIT_RTOL))


def compare(dev, ora, what=""):
    """Per-field worst relative differences; exact fields must agree."""
    for k in EXACT_KEYS:
        assert dev[k] == ora[k], (what, k, dev[k], ora[k])
    out = {}
    for k in PRIMARY_KEYS + DERIVED_KEYS:
        r = _rel(dev[k], ora[k])
        out[k] = float(r.max()) if r.size else 0.0
    # chi^2 differences against the larger chi^2 of the pair
    rc = np.asarray(ora["rankdrop_chisq"], np.float64)
    nxt = np.append(rc[1:], np.nan)
    scale = np.fmax(np.abs(rc), np.abs(nxt))
    r = _rel(dev["rankdrop_chisqdiff"], ora["rankdrop_chisqdiff"], scale)
    out["rankdrop_chisqdiff"] = float(r.max()) if r.size else 0.0
    return out


def check_fields(worst: dict, what: str):
    lines = ", ".join(f"{k} {v:.1e}" for k, v in worst.items() if v > 0)
    print(f"{what}: {lines}")
    for k, v in worst.items():
        bound = DERIVED_RTOL if k in DERIVED_KEYS else FIT_RTOL
        assert v <= bound, (what, k, v, bound)


def merge(acc: dict, new: dict, tag=None, where=None):
    for k, v in new.items():
        if v > acc.get(k, -1.0):
            acc[k] = v
            if where is not None:
                where[k] = tag
    return acc


def _upload(f2, bs, vpair=None, pops=None):
    from steppe_amd import _core
    P = f2.shape[0]
    return _core.upload_f2(f2, vpair, bs, pops or [f"p{i}" for i in range(P)])


def test_fit0_device_matches_golden_and_oracle(cref):
    from steppe_amd import _core
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fit0_9pop.bin"))
    g = json.load(open(os.path.join(AT2, "golden_fit0.json")))["fixture_f2_object_path"]
    h = _upload(f2, bs)
    r = _core.run_qpadm(h, 0, [1, 2], [3, 4, 5, 6, 7, 8])
    assert r["status"] == "ok" and r["dof"] == 4 and r["est_rank"] == 1
    for got, want in zip(r["weight"], g["weight"]):
        assert abs(got - want) <= 1e-6 * abs(want)
    assert abs(r["chisq"] - g["rankdrop"]["chisq"][0]) <= 1e-6 * g["rankdrop"]["chisq"][0]
    assert abs(r["p"] - g["rankdrop"]["p"][0]) <= 1e-3 * g["rankdrop"]["p"][0]
    for got, want in zip(r["se"], g["se"]):
        assert abs(got - want) <= 1e-3 * abs(want)
    assert r["popdrop_pat"] == g["popdrop"]["pat"] and r["popdrop_feasible"] == g["popdrop"]["feasible"]
    check_fields(compare(r, cref.qpadm(f2, None, bs, 0, [1, 2], [3, 4, 5, 6, 7, 8]), "fit0"), "fit0")


def test_missing_block_tensor(cref):
    from steppe_amd import _core
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fitNA.bin"))
    g = json.load(open(os.path.join(AT2, "golden_fitNA.json")))
    vpair = np.where(np.isnan(f2), 0.0, 1.0)
    f2 = np.nan_to_num(f2, nan=0.0)
    h = _upload(f2, bs, vpair)
    right = [3, 4, 5, 6, 7, 8, 9]
    r = _core.run_qpadm(h, 0, [1, 2], right)
    for got, want in zip(r["weight"], g["weights"]["weight"]):
        assert abs(got - want) <= 1e-6 * abs(want)
    check_fields(compare(r, cref.qpadm(f2, vpair, bs, 0, [1, 2], right), "fitNA"), "fitNA")


def test_rotation_batch(cref):
    from steppe_amd import _core
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_rot.bin"))
    g = json.load(open(os.path.join(AT2, "golden_rot.json")))
    order = g["fixture"]["pop_order"]
    h = _upload(f2, bs, pops=order)
    lefts = [m["left"] for m in g["models"]]
    res = _core.run_qpadm_search(h, g["metadata"]["target"], lefts, g["metadata"]["right"])
    assert len(res) == 84
    idx = {n: i for i, n in enumerate(order)}
    right = [idx[n] for n in g["metadata"]["right"]]
    worst = {}
    for k, (m, r) in enumerate(zip(g["models"], res)):
        assert r["model_index"] == k and r["status"] == m["status"]
        wrtol = 1e-5 if m["feasible"] else 1e-4
        for got, want in zip(r["weight"], m["weight"]):
            assert abs(got - want) <= wrtol * abs(want) + 1e-5
        assert r["popdrop_feasible"][0] == m["feasible"]
        o = cref.qpadm(f2, None, bs, idx[g["metadata"]["target"]], [idx[n] for n in m["left"]], right)
        merge(worst, compare(r, o, f"rotation model {k}"))
    check_fields(worst, "rotation (84 models)")


@pytest.fixture(scope="module")
def synthetic_tensor(cref):
    from steppe_amd import synth
    p = synth.scaled_config(P=36, M=24_000, seed=404, block_lo=150, block_hi=500)
    q, v, n = cref.decode_af(p, "sample")
    f2, vp, bs = cref.f2_blocks(q, v, n, p.block_id, p.n_block, n_threads=8)
    return np.asfortranarray(f2), np.asfortranarray(vp), bs


def _random_models(rng, P, count):
    models = []
    for _ in range(count):
        ns = int(rng.integers(2, 6))
        nrt = int(rng.integers(6, 13))
        pick = rng.choice(P, 1 + ns + nrt, replace=False)
        models.append((int(pick[0]), [int(x) for x in pick[1:1 + ns]], [int(x) for x in pick[1 + ns:]]))
    return models


def test_random_batch_against_oracle(cref, synthetic_tensor):
    from steppe_amd import _core
    f2, vp, bs = synthetic_tensor
    P = f2.shape[0]
    h = _upload(f2, bs, vp)
    models = _random_models(np.random.default_rng(7), P, 160)
    res = _core.qpadm_models(h, models)
    worst, where = {}, {}
    for k, ((t, l, r), d) in enumerate(zip(models, res)):
        o = cref.qpadm(f2, vp, bs, t, l, r)
        merge(worst, compare(d, o, f"model {k} (nl={len(l)}, nr={len(r) - 1})"), k, where)
    print("random batch: worst model per field", where)
    check_fields(worst, "random batch (160 models)")


def test_jackknife_policies_and_ranks(cref, synthetic_tensor):
    from steppe_amd import _core
    f2, vp, bs = synthetic_tensor
    h = _upload(f2, bs, vp)
    models = _random_models(np.random.default_rng(11), f2.shape[0], 24)
    for pol, code in (("none", 0), ("feasible", 1)):
        res = _core.qpadm_models(h, models, jackknife=pol)
        for (t, l, r), d in zip(models, res):
            o = cref.qpadm(f2, vp, bs, t, l, r, jackknife=code)
            check_fields(compare(d, o, pol), f"jackknife={pol}")
    # an explicit rank below nl - 1
    res = _core.qpadm_models(h, models[:8], rank=0)
    for (t, l, r), d in zip(models[:8], res):
        o = cref.qpadm(f2, vp, bs, t, l, r, rank=0)
        check_fields(compare(d, o, "rank 0"), "rank 0")


def test_invalid_models_raise():
    from steppe_amd import _core
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fit0_9pop.bin"))
    h = _upload(f2, bs)
    with pytest.raises(ValueError):  # more sources than right populations beyond the base
        _core.run_qpadm(h, 0, [1, 2, 3, 4], [5, 6, 7])
    with pytest.raises(ValueError):
        _core.run_qpadm(h, 0, [], [3, 4, 5])
    with pytest.raises(KeyError):
        _core.run_qpadm(h, "nope", [1, 2], [3, 4, 5, 6])
    # the context stays usable after a rejected call
    assert _core.run_qpadm(h, 0, [1, 2], [3, 4, 5, 6, 7, 8])["status"] == "ok"
