"""Pin the restated qpAdm oracle (oracle/cpu_ref_qpadm.cpp).

* against the ADMIXTOOLS 2 goldens committed by the reference, at the reference's
  own tolerances (tests/reference/test_qpadm_parity.cu:231-284): weights and chi2
  rtol 1e-6, SE / z / p rtol 1e-3, dof / f4rank / popdrop patterns exact;
* bit-for-bit against the real reference driver (oracle/_ref) on the same
  fixtures, including the 84-model rotation and the missing-block case.
"""
import json
import os

import numpy as np
import pytest

from fixtures import read_raw_fixture

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")


def _close(a, b, rtol, atol=0.0):
    return abs(a - b) <= max(rtol * abs(b), atol)


def _bits_equal_dict(a, b):
    for k in a:
        x, y = a[k], b[k]
        if isinstance(x, list) and x and isinstance(x[0], float):
            xa, ya = np.asarray(x), np.asarray(y)
            if not np.array_equal(xa.view(np.int64), ya.view(np.int64)):
                return k
        elif isinstance(x, float):
            if np.float64(x).view(np.int64) != np.float64(y).view(np.int64):
                return k
        elif x != y:
            return k
    return None


@pytest.fixture(scope="module")
def fit0():
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fit0_9pop.bin"))
    g = json.load(open(os.path.join(AT2, "golden_fit0.json")))["fixture_f2_object_path"]
    return f2, bs, g


def test_fit0_matches_admixtools_golden(cref, fit0):
    f2, bs, g = fit0
    r = cref.qpadm(f2, None, bs, 0, [1, 2], [3, 4, 5, 6, 7, 8])
    assert r["status"] == "ok" and r["dof"] == 4 and r["est_rank"] == 1
    for got, want in zip(r["weight"], g["weight"]):
        assert _close(got, want, 1e-6)
    assert _close(r["chisq"], g["rankdrop"]["chisq"][0], 1e-6)
    assert _close(r["p"], g["rankdrop"]["p"][0], 1e-3)
    for got, want in zip(r["se"], g["se"]):
        assert _close(got, want, 1e-3)
    for got, want in zip(r["z"], g["z"]):
        assert _close(got, want, 1e-3)
    rd = g["rankdrop"]
    assert r["rankdrop_f4rank"] == rd["f4rank"] and r["rankdrop_dof"] == rd["dof"]
    for got, want in zip(r["rankdrop_chisq"], rd["chisq"]):
        assert _close(got, want, 1e-6)
    assert r["rankdrop_dofdiff"][0] == rd["dofdiff"][0]
    assert _close(r["rankdrop_chisqdiff"][0], rd["chisqdiff"][0], 1e-6)
    pd = g["popdrop"]
    assert r["popdrop_pat"] == pd["pat"] and r["popdrop_wt"] == pd["wt"]
    assert r["popdrop_dof"] == pd["dof"] and r["popdrop_f4rank"] == pd["f4rank"]
    assert r["popdrop_feasible"] == pd["feasible"]
    for got, want in zip(r["popdrop_chisq"], pd["chisq"]):
        assert _close(got, want, 1e-6)
    for got, want in zip(r["popdrop_p"], pd["p"]):
        assert _close(got, want, 1e-3, atol=1e-300)


def test_fit0_bit_identical_to_reference(cref, ref, fit0):
    f2, bs, _ = fit0
    a = cref.qpadm(f2, None, bs, 0, [1, 2], [3, 4, 5, 6, 7, 8])
    b = ref.qpadm(f2, None, bs, 0, [1, 2], [3, 4, 5, 6, 7, 8])
    assert _bits_equal_dict(a, b) is None


def test_missing_block_is_dropped(cref, ref):
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fitNA.bin"))
    g = json.load(open(os.path.join(AT2, "golden_fitNA.json")))
    vpair = np.where(np.isnan(f2), 0.0, 1.0)
    f2 = np.nan_to_num(f2, nan=0.0)
    right = [3, 4, 5, 6, 7, 8, 9]
    r = cref.qpadm(f2, vpair, bs, 0, [1, 2], right)
    for got, want in zip(r["weight"], g["weights"]["weight"]):
        assert _close(got, want, 1e-6)
    for got, want in zip(r["se"], g["weights"]["se"]):
        assert _close(got, want, 1e-3)
    assert r["popdrop_dof"] == g["popdrop"]["dof"]
    for got, want in zip(r["popdrop_chisq"], g["popdrop"]["chisq"]):
        assert _close(got, want, 1e-6)
    assert _bits_equal_dict(r, ref.qpadm(f2, vpair, bs, 0, [1, 2], right)) is None


def test_rotation_84_models(cref, ref):
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_rot.bin"))
    g = json.load(open(os.path.join(AT2, "golden_rot.json")))
    order = g["fixture"]["pop_order"]
    idx = {n: i for i, n in enumerate(order)}
    right = [idx[n] for n in g["metadata"]["right"]]
    target = idx[g["metadata"]["target"]]
    assert len(g["models"]) == 84
    for k, m in enumerate(g["models"]):
        left = [idx[n] for n in m["left"]]
        r = cref.qpadm(f2, None, bs, target, left, right)
        assert r["status"] == m["status"]
        # The reference's tiers for this set (test_qpadm_rotation.cu:433-447,728-750):
        # weights rtol 1e-5 (feasible) / 1e-4 (infeasible extrapolations) + atol 1e-5,
        # p rtol 1e-3 + atol 1e-6.
        wrtol = 1e-5 if m["feasible"] else 1e-4
        for got, want in zip(r["weight"], m["weight"]):
            assert abs(got - want) <= wrtol * abs(want) + 1e-5, (k, got, want)
        # SE: rtol 1e-3 on feasible models.  The infeasible extrapolations (weights
        # like [171, -0.6, -170], SE up to ~3e4) have jackknife SEs that are not
        # meaningful and differ between implementations; they are not compared.
        if m["feasible"]:
            for got, want in zip(r["se"], m["se"]):
                assert abs(got - want) <= 1e-3 * abs(want) + 1e-5, (k, got, want)
        assert abs(r["p"] - m["p"]) <= 1e-3 * abs(m["p"]) + 1e-6
        # AT2's per-model f4rank is the fitted rank len(left)-1, which the result
        # carries as est_rank and popdrop_f4rank[0] (test_qpadm_rotation.cu:456-466).
        assert r["est_rank"] == m["f4rank"] and r["popdrop_f4rank"][0] == m["f4rank"]
        assert r["popdrop_feasible"][0] == m["feasible"]
        if k % 12 == 0:
            assert _bits_equal_dict(r, ref.qpadm(f2, None, bs, target, left, right)) is None
