"""Pin the restated f-statistics oracle (oracle/cpu_ref_fstats.cpp).

* against the ADMIXTOOLS 2 goldens the reference's own Python tests use
  (tests/python/test_py_{f4,qpdstat,f3,f4ratio}.py: rtol 1e-6, atol 1e-9; p atol 1e-12);
* bit-for-bit against the real reference drivers run_f4 / run_f3 / run_f4ratio
  (oracle/_ref) on the golden fixture, on random tensors and with a dropped block;
* against what the mathematics fixes: the jackknife of a weighted-mean statistic is that
  mean, with variance (1/nb) sum_b w_b / (n - w_b) (X_b - mean)^2; f4 changes sign when
  p1 and p2 swap; an f4-ratio whose numerator and denominator coincide is exactly 1 with
  no spread;
* the sweep restatement against the explicit enumeration of every combination (the
  combinatorial number system order, filter and |z| ordering).
"""
import csv
import itertools
import math
import os

import numpy as np
import pytest

from fixtures import random_tensor, read_raw_fixture

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
POPS_9 = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
          "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]
IDX = {n: i for i, n in enumerate(POPS_9)}


def _golden(name, width):
    rows = list(csv.DictReader(open(os.path.join(AT2, name))))
    idx = np.array([[IDX[r[f"pop{j + 1}"]] for j in range(width)] for r in rows], np.int32)
    return rows, idx


def _close(got, want, rtol, atol):
    return abs(got - want) <= atol + rtol * abs(want)


@pytest.fixture(scope="module")
def fit0():
    return read_raw_fixture(os.path.join(AT2, "f2_fit0_9pop.bin"))


def _bits(a):
    return np.asarray(a, np.float64).view(np.int64)


@pytest.mark.parametrize("name,kind,width", [
    ("golden_fit0_f4_readf2.csv", "f4", 4),
    ("golden_fit0_qpdstat_readf2.csv", "f4", 4),
    ("golden_fit0_f3_readf2.csv", "f3", 3),
])
def test_fstat_matches_admixtools_golden(cref, fit0, name, kind, width):
    f2, bs = fit0
    rows, idx = _golden(name, width)
    r = getattr(cref, kind)(f2, None, bs, idx)
    for i, g in enumerate(rows):
        for k in ("est", "se", "z"):
            assert _close(r[k][i], float(g[k]), 1e-6, 1e-9), (name, i, k, r[k][i], g[k])
        assert _close(r["p"][i], float(g["p"]), 1e-6, 1e-12), (name, i, r["p"][i], g["p"])


def test_f4ratio_matches_admixtools_golden(cref, fit0):
    f2, bs = fit0
    rows, idx = _golden("golden_fit0_f4ratio_readf2.csv", 5)
    r = cref.f4ratio(f2, None, bs, idx)
    for i, g in enumerate(rows):
        for k in ("alpha", "se", "z"):
            assert _close(r[k][i], float(g[k]), 1e-6, 1e-9), (i, k, r[k][i], g[k])


@pytest.mark.parametrize("name,kind,width", [
    ("golden_fit0_f4_readf2.csv", "f4", 4),
    ("golden_fit0_f3_readf2.csv", "f3", 3),
    ("golden_fit0_f4ratio_readf2.csv", "f4ratio", 5),
])
def test_oracle_bit_identical_to_reference_on_golden(cref, ref, fit0, name, kind, width):
    f2, bs = fit0
    _, idx = _golden(name, width)
    a = getattr(cref, kind)(f2, None, bs, idx)
    b = getattr(ref, kind)(f2, None, bs, idx)
    for k in a:
        assert np.array_equal(_bits(a[k]), _bits(b[k])), k


def _random_rows(rng, P, n, width):
    return np.array([rng.choice(P, width, replace=False) for _ in range(n)], np.int32)


def test_oracle_bit_identical_to_reference_random_and_dropped_block(cref, ref):
    P, nb = 7, 60
    f2, bs = random_tensor(P, nb, 3)
    vp = np.ones_like(f2)
    vp[1, 4, 10] = vp[4, 1, 10] = 0.0          # block 10 partly missing: dropped
    vp[:, :, 20] = 0.0                         # block 20 wholly missing: kept
    rng = np.random.default_rng(5)
    for kind, width in (("f4", 4), ("f3", 3), ("f4ratio", 5)):
        idx = _random_rows(rng, P, 40, width)
        for v in (None, vp):
            a = getattr(cref, kind)(f2, v, bs, idx)
            b = getattr(ref, kind)(f2, v, bs, idx)
            for k in a:
                assert np.array_equal(_bits(a[k]), _bits(b[k])), (kind, k)
    # the dropped block is as if it were not there
    keep = [b for b in range(nb) if b != 10]
    idx = _random_rows(rng, P, 30, 4)
    a = cref.f4(f2, vp, bs, idx)
    c = cref.f4(f2[:, :, keep], None, bs[keep], idx)
    for k in a:
        assert np.array_equal(_bits(a[k]), _bits(c[k])), k


def test_jackknife_closed_form(cref):
    """est = sum_b w_b X_b / n; var = (1/nb) sum_b w_b/(n - w_b) (X_b - est)^2."""
    P, nb = 8, 90
    f2, bs = random_tensor(P, nb, 11)
    rng = np.random.default_rng(12)
    for kind, width in (("f4", 4), ("f3", 3)):
        idx = _random_rows(rng, P, 25, width)
        r = getattr(cref, kind)(f2, None, bs, idx)
        w = bs.astype(np.float64)
        n = w.sum()
        for i, q in enumerate(idx):
            if width == 4:
                x = 0.5 * (f2[q[1], q[2]] + f2[q[0], q[3]] - f2[q[0], q[2]] - f2[q[1], q[3]])
            else:
                x = 0.5 * (f2[q[0], q[1]] + f2[q[0], q[2]] - f2[q[1], q[2]])
            est = float(np.sum(w * x) / n)
            var = float(np.sum(w / (n - w) * (x - est) ** 2) / nb)
            assert abs(r["est"][i] - est) <= 1e-12 * (abs(est) + math.sqrt(var))
            assert abs(r["se"][i] ** 2 - var) <= 1e-10 * var
            assert abs(r["p"][i] - math.erfc(abs(r["z"][i]) / math.sqrt(2))) <= 1e-15


def test_f4_antisymmetry_and_f3_symmetry(cref):
    P, nb = 6, 70
    f2, bs = random_tensor(P, nb, 21)
    q = np.array([[0, 1, 2, 3], [1, 0, 2, 3], [2, 3, 0, 1], [0, 2, 4, 5], [2, 0, 4, 5]], np.int32)
    r = cref.f4(f2, None, bs, q)
    for a, b, sign in ((0, 1, -1.0), (0, 2, 1.0), (3, 4, -1.0)):
        assert abs(r["est"][a] - sign * r["est"][b]) <= 1e-13 * r["se"][a]
        assert abs(r["se"][a] - r["se"][b]) <= 1e-11 * r["se"][a]
    t = cref.f3(f2, None, bs, np.array([[0, 1, 2], [0, 2, 1]], np.int32))
    assert abs(t["est"][0] - t["est"][1]) <= 1e-13 * t["se"][0]


def test_f4ratio_identity_and_degenerate(cref, ref):
    P, nb = 6, 50
    f2, bs = random_tensor(P, nb, 31)
    tup = np.array([[0, 1, 2, 3, 2], [0, 0, 2, 3, 4], [0, 1, 2, 3, 4]], np.int32)
    a = cref.f4ratio(f2, None, bs, tup)
    assert a["alpha"][0] == 1.0 and np.isnan(a["se"][0]) and np.isnan(a["z"][0])
    # p1 == p2: every denominator block is 0 < 1e-6, so every block is set missing
    assert np.isnan(a["alpha"][1]) and np.isnan(a["se"][1])
    b = ref.f4ratio(f2, None, bs, tup)
    for k in a:
        assert np.array_equal(_bits(a[k]), _bits(b[k])), k


def test_f4ratio_setmiss_blocks(cref, ref):
    """Denominator blocks below 1e-6 drop out per tuple (small-scale tensor)."""
    P, nb = 6, 80
    f2, bs = random_tensor(P, nb, 41, scale=2e-6)
    rng = np.random.default_rng(42)
    tup = np.array([rng.choice(P, 5, replace=False) for _ in range(30)], np.int32)
    x = 0.5 * (f2[tup[:, 1], tup[:, 4]] + f2[tup[:, 0], tup[:, 3]] - f2[tup[:, 0], tup[:, 4]]
               - f2[tup[:, 1], tup[:, 3]])
    small = np.abs(x) < 1e-6
    assert small.any() and not small.all()
    a = cref.f4ratio(f2, None, bs, tup)
    b = ref.f4ratio(f2, None, bs, tup)
    for k in a:
        assert np.array_equal(_bits(a[k]), _bits(b[k])), k


def test_unrank_is_colex_order(cref):
    for k, n in ((3, 9), (4, 9)):
        combos = sorted(itertools.combinations(range(n), k), key=lambda c: c[::-1])
        for r, c in enumerate(combos):
            assert cref.unrank(r, k) == c
            assert sum(math.comb(c[i], i + 1) for i in range(k)) == r


@pytest.mark.parametrize("k", [3, 4])
def test_sweep_matches_explicit_enumeration(cref, k):
    P, nb = 9, 60
    f2, bs = random_tensor(P, nb, 51 + k)
    sub = np.array([8, 2, 5, 0, 7, 3, 1], np.int32)
    for subset in (None, sub):
        rng_pops = list(range(P)) if subset is None else list(subset)
        combos = sorted(itertools.combinations(range(len(rng_pops)), k), key=lambda c: c[::-1])
        idx = np.array([[rng_pops[c] for c in cmb] for cmb in combos], np.int32)
        full = cref.f4(f2, None, bs, idx) if k == 4 else cref.f3(f2, None, bs, idx)
        az = np.abs(full["z"])
        for filt, min_z, top_k in (("min_z", 3.0, 100), ("min_z", 0.0, 100), ("top_k", 3.0, 7),
                                   ("top_k", 0.0, 5)):
            s = cref.fstat_sweep(f2, None, bs, k, subset=subset, filter=filt, min_z=min_z, top_k=top_k)
            assert s["status"] == 0 and s["enumerated"] == len(combos) and not s["capped"]
            sel = [r for r in range(len(combos)) if az[r] >= min_z]
            sel.sort(key=lambda r: (-az[r], r))
            if filt == "top_k":
                sel = sel[:top_k]
            assert s["survivors"] == len(sel)
            for i, r in enumerate(sel):
                assert list(s["keys"][i][:k]) == list(idx[r])
                assert s["est"][i] == full["est"][r] and s["z"][i] == full["z"][r]
                assert s["p"][i] == full["p"][r]


def test_sweep_cap_and_sure(cref):
    f2, bs = random_tensor(5, 3, 61)
    assert math.comb(230, 4) > 100000000
    s = cref.fstat_sweep(f2, None, bs, 4, subset=np.arange(5, dtype=np.int32).repeat(46), min_z=3.0)
    assert s["capped"] and s["status"] == 5 and s["enumerated"] == math.comb(230, 4)
    s = cref.fstat_sweep(f2, None, bs, 4, subset=np.array([0, 1, 2], np.int32))
    assert s["status"] == 0 and s["enumerated"] == 0 and s["survivors"] == 0
