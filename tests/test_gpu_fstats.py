"""Standalone f-statistics on the device-resident f2 tensor against the oracle and goldens.

* the ADMIXTOOLS 2 goldens (f4, qpDstat, f3, f4-ratio over f2_fit0_9pop.bin) through the
  Python API at the reference's own tolerances (tests/python/test_py_*.py: rtol 1e-6);
* the device against the restated CpuBackend (oracle/cpu_ref_fstats.cpp) on the same
  tensors, for every register width of the block series (50 .. 2500 surviving blocks),
  with dropped blocks, set-missing f4-ratio blocks and degenerate rows;
* the all-combination sweeps against the oracle's sweep (order, filter, top-K across
  several launches), in the direct form and in the Gram form (FP64 MFMA), including
  copied populations whose variances go back to the direct kernel;
* qpWave against the reference's pinned f2-object chi-square values
  (tests/reference/test_qpwave_parity.cu:125-165).

Tolerances (device vs oracle; the device sums in double-double where the reference
sums in long double): est within 1e-12 (|est| + se); se within 1e-10 relative;
z within 1e-10 (1 + |z|); p within 1e-8 relative (device erfc); f4-ratio alpha within
1e-10 (|alpha| + se), se 1e-9 relative.  The Gram form of the f4 sweep forms each
statistic in another operation order: se within 1e-9 relative, est within 1e-9 (|est| +
se) plus 4e-15 of the f2 means it is formed from (GRAM_TOL below).
"""
import csv
import math
import os

import numpy as np
import pytest

import steppe_amd as st
from steppe_amd import _core
from fixtures import random_tensor, read_raw_fixture

pytestmark = pytest.mark.gpu

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
POPS_9 = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
          "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]


def names(P):
    return [f"pop{i}" for i in range(P)]


def _close(got, want, rtol, atol):
    return abs(got - want) <= atol + rtol * abs(want)


@pytest.fixture(scope="module")
def fit0():
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_fit0_9pop.bin"))
    return f2, bs, st.upload_f2(f2, None, bs, POPS_9)


def check_fstat(dev, ora, what):
    est, se, z, p = (np.asarray(dev[k], np.float64) for k in ("est", "se", "z", "p"))
    oe, os_, oz, op = (np.asarray(ora[k], np.float64) for k in ("est", "se", "z", "p"))
    for a, b, k in ((est, oe, "est"), (se, os_, "se"), (z, oz, "z"), (p, op, "p")):
        assert np.array_equal(np.isnan(a), np.isnan(b)), (what, k)
    ok = ~np.isnan(os_)
    scale = np.abs(oe) + np.where(ok, os_, 0.0)
    de = np.abs(est - oe)
    assert np.all((de <= 1e-12 * scale) | np.isnan(oe)), (what, "est", float(np.nanmax(de / np.maximum(scale, 1e-300))))
    assert np.all(np.abs(se[ok] - os_[ok]) <= 1e-10 * os_[ok]), (what, "se")
    zk = ~np.isnan(oz)
    assert np.all(np.abs(z[zk] - oz[zk]) <= 1e-10 * (1.0 + np.abs(oz[zk]))), (what, "z")
    pk = ~np.isnan(op)
    assert np.all(np.abs(p[pk] - op[pk]) <= 1e-8 * op[pk] + 1e-300), (what, "p")
    r = np.nanmax(de / np.maximum(scale, 1e-300)) if de.size else 0.0
    print(f"{what}: {len(oe)} rows, worst |d est| / (|est| + se) = {r:.2e}")


@pytest.mark.parametrize("name,fn,width", [
    ("golden_fit0_f4_readf2.csv", "f4", 4),
    ("golden_fit0_qpdstat_readf2.csv", "qpdstat", 4),
    ("golden_fit0_f3_readf2.csv", "f3", 3),
])
def test_goldens_through_python_api(cref, fit0, name, fn, width):
    f2, bs, h = fit0
    rows = list(csv.DictReader(open(os.path.join(AT2, name))))
    tuples = [tuple(r[f"pop{j + 1}"] for j in range(width)) for r in rows]
    res = getattr(st, fn)(h, tuples)
    assert res.status == "ok" and len(res) == len(rows)
    for i, g in enumerate(rows):
        assert [getattr(res, f"pop{j + 1}")[i] for j in range(width)] == list(tuples[i])
        for k in ("est", "se", "z"):
            assert _close(getattr(res, k)[i], float(g[k]), 1e-6, 1e-9), (name, i, k)
        assert _close(res.p[i], float(g["p"]), 1e-6, 1e-12), (name, i)
    idx = np.array([[POPS_9.index(x) for x in t] for t in tuples], np.int32)
    ora = cref.f4(f2, None, bs, idx) if width == 4 else cref.f3(f2, None, bs, idx)
    check_fstat({k: getattr(res, k) for k in ("est", "se", "z", "p")}, ora, name)


def test_f4ratio_golden(cref, fit0):
    f2, bs, h = fit0
    rows = list(csv.DictReader(open(os.path.join(AT2, "golden_fit0_f4ratio_readf2.csv"))))
    tuples = [tuple(r[f"pop{j + 1}"] for j in range(5)) for r in rows]
    res = st.f4ratio(h, tuples)
    idx = np.array([[POPS_9.index(x) for x in t] for t in tuples], np.int32)
    ora = cref.f4ratio(f2, None, bs, idx)
    for i, g in enumerate(rows):
        for k in ("alpha", "se", "z"):
            assert _close(getattr(res, k)[i], float(g[k]), 1e-6, 1e-9), (i, k)
        assert abs(res.alpha[i] - ora["alpha"][i]) <= 1e-10 * (abs(ora["alpha"][i]) + ora["se"][i])
        assert abs(res.se[i] - ora["se"][i]) <= 1e-9 * ora["se"][i]
    df = st.f4(h, [tuples[0][:4]], as_dataframe=True)
    assert list(df.columns) == ["pop1", "pop2", "pop3", "pop4", "est", "se", "z", "p"]


@pytest.mark.parametrize("nb", [50, 200, 700, 1500, 2500])
def test_fstat_every_series_width(cref, nb):
    """50 / 200 / 700 / 1500 surviving blocks take the 2 / 4 / 16 / 32-per-lane register
    forms, 2500 the streaming form; block 7 is partly missing and dropped."""
    P = 10
    f2, bs = random_tensor(P, nb, 100 + nb)
    vp = np.ones_like(f2)
    vp[2, 5, 7] = vp[5, 2, 7] = 0.0
    h = st.upload_f2(f2, vp, bs, names(P))
    rng = np.random.default_rng(nb)
    q = np.array([rng.choice(P, 4, replace=False) for _ in range(300)], np.int32)
    t = np.array([rng.choice(P, 3, replace=False) for _ in range(300)], np.int32)
    q[0] = [1, 1, 2, 3]     # X == 0 in every block: est 0, se NaN
    check_fstat(st.run_f4(h, q), cref.f4(f2, vp, bs, q), f"f4 nb={nb}")
    check_fstat(st.run_f3(h, t), cref.f3(f2, vp, bs, t), f"f3 nb={nb}")


def test_f4ratio_random_and_setmiss(cref):
    P, nb = 8, 300
    for scale in (0.05, 2e-6):
        f2, bs = random_tensor(P, nb, 7, scale=scale)
        h = st.upload_f2(f2, None, bs, names(f2.shape[0]))
        rng = np.random.default_rng(8)
        tup = np.array([rng.choice(P, 5, replace=False) for _ in range(200)], np.int32)
        tup[0] = [0, 1, 2, 3, 2]        # alpha exactly 1, no spread
        tup[1] = [0, 0, 2, 3, 4]        # every block set missing
        d = st.run_f4ratio(h, tup)
        o = cref.f4ratio(f2, None, bs, tup)
        a, s, z = (np.asarray(d[k]) for k in ("alpha", "se", "z"))
        for x, y in ((a, o["alpha"]), (s, o["se"]), (z, o["z"])):
            assert np.array_equal(np.isnan(x), np.isnan(y))
        ok = ~np.isnan(o["se"])
        assert a[0] == 1.0 and np.isnan(s[0])
        assert np.all(np.abs(a[ok] - o["alpha"][ok]) <= 1e-10 * (np.abs(o["alpha"][ok]) + o["se"][ok]))
        assert np.all(np.abs(s[ok] - o["se"][ok]) <= 1e-9 * o["se"][ok])


def test_no_surviving_block_and_bad_index(cref):
    P, nb = 6, 4
    f2, bs = random_tensor(P, nb, 9)
    vp = np.ones_like(f2)
    for b in range(nb):
        vp[0, 1, b] = vp[1, 0, b] = 0.0
    h = st.upload_f2(f2, vp, bs, names(P))
    d = st.run_f4(h, [[0, 1, 2, 3]])
    assert all(math.isnan(d[k][0]) for k in ("est", "se", "z", "p"))
    h2 = st.upload_f2(f2, None, bs, names(f2.shape[0]))
    with pytest.raises(KeyError):
        st.run_f4(h2, [[0, 1, 2, P]])


DIRECT_TOL = {"est": 1e-12, "se": 1e-10, "z": 2e-10, "cut": 1e-9}
# Gram form: se within 1e-9; est within 1e-9 se plus the rounding floor of the four f2
# means it is formed from (4e-15 of their magnitudes, the level of the reference's own
# block-value rounding); |z| order and the filter cut within 5e-9 (1 + |z|).
GRAM_TOL = {"est": 1e-9, "se": 1e-9, "z": 5e-9, "cut": 5e-9}


def _sweep_compare(dev, ora, k, cut, what, tol=DIRECT_TOL, floor=None):
    """Same survivors in the same order, but rows whose |z| lies within tolerance of the cut
    (min_z or the K-th |z|) may fall on either side of it, and rows whose |z| agree within
    tolerance (copied populations give exact ties in the oracle) in either order."""
    dk = [tuple(int(v) for v in r[:k]) for r in dev["keys"]]
    ok = [tuple(int(v) for v in r[:k]) for r in ora["keys"]]
    zo = {key: abs(z) for key, z in zip(ok, ora["z"])}
    zd = {key: abs(z) for key, z in zip(dk, dev["z"])}
    for key in set(dk) ^ set(ok):
        zz = zo.get(key, zd.get(key))
        assert abs(zz - cut) <= tol["cut"] * (1.0 + cut), (what, key, zz, cut)
    common = [key for key in ok if key in zd]
    seq = [zo[key] for key in dk if key in zo]
    bad = [i for i in range(len(seq) - 1) if seq[i] < seq[i + 1] - tol["z"] * (1.0 + seq[i + 1])]
    assert not bad, (what, bad[:3], [(seq[i], seq[i + 1]) for i in bad[:3]])
    exact = [key for key in dk if key in zo] == common
    az = np.abs(np.asarray(dev["z"]))
    assert np.all(az[:-1] >= az[1:]), what
    oi = {key: i for i, key in enumerate(ok)}
    di = {key: i for i, key in enumerate(dk)}
    worst = 0.0
    for key in common:
        i, j = di[key], oi[key]
        se = ora["se"][j]
        fl = floor(key) if floor else 0.0
        de = abs(dev["est"][i] - ora["est"][j])
        assert de <= tol["est"] * (abs(ora["est"][j]) + se) + fl, (what, key, de, se, fl)
        assert abs(dev["se"][i] - se) <= tol["se"] * se, (what, key)
        assert abs(dev["p"][i] - ora["p"][j]) <= 1e2 * tol["z"] * (1.0 + abs(ora["z"][j])) ** 2 * ora["p"][j] + 1e-300, \
            (what, key)
        worst = max(worst, abs(dev["se"][i] - se) / se)
    print(f"{what}: {len(dk)} survivors, {len(set(dk) ^ set(ok))} at the cut, same order: {exact}, "
          f"worst se rel {worst:.1e}")


def mean_floor(f2, bs):
    """4e-15 times the magnitudes of the four f2 block means an f4 is formed from."""
    w = bs.astype(np.float64)
    A = np.abs((f2 * w).sum(axis=2) / w.sum())

    def fl(key):
        c0, c1, c2, c3 = key
        return 4e-15 * (A[c1, c2] + A[c1, c3] + A[c0, c2] + A[c0, c3])
    return fl


@pytest.mark.parametrize("k", [4, 3])
def test_sweeps_against_oracle(cref, k, monkeypatch):
    P, nb = 14, 400
    f2, bs = random_tensor(P, nb, 200 + k)
    h = st.upload_f2(f2, None, bs, names(f2.shape[0]))
    sub = np.array([13, 2, 7, 0, 11, 5, 9, 1, 4], np.int32)
    for subset in (None, sub):
        for mode, min_z, top_k in (("min_z", 3.0, 0), ("min_z", 0.5, 0), ("top_k", 3.0, 25), ("top_k", 0.0, 9)):
            for chunk in (None, "37"):
                if chunk:
                    monkeypatch.setenv("STEPPE_FSTAT_CHUNK", chunk)
                else:
                    monkeypatch.delenv("STEPPE_FSTAT_CHUNK", raising=False)
                pops = None if subset is None else [int(x) for x in subset]
                d = st.run_fstat_sweep(h, k, pops, min_z=min_z, top_k=top_k, method="direct")
                o = cref.fstat_sweep(f2, None, bs, k, subset=subset, filter=mode, min_z=min_z,
                                     top_k=top_k if top_k > 0 else 100)
                assert d["enumerated"] == o["enumerated"] and d["status"] == "ok"
                cut = min_z
                if mode == "top_k" and len(o["z"]) == top_k:
                    cut = abs(o["z"][-1])
                _sweep_compare(d, o, k, cut, f"k={k} {mode} min_z={min_z} top_k={top_k} chunk={chunk}")


def test_sweep_cap():
    f2, bs = random_tensor(6, 3, 5)
    h = st.upload_f2(f2, None, bs, names(f2.shape[0]))
    d = st.run_fstat_sweep(h, 4, list(range(6)) * 39, min_z=3.0)     # C(234, 4) > 1e8
    assert d["capped"] and d["status"] == "invalid_config" and d["survivors"] == 0


def test_qpwave_reference_values(cref, fit0):
    """M1: left {0,1,2}, M2: left {1,2}; right {3..8} (test_qpwave_parity.cu:125-165)."""
    f2, bs, h = fit0
    right = POPS_9[3:]
    m1 = _core.run_qpwave(h, POPS_9[:3], right)
    assert st.qpwave(h, left=POPS_9[:3], right=right).f4rank == 1
    assert m1["f4rank"] == 1 and m1["est_rank"] == 1 and m1["status"] == "ok"
    assert m1["rankdrop_f4rank"] == [1, 0]
    for got, want in zip(m1["rankdrop_chisq"], [3.95682062790988, 1474.03320584515]):
        assert _close(got, want, 1e-6, 1e-9)
    assert _close(m1["rankdrop_chisqdiff"][0], 1470.07638521724, 1e-6, 1e-9)
    assert math.isnan(m1["rankdrop_chisqdiff"][-1])
    assert m1["rank_chisq"][0] == m1["rankdrop_chisq"][-1] and len(m1["rank_chisq"]) == 2
    m2 = _core.run_qpwave(h, POPS_9[1:3], right)
    assert m2["f4rank"] == 0 and m2["rankdrop_f4rank"] == [0]
    assert _close(m2["rankdrop_chisq"][0], 1401.571655111, 1e-6, 1e-9)
    o = cref.qpadm(f2, None, bs, 0, [1, 2], [3, 4, 5, 6, 7, 8], jackknife=0)
    for got, want in zip(m1["rankdrop_chisq"], o["rankdrop_chisq"]):
        assert abs(got - want) <= 1e-9 * abs(want)


def test_gram_sweep_against_oracle(cref):
    """The Gram form of the f4 sweep (FP64 MFMA block Gram matrices) against the oracle's
    sweep: same survivors, order and values; populations 5 and 6 are exact copies and 7 / 8
    near-copies, whose f4 variances would be formed with heavy cancellation and are
    recomputed by the direct kernel (the exact copies give se = NaN, so never survive even
    at min_z = 0)."""
    P, nb = 36, 300
    f2, bs = random_tensor(P, nb, 301)
    f2[6, :, :] = f2[5, :, :]
    f2[:, 6, :] = f2[:, 5, :]
    f2[5, 6, :] = f2[6, 5, :] = 0.0
    f2[6, 6, :] = 0.0
    rng = np.random.default_rng(302)
    noise = rng.normal(0.0, 1e-9, (P, nb))
    f2[8, :, :] = f2[7, :, :] + noise
    f2[:, 8, :] = f2[8, :, :]
    f2[7, 8, :] = f2[8, 7, :] = np.abs(noise[0])
    f2[8, 8, :] = 0.0
    h = st.upload_f2(f2, None, bs, names(P))
    sub = np.array([35, 7, 2, 8, 19, 0, 5, 30, 6, 11, 24, 13, 1, 28, 17, 3, 22, 9, 33, 14, 26, 4, 31, 12], np.int32)
    for subset in (None, sub):
        pops = None if subset is None else [int(x) for x in subset]
        for mode, min_z, top_k in (("min_z", 3.0, 0), ("min_z", 0.0, 0), ("top_k", 2.0, 40)):
            d = st.run_fstat_sweep(h, 4, pops, min_z=min_z, top_k=top_k, method="gram")
            o = cref.fstat_sweep(f2, None, bs, 4, subset=subset, filter=mode, min_z=min_z,
                                 top_k=top_k if top_k > 0 else 100)
            assert d["enumerated"] == o["enumerated"] and d["status"] == "ok"
            cut = min_z
            if mode == "top_k" and len(o["z"]) == top_k:
                cut = abs(o["z"][-1])
            _sweep_compare(d, o, 4, cut, f"gram {'subset' if pops else 'all'} {mode} min_z={min_z} top_k={top_k}",
                           GRAM_TOL, mean_floor(f2, bs))


def test_gram_sweep_two_row_blocks(cref):
    """140 populations: pairs with c2 >= 128 take two 128-row blocks of S (three items).
    The Gram and direct forms agree on every survivor; a sample is checked against f4."""
    P, nb = 140, 48
    f2, bs = random_tensor(P, nb, 401)
    h = st.upload_f2(f2, None, bs, names(P))
    g = st.run_fstat_sweep(h, 4, None, min_z=6.0, method="gram")
    d = st.run_fstat_sweep(h, 4, None, min_z=6.0, method="direct")
    assert g["enumerated"] == math.comb(P, 4) and g["survivors"] > 1000
    _sweep_compare(g, d, 4, 6.0, "gram vs direct, P=140", GRAM_TOL, mean_floor(f2, bs))
    rows = np.asarray(g["keys"])
    assert rows[:, 3].max() >= 129 and np.any(rows[:, 1] >= 128)
    pick = np.random.default_rng(402).choice(len(rows), 500, replace=False)
    o = cref.f4(f2, None, bs, rows[pick].astype(np.int32))
    fl = mean_floor(f2, bs)
    for i, j in enumerate(pick):
        assert abs(g["est"][j] - o["est"][i]) <= 1e-9 * (abs(o["est"][i]) + o["se"][i]) + fl(tuple(rows[j]))
        assert abs(g["se"][j] - o["se"][i]) <= 1e-9 * o["se"][i]
