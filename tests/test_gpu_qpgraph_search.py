"""The qpGraph topology search on the device against ADMIXTOOLS 2 and the reference.

The AT2 qpGraph fixture (f2_qpgraph_9pop.bin) searched over the reference's five
populations: 105 trees and 1485 one-admixture graphs, every candidate fitted in one device
fleet.  Bars (the reference's own, tests/reference/test_qpgraph_search_parity.cu, applied
to this search): the argmin is ADMIXTOOLS 2's graph and its score matches AT2 at rtol 1e-6;
every candidate lies within the 2% envelope of its AT2 score and at least 400 match to
1e-6; against the reference's CPU search (tests/golden/qpgraph_search_ref_5pop.csv) the
argmin and its score agree to 1e-6, at least 99% of the candidates score no worse than
1e-6 relative and none worse than 1e-5 (the device fits from the reference's start points
plus grid seeds reach the same or a lower minimum; the score is piecewise smooth in the
admixture weights, with kinks where the NNLS active set changes, and two optimizers can
stop at neighbouring kink minima a few 1e-6 apart), and the hill-climb recovers the
global best.  The best graph's fit equals run_qpgraph on its edges.
"""
import csv
import json
import os
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (first import on a fresh box is slow; keep it out of test timeouts)

import steppe_amd as st
from steppe_amd import _core
from steppe_amd import qpgraph as qg
from fixtures import read_raw_fixture

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
         "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]
POPS = ["Mbuti", "Han", "Iran_GanjDareh_N", "Israel_Natufian", "Czechia_EBA_CordedWare"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def handle():
    f2, bs = read_raw_fixture(os.path.join(GOLD, "at2", "f2_qpgraph_9pop.bin"))
    return _core.upload_f2(f2, None, bs, NAMES)


@pytest.fixture(scope="module")
def search(handle):
    _say("[search] 5 populations, max_nadmix 1, numstart 10, hill-climb on")
    return qg.run_qpgraph_search(handle, POPS, details=True)


def test_counts_and_order(search):
    d = search
    assert (d["n_trees"], d["n_admix1"], d["n_candidates"]) == (105, 1485, 1590)
    c = qg.enumerate_graphs(POPS, 1)
    assert d["cand_hash"] == [x["hash"] for x in c] and d["cand_id"] == [x["id"] for x in c]
    _say(f"fit of all candidates: {d['fit_all_wall_ms']:.1f} ms ({d['topologies_per_s']:.0f} topologies/s)")


def test_matches_admixtools(search):
    d = search
    with open(os.path.join(GOLD, "at2", "at2_scores_5pop.csv")) as f:
        at2 = {int(r["hash"]): float(r["score"]) for r in csv.DictReader(f)}
    best = min(at2, key=at2.get)
    assert d["best_hash"] == best and d["best_nadmix"] == 1
    assert abs(d["best_score"] - at2[best]) <= 1e-9 + 1e-6 * at2[best]
    rel = np.array([abs(s - at2[h]) / abs(at2[h]) for h, s in zip(d["cand_hash"], d["cand_score"])])
    _say(f"vs AT2: max rel {rel.max():.2e}, median {np.median(rel):.2e}, {int((rel <= 1e-6).sum())} within 1e-6")
    assert rel.max() <= 2e-2
    assert int((rel <= 1e-6).sum()) >= 400


def test_matches_reference_search(search):
    d = search
    with open(os.path.join(GOLD, "qpgraph_search_ref_5pop.csv")) as f:
        summary = json.loads(f.readline()[1:])
        ref = {int(r["hash"]): float(r["score"]) for r in csv.DictReader(f)}
    assert d["best_hash"] == summary["best_hash"]
    assert abs(d["best_score"] - summary["best_score"]) <= 1e-9 + 1e-6 * summary["best_score"]
    worse = np.array([(s - ref[h]) / abs(ref[h]) for h, s in zip(d["cand_hash"], d["cand_score"])])
    _say(f"vs reference search: worst excess {worse.max():.2e}, best gain {-worse.min():.2e}, "
         f"{int((np.abs(worse) <= 1e-6).sum())} equal to 1e-6")
    assert worse.max() <= 1e-5
    assert (worse <= 1e-6).mean() >= 0.99


def test_second_best_and_hill_climb(search):
    d = search
    s = sorted(d["cand_score"])
    assert d["best_score"] == s[0] and d["second_best_score"] == s[1]
    assert d["heuristic_recovered"] and len(d["seed_hashes"]) == 8
    assert d["best_hash"] in d["seed_hashes"]


def test_best_fit_equals_single_graph_fit(search, handle):
    d = search
    r = qg.run_qpgraph(handle, d["best_edges"])
    bf = d["best_fit"]
    assert abs(r["score"] - d["best_score"]) <= 1e-9 * max(1.0, abs(r["score"]))
    assert np.allclose(r["weight"], bf["weight"], rtol=0, atol=1e-7)
    assert r["leaves"] == bf["leaves"] or sorted(r["leaves"]) == sorted(bf["leaves"])


def test_trees_only_and_facade(handle):
    r = st.qpgraph_search(handle, pops=POPS[:4], max_nadmix=0, run_heuristic=False)
    assert r.n_trees == 15 and r.n_admix1 == 0 and r.n_candidates == 15
    assert r.candidates["score"].min() == r.best_score and r.argmin == int(np.argmin(r.candidates["score"]))


def test_errors(handle):
    with pytest.raises(ValueError, match="not an f2 population"):
        qg.run_qpgraph_search(handle, ["Mbuti", "Han", "Nowhere"])
    with pytest.raises(ValueError, match="twice"):
        qg.run_qpgraph_search(handle, ["Mbuti", "Han", "Mbuti"])
