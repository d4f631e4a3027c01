"""The qpGraph topology space (enumeration, shape hash, hill-climb moves) against the
reference's enumerator compiled from its sources (oracle/_ref) and against ADMIXTOOLS 2's
own enumeration (tests/golden/at2/at2_enum_5pop.txt: 105 trees and 1485 one-admixture
graphs on 5 leaves), plus the structural facts any enumeration of rooted binary trees and
one-admixture graphs must satisfy.  Host code only: no GPU.

Also pins the reference-search fixture (tests/golden/qpgraph_search_ref_5pop.csv, made by
tests/golden/make_qpgraph_search_ref.py) that the GPU search test compares with: its
candidates are this enumeration, its argmin is ADMIXTOOLS 2's, and its scores lie within
the reference's documented 2% envelope of ADMIXTOOLS 2's (the reference's
tests/reference/test_qpgraph_search_parity.cu).
"""
import csv
import json
import os
import random

import pytest

from steppe_amd import qpgraph as qg

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
POPS5 = ["Mbuti", "Han", "Iran_GanjDareh_N", "Israel_Natufian", "Czechia_EBA_CordedWare"]


def _double_factorial(k):
    out = 1
    while k > 1:
        out *= k
        k -= 2
    return out


def _at2_enum():
    trees, admix = set(), set()
    with open(os.path.join(GOLD, "at2", "at2_enum_5pop.txt")) as f:
        for line in f:
            if not line.strip() or line.startswith("#"):
                continue
            kind, _, es = line.rstrip("\n").split("\t")[:3]
            edges = [tuple(t.split(">", 1)) for t in es.split(";") if ">" in t]
            (trees if kind == "T" else admix).add(qg.graph_hash(edges))
    return trees, admix


def _at2_scores():
    with open(os.path.join(GOLD, "at2", "at2_scores_5pop.csv")) as f:
        return [(int(r["nadmix"]), int(r["id"]), int(r["hash"]), float(r["score"])) for r in csv.DictReader(f)]


def _ref_fixture():
    path = os.path.join(GOLD, "qpgraph_search_ref_5pop.csv")
    with open(path) as f:
        summary = json.loads(f.readline()[1:])
        rows = [(int(r["nadmix"]), int(r["id"]), int(r["hash"]), float(r["score"]), float(r["restart_spread"]))
                for r in csv.DictReader(f)]
    return summary, rows


@pytest.mark.parametrize("n", [3, 4, 5, 6])
def test_enumeration_matches_reference(ref, n):
    leaves = [f"Pop{i}" for i in range(n)][::-1]
    ours, theirs = qg.enumerate_graphs(leaves, 1), ref.qg_enumerate(leaves, 1)
    assert len(ours) == len(theirs)
    for a, b in zip(ours, theirs):
        assert (a["nadmix"], a["id"], a["hash"], a["edges"]) == (b["nadmix"], b["id"], b["hash"], b["edges"])
    assert qg.enumerate_graphs(leaves, 0) == [c for c in ours if c["nadmix"] == 0]


def test_enumeration_structure():
    # (2n-3)!! rooted binary trees of 2n-2 edges; one-admixture graphs add three edges and
    # one node of in-degree 2; all shapes distinct
    for n in (3, 4, 5, 6):
        leaves = [f"L{i}" for i in range(n)]
        c = qg.enumerate_graphs(leaves, 1)
        trees = [x for x in c if x["nadmix"] == 0]
        assert len(trees) == _double_factorial(2 * n - 3)
        assert len({x["hash"] for x in c}) == len(c)
        for x in c:
            assert len(x["edges"]) == 2 * n - 2 + 3 * x["nadmix"]
            indeg = {}
            for a, b in x["edges"]:
                indeg[b] = indeg.get(b, 0) + 1
            assert sorted(indeg.values()).count(2) == x["nadmix"]
            assert {b for _, b in x["edges"]} - {a for a, _ in x["edges"]} == set(leaves)
        assert [x["id"] for x in trees] == list(range(len(trees)))


def test_enumeration_matches_admixtools():
    at2_trees, at2_admix = _at2_enum()
    assert len(at2_trees) == 105 and len(at2_admix) == 1485
    c = qg.enumerate_graphs(POPS5, 1)
    assert {x["hash"] for x in c if x["nadmix"] == 0} == at2_trees
    assert {x["hash"] for x in c if x["nadmix"] == 1} == at2_admix
    # the AT2 score table is keyed by this enumeration's (nadmix, id, hash), row for row
    rows = _at2_scores()
    assert [(r[0], r[1], r[2]) for r in rows] == [(x["nadmix"], x["id"], x["hash"]) for x in c]


def test_every_candidate_parses():
    for x in qg.enumerate_graphs(POPS5, 1):
        t = qg.Topology(x["edges"], POPS5)
        assert t.nadmix == x["nadmix"] and sorted(t.leaves) == sorted(POPS5)


def test_graph_hash_matches_reference_and_is_a_shape_invariant(ref):
    rng = random.Random(7)
    c = qg.enumerate_graphs(POPS5, 1)
    for x in c[::7]:
        h = qg.graph_hash(x["edges"])
        assert h == ref.qg_graph_hash(x["edges"])
        # renaming internal nodes and reordering edges keep the shape
        inner = {a for a, _ in x["edges"]}
        rename = {v: f"X{rng.randrange(10**6)}_{i}" for i, v in enumerate(sorted(inner))}
        moved = [(rename.get(a, a), rename.get(b, b)) for a, b in x["edges"]]
        rng.shuffle(moved)
        assert qg.graph_hash(moved) == h
    # a different leaf labelling of a non-symmetric tree is a different key
    t = c[0]["edges"]
    swap = {POPS5[0]: POPS5[1], POPS5[1]: POPS5[0]}
    assert qg.graph_hash([(swap.get(a, a), swap.get(b, b)) for a, b in t]) != qg.graph_hash(t)


def test_neighbors_match_reference(ref):
    c = qg.enumerate_graphs(POPS5, 1)
    picks = c[::53] + [c[0], c[104], c[105], c[-1]]
    for x in picks:
        for m in (0, 1):
            ours, theirs = qg.topology_neighbors(x, POPS5, m), ref.qg_neighbors(x, POPS5, m)
            assert [(a["nadmix"], a["hash"], a["edges"]) for a in ours] == \
                   [(b["nadmix"], b["hash"], b["edges"]) for b in theirs]
    # moves cross the admixture boundary both ways
    nb = qg.topology_neighbors(c[0], POPS5, 1)
    assert {a["nadmix"] for a in nb} == {0, 1}
    assert any(a["nadmix"] == 0 for a in qg.topology_neighbors(c[105], POPS5, 1))


def test_reference_search_fixture_is_pinned():
    summary, rows = _ref_fixture()
    c = qg.enumerate_graphs(POPS5, 1)
    assert [(r[0], r[1], r[2]) for r in rows] == [(x["nadmix"], x["id"], x["hash"]) for x in c]
    assert summary["n_trees"] == 105 and summary["n_admix1"] == 1485 and summary["n_candidates"] == 1590
    at2 = {r[2]: r[3] for r in _at2_scores()}
    at2_best = min(at2, key=at2.get)
    assert summary["best_hash"] == at2_best
    assert abs(summary["best_score"] - at2[at2_best]) <= 1e-9 + 1e-6 * at2[at2_best]
    rel = [abs(r[3] - at2[r[2]]) / abs(at2[r[2]]) for r in rows]
    assert max(rel) <= 2e-2
    assert sum(x <= 1e-6 for x in rel) >= 400


def test_search_argument_errors():
    with pytest.raises(ValueError):
        qg.run_qpgraph_search(type("H", (), {"pops": POPS5})(), POPS5[:2])


def test_enumeration_edge_cases(ref):
    # one and two leaves: a single (edge-free / cherry) tree, no admixture graph, as the reference
    for n in (1, 2):
        leaves = [f"L{i}" for i in range(n)]
        ours, theirs = qg.enumerate_graphs(leaves, 1), ref.qg_enumerate(leaves, 1)
        assert [(a["nadmix"], a["hash"], a["edges"]) for a in ours] == \
               [(b["nadmix"], b["hash"], b["edges"]) for b in theirs]
    assert qg.graph_hash([]) == ref.qg_graph_hash([])
    with pytest.raises(ValueError):
        qg.enumerate_graphs([], 1)
    with pytest.raises(ValueError):
        qg.enumerate_graphs([f"L{i}" for i in range(10)], 1)
    with pytest.raises(ValueError):
        qg.topology_neighbors({"edges": [("R", "A"), ("R", "B")], "nadmix": 2}, ["A", "B"], 1)
