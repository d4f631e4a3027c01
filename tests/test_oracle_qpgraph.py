"""Pin the restated qpGraph oracle (oracle/cpu_ref_qpgraph.cpp) and the torch FP64
reference of the score (tests/qpgraph_ref.py).

* the ADMIXTOOLS 2 golden fit (f2_qpgraph_9pop.bin with golden_qpgraph_topology.csv) at
  the reference's tolerances (tests/reference/test_qpgraph_parity.cu:94-119): score
  +-1e-4, weights +-1e-5, a tight restart spread; the golden edge lengths;
* bit for bit against the reference's own run_qpgraph and qpgraph_score (oracle/_ref)
  on the golden and on simulated graphs;
* the torch reference against the oracle's score, and the NNLS solution against the
  Karush-Kuhn-Tucker conditions of its scaled problem and scipy's Lawson-Hanson NNLS;
* the product's graph parser (host code of the engine, no GPU needed) against the
  oracle's tables.
"""
import csv
import os

import numpy as np
import pytest

from fixtures import read_raw_fixture

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
POPS = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
        "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]


def golden_edges():
    return [(r["from"], r["to"]) for r in csv.DictReader(open(os.path.join(AT2, "golden_qpgraph_topology.csv")))]


@pytest.fixture(scope="module")
def golden():
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_qpgraph_9pop.bin"))
    return f2, bs, golden_edges()


def simulated(seed, L, A):
    from steppe_amd import synth
    rng = np.random.default_rng(seed)
    edges, theta, drift = synth.random_admixture_graph(rng, L, A)
    f2, bs, names = synth.simulate_graph_f2(edges, theta, drift, n_block=120, rng=rng)
    return f2, bs, edges, names


def test_golden_fit(cref, golden):
    f2, bs, edges = golden
    r = cref.qpgraph_fit(f2, None, bs, edges, POPS)
    assert r["status"] == 0
    assert abs(r["score"] - 80.0674246076313) <= 1e-4
    # pSteppe->aCW is the first-listed parent edge of aCW: theta_0
    assert abs(r["weight"][0] - 0.153483829987482) <= 1e-5
    assert r["restart_spread"] < 1e-4
    gold = {(row["from"], row["to"]): float(row["weight"])
            for row in csv.DictReader(open(os.path.join(AT2, "golden_qpgraph_edges.csv"))) if row["type"] == "edge"}
    dims = cref.qpgraph_model(edges, POPS)
    normal = [e for e in edges if e in gold]
    assert len(normal) == dims["nedge_norm"]
    for k, e in enumerate(normal):
        assert abs(r["edge_length"][k] - gold[e]) <= 1e-6 * max(abs(gold[e]), 1e-3), (e, r["edge_length"][k], gold[e])


def test_bit_identical_to_reference(cref, ref, golden):
    f2, bs, edges = golden
    cases = [(f2, bs, edges, POPS)] + [simulated(s, L, A) for s, L, A in ((1, 8, 1), (2, 10, 2), (3, 12, 3))]
    for f2_, bs_, e_, names in cases:
        a = cref.qpgraph_fit(f2_, None, bs_, e_, names)
        b = ref.qpgraph_fit(f2_, None, bs_, e_, names)
        for k in ("score", "restart_spread"):
            assert np.float64(a[k]).tobytes() == np.float64(b[k]).tobytes(), k
        for k in ("weight", "weight_lo", "weight_hi", "edge_length"):
            assert np.array_equal(a[k], b[k]), k
        f_obs, _, qinv = cref.qpgraph_basis(f2_, None, bs_, e_, names)
        rng = np.random.default_rng(9)
        D = cref.qpgraph_model(e_, names)["nadmix"]
        for _ in range(4):
            th = rng.uniform(0, 1, D)
            for con in (True, False):
                sa, bla, fa = cref.qpgraph_score(e_, names, th, f_obs, qinv, constrained=con)
                sb, blb, fb = ref.qpgraph_score(e_, names, th, f_obs, qinv, constrained=con)
                assert np.float64(sa).tobytes() == np.float64(sb).tobytes()
                assert np.array_equal(bla, blb) and np.array_equal(fa, fb)


def _scaled_system(t, th, f_obs, G, fudge=1e-4):
    import torch
    import qpgraph_ref as qr
    P = qr.pair_products(t, torch.as_tensor(th)).numpy()
    cc = P.T @ G @ P
    q = P.T @ (G @ f_obs)
    C = cc + fudge * np.trace(cc) / cc.shape[0] * np.eye(cc.shape[0])
    s = np.sqrt(np.diag(C))
    return C / np.outer(s, s), q / s, s


def test_torch_reference_and_nnls(cref, golden):
    import torch
    import qpgraph_ref as qr
    from scipy.optimize import nnls
    f2, bs, edges = golden
    cases = [(f2, bs, edges, POPS)] + [simulated(s, L, A) for s, L, A in ((4, 9, 2), (5, 11, 4))]
    for f2_, bs_, e_, names in cases:
        t = cref.qpgraph_tables(e_, names)
        f_obs, _, G = cref.qpgraph_basis(f2_, None, bs_, e_, names)
        rng = np.random.default_rng(3)
        for _ in range(5):
            th = rng.uniform(0.05, 0.95, t["nadmix"])
            s, bl, _ = cref.qpgraph_score(e_, names, th, f_obs, G)
            val, blt = qr.profiled_score(t, torch.as_tensor(th), f_obs, G, bl > 0)
            assert abs(float(val) - s) <= 1e-10 * abs(s)
            assert np.allclose(blt.numpy(), bl, rtol=1e-8, atol=1e-14)
            # KKT of min 1/2 x^T A x - q1^T x, x >= 0, at the scaled solution
            A, q1, sc = _scaled_system(t, th, f_obs, G)
            x = bl * sc
            grad = A @ x - q1
            scale = np.abs(q1).max()
            assert np.all(np.abs(grad[x > 0]) <= 1e-8 * scale)
            assert np.all(grad[x == 0] >= -1e-8 * scale)
            # the same minimiser from scipy's Lawson-Hanson on the Cholesky form
            As = 0.5 * (A + A.T)
            Lc = np.linalg.cholesky(As)
            xs, _ = nnls(Lc.T, np.linalg.solve(Lc, q1))
            assert np.allclose(xs, x, rtol=1e-6, atol=1e-9 * np.abs(x).max())


def test_engine_parser_matches_oracle(cref, golden):
    """The engine's host parser (no device needed) builds the oracle's model."""
    from steppe_amd import qpgraph
    _, _, edges = golden
    cases = [(edges, POPS)] + [simulated(s, L, A)[2:] for s, L, A in ((6, 8, 1), (7, 12, 4))]
    for e_, names in cases:
        t = qpgraph.Topology(e_, names)
        d = cref.qpgraph_model(e_, names)
        assert (t.npop, t.nedge, t.nadmix, t.npair, t.npath, t.base) == (
            d["npop"], d["nedge_norm"], d["nadmix"], d["npair"], d["npath"], d["base_leaf"])
    t = qpgraph.Topology(edges, POPS)
    assert t.admix_from == ["pSteppe"] and t.admix_to == ["aCW"]
    assert t.leaves[0] == "Mbuti"
    with pytest.raises(ValueError):
        qpgraph.Topology([("R", "A"), ("R", "B"), ("S", "B")], ["A", "B"])   # two roots
    with pytest.raises(ValueError):
        qpgraph.Topology([("R", "A"), ("R", "Z")], ["A", "B"])               # leaf not in f2
