"""qpGraph on the device: f3 basis and covariance, batched score, the hand-written
adjoint and the L-BFGS fit, against the oracle, the torch FP64 reference gradient,
central finite differences (h = 1e-6) and the ADMIXTOOLS 2 golden.

Tolerances: f_obs / Q diagonal 1e-12 relative; Qinv 1e-9 of its largest entry; score
1e-9 relative; adjoint vs the torch FP64 gradient 1e-7 relative (plus 1e-9 of the
gradient's scale); adjoint vs central differences (h = 1e-6) of the oracle's score:
1e-4 relative plus 1e-13 s / h absolute, the difference quotient's rounding for a score s
computed to ~1e-14 relative (a small gradient component of a large score is otherwise
below the quotient's noise: at s = 1446, d/dtheta = 3.5e-3 moves by 5e-6 between
h = 1e-5 and 1e-6);
fitted optimum: no worse than the oracle's finite-difference search by more than 1e-7
relative, the golden score within 1e-4 and its weight within 1e-5.
"""
import csv
import os

import numpy as np
import pytest

from fixtures import read_raw_fixture

pytestmark = pytest.mark.gpu

AT2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
FD_NOISE = 1e-13
POPS = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
        "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]


def golden_edges():
    return [(r["from"], r["to"]) for r in csv.DictReader(open(os.path.join(AT2, "golden_qpgraph_topology.csv")))]


def simulated(seed, L, A, n_block=300):
    from steppe_amd import synth
    rng = np.random.default_rng(seed)
    edges, theta, drift = synth.random_admixture_graph(rng, L, A)
    f2, bs, names = synth.simulate_graph_f2(edges, theta, drift, n_block=n_block, rng=rng)
    return f2, bs, edges, names


@pytest.fixture(scope="module")
def cases():
    f2, bs = read_raw_fixture(os.path.join(AT2, "f2_qpgraph_9pop.bin"))
    out = [(f2, bs, golden_edges(), POPS)]
    out += [simulated(s, L, A) for s, L, A in ((11, 8, 1), (12, 10, 2), (13, 12, 3), (14, 12, 4))]
    return out


@pytest.fixture(scope="module")
def gset(cases):
    from steppe_amd import _core, qpgraph
    topos, hs = [], []
    for f2, bs, edges, names in cases:
        h = _core.upload_f2(f2, None, bs, names)
        hs.append(h)
        topos.append(qpgraph.Topology(edges, names))
    return qpgraph.GraphSet(topos, hs)


def test_basis_matches_oracle(cref, cases, gset):
    assert np.all(gset.status == 0)
    for g, (f2, bs, edges, names) in enumerate(cases):
        fo, Q, Qi = cref.qpgraph_basis(f2, None, bs, edges, names)
        f_dev, qd_dev, qi_dev = gset.basis(g)
        assert np.allclose(f_dev, fo, rtol=1e-12, atol=1e-18)
        assert np.allclose(qd_dev, np.diag(Q), rtol=1e-12, atol=0)
        err = np.abs(qi_dev - Qi).max() / np.abs(Qi).max()
        print(f"graph {g}: Qinv max diff / max |Qinv| = {err:.2e}")
        assert err <= 1e-9


def _thetas(D, n, seed):
    return np.random.default_rng(seed).uniform(0.05, 0.95, (n, D))


def test_score_matches_oracle(cref, cases, gset):
    worst = 0.0
    for g, (f2, bs, edges, names) in enumerate(cases):
        fo, _, Qi = cref.qpgraph_basis(f2, None, bs, edges, names)
        D = gset.topologies[g].nadmix
        th = _thetas(D, 6, g)
        for con in (True, False):
            r = gset.eval([g] * len(th), th, mode=0, constrained=con, want_fit=True)
            for k in range(len(th)):
                s, bl, fit = cref.qpgraph_score(edges, names, th[k], fo, Qi, constrained=con)
                rel = abs(r["score"][k] - s) / abs(s)
                worst = max(worst, rel)
                assert rel <= 1e-9, (g, con, r["score"][k], s)
                ne = len(bl)
                assert np.allclose(r["bl"][k, :ne], bl, rtol=1e-7, atol=1e-12 * np.abs(bl).max())
    print(f"score: worst relative difference {worst:.2e}")


def _active_stable(cref, edges, names, th, fo, Qi, d, h):
    a = cref.qpgraph_score(edges, names, th, fo, Qi)[1] > 0
    for sgn in (1, -1):
        t2 = th.copy()
        t2[d] += sgn * h
        if not np.array_equal(cref.qpgraph_score(edges, names, t2, fo, Qi)[1] > 0, a):
            return False
    return True


def test_adjoint_against_torch_and_fd(cref, cases, gset):
    import torch
    import qpgraph_ref as qr
    h = 1e-6
    worst_t, worst_fd, n_fd = 0.0, 0.0, 0
    for g, (f2, bs, edges, names) in enumerate(cases):
        t = cref.qpgraph_tables(edges, names)
        fo, _, Qi = cref.qpgraph_basis(f2, None, bs, edges, names)
        D = t["nadmix"]
        th = _thetas(D, 5, 100 + g)
        r = gset.eval([g] * len(th), th, mode=1)
        for k in range(len(th)):
            gdev = r["grad_theta"][k, :D]
            s, bl, _ = cref.qpgraph_score(edges, names, th[k], fo, Qi)
            tt = torch.tensor(th[k], requires_grad=True)
            val, _ = qr.profiled_score(t, tt, fo, Qi, bl > 0)
            (gref,) = torch.autograd.grad(val, tt)
            gref = gref.numpy()
            scale = np.abs(gref).max() + 1e-300
            err = np.abs(gdev - gref).max() / scale
            worst_t = max(worst_t, err)
            assert err <= 1e-7, (g, k, gdev, gref)
            for d in range(D):
                if not _active_stable(cref, edges, names, th[k], fo, Qi, d, h):
                    continue
                tp, tm = th[k].copy(), th[k].copy()
                tp[d] += h
                tm[d] -= h
                fd = (cref.qpgraph_score(edges, names, tp, fo, Qi)[0] - cref.qpgraph_score(edges, names, tm, fo, Qi)[0]) / (2 * h)
                # 1e-4 relative, plus the difference quotient's own rounding: the score
                # carries ~1e-14 relative error, amplified by s / h
                tol = 1e-4 * abs(fd) + FD_NOISE * s / h
                e = abs(gdev[d] - fd) / tol
                worst_fd = max(worst_fd, e)
                n_fd += 1
                assert e <= 1.0, (g, k, d, gdev[d], fd, tol)
    print(f"adjoint: worst vs torch FP64 {worst_t:.2e}; worst vs central FD {worst_fd:.2e} of tolerance "
          f"over {n_fd} checks")


def test_joint_adjoint(cref, cases, gset):
    import torch
    import qpgraph_ref as qr
    h = 1e-6
    for g, (f2, bs, edges, names) in enumerate(cases):
        t = cref.qpgraph_tables(edges, names)
        fo, _, Qi = cref.qpgraph_basis(f2, None, bs, edges, names)
        D, ne = t["nadmix"], t["nedge_norm"]
        rng = np.random.default_rng(200 + g)
        th = rng.uniform(0.05, 0.95, (3, D))
        bl = rng.exponential(0.01, (3, ne))
        r = gset.eval([g] * 3, th, bl=bl, mode=2)
        for k in range(3):
            tt = torch.tensor(th[k], requires_grad=True)
            bb = torch.tensor(bl[k], requires_grad=True)
            val = qr.joint_score(t, tt, bb, fo, Qi)
            v = float(val.detach())
            assert abs(r["score"][k] - v) <= 1e-10 * abs(v)
            gt, gb = torch.autograd.grad(val, (tt, bb))
            assert np.allclose(r["grad_theta"][k, :D], gt.numpy(), rtol=1e-8, atol=1e-9 * np.abs(gt.numpy()).max())
            assert np.allclose(r["grad_bl"][k, :ne], gb.numpy(), rtol=1e-8, atol=1e-9 * np.abs(gb.numpy()).max())
            e = int(rng.integers(ne))
            bp, bm = bl[k].copy(), bl[k].copy()
            bp[e] += h
            bm[e] -= h
            fd = (float(qr.joint_score(t, tt.detach(), torch.tensor(bp), fo, Qi)) -
                  float(qr.joint_score(t, tt.detach(), torch.tensor(bm), fo, Qi))) / (2 * h)
            assert abs(r["grad_bl"][k, e] - fd) <= 1e-4 * abs(fd) + FD_NOISE * abs(v) / h


def test_autograd_function(cases, gset):
    import torch
    from steppe_amd import qpgraph
    Score = qpgraph.score_function()
    g = 3
    D = gset.topologies[g].nadmix
    th = torch.tensor(_thetas(D, 4, 7), requires_grad=True)
    val = Score.apply(th, gset, np.full(4, g, np.int32))
    (grad,) = torch.autograd.grad(val.sum(), th)
    r = gset.eval([g] * 4, th.detach().numpy(), mode=1)
    assert np.allclose(grad.numpy(), r["grad_theta"][:, :D], rtol=0, atol=0)


def test_fits(cref, cases, gset):
    from steppe_amd import qpgraph
    fits = qpgraph.fit_sets(gset, numstart=10)
    for g, (f2, bs, edges, names) in enumerate(cases):
        o = cref.qpgraph_fit(f2, None, bs, edges, names)
        d = fits[g]
        print(f"graph {g}: L-BFGS {d['score']:.10f} ({d['iterations']} it) vs oracle search {o['score']:.10f}")
        assert d["score"] <= o["score"] * (1 + 1e-7) + 1e-12
    # the golden graph through the single-fit entry point
    from steppe_amd import _core
    f2, bs, edges, names = cases[0]
    h = _core.upload_f2(f2, None, bs, names)
    r = qpgraph.run_qpgraph(h, edges)
    assert r["status"] == "ok"
    assert abs(r["score"] - 80.0674246076313) <= 1e-4
    assert abs(r["weight"][0] - 0.153483829987482) <= 1e-5
    assert r["restart_spread"] < 1e-4
    assert r["admix_from"] == ["pSteppe"] and r["admix_to"] == ["aCW"]


def test_device_fit_entry_points(cref, cases, gset):
    """stp_qpgraph_fit from explicit start points: the returned score is the score of the
    returned theta, and no restart ends above its start; a set built from a given basis
    (the reference fleet's inputs) fits to the same optimum as the f2-built set; a tree
    (no admixture) fits in one evaluation."""
    from steppe_amd import _core, qpgraph, synth
    rng = np.random.default_rng(31)
    ig = np.repeat(np.arange(len(cases), dtype=np.int32), 4)
    x0 = rng.uniform(0.0, 1.0, (len(ig), gset.dmax))
    th, sc, it = gset.fit(ig, x0)
    assert np.all(it >= 0)
    again = gset.eval(ig, th, mode=0)["score"]
    assert np.array_equal(again, sc)
    start = gset.eval(ig, x0, mode=0)["score"]
    assert np.all(sc <= start)
    # the same graphs from the oracle's basis
    basis = []
    for f2, bs, edges, names in cases:
        fo, _, Qi = cref.qpgraph_basis(f2, None, bs, edges, names)
        basis.append((fo, Qi))
    gb = qpgraph.GraphSet(gset.topologies, basis=basis)
    a = gset.fit_fleet()
    b = gb.fit_fleet()
    assert np.all(a["status"] == 0) and np.all(b["status"] == 0)
    rel = np.abs(a["score"] - b["score"]) / np.abs(b["score"])
    print(f"fleet from f2 vs from the oracle's basis: max rel score diff {rel.max():.2e}")
    assert rel.max() <= 1e-8
    # a tree: no admixture weights, one restart
    trng = np.random.default_rng(5)
    edges, theta, drift = synth.random_admixture_graph(trng, 7, 0)
    f2, bs, names = synth.simulate_graph_f2(edges, theta, drift, rng=trng)
    h = _core.upload_f2(f2, None, bs, names)
    tt = qpgraph.Topology(edges, names)
    assert tt.nadmix == 0
    gt = qpgraph.GraphSet([tt], [h])
    r = gt.fit_fleet()
    o = cref.qpgraph_fit(f2, None, bs, edges, names)
    assert r["status"][0] == 0 and r["iterations"][0] == 0 and r["restart_spread"][0] == 0.0
    assert abs(r["score"][0] - o["score"]) <= 1e-9 * abs(o["score"])
    assert np.allclose(r["edge_length"][0, :tt.nedge], o["edge_length"], rtol=1e-7,
                       atol=1e-12 * np.abs(o["edge_length"]).max())
