"""qpGraph on the MI355X engine.

* ``Topology``: a parsed admixture graph (the reference's parse_qpgraph,
  src/core/qpadm/qpgraph_model.cpp:41-243) held by the engine.
* ``GraphSet``: graphs resident on the device with their f3 basis and jackknife
  covariance (computed there once); ``GraphSet.eval`` scores batches of instances and,
  on request, returns the hand-written adjoint (d score / d theta through the
  edge-length NNLS, or d score / d (theta, edge lengths) in the joint form).
* ``QpGraphScore`` / ``QpGraphJointScore``: ``torch.autograd.Function`` wrappers of the
  device score whose backward is the device adjoint.
* ``GraphSet.fit`` / ``GraphSet.fit_fleet`` / ``fit_sets``: projected L-BFGS on the
  adjoint gradient, run to convergence on the device (one workgroup per restart), from
  the reference's splitmix multistart points.
* ``run_qpgraph``: the reference's single-graph fit (bindings/bind_qpgraph.cpp:14-36) with
  the finite-difference coordinate search replaced by that device L-BFGS.
* ``enumerate_graphs`` / ``graph_hash`` / ``topology_neighbors``: the bounded topology
  space and its moves (src/core/qpadm/qpgraph_enumerate.cpp), and ``run_qpgraph_search``:
  every candidate fitted in one device fleet against one shared f3 basis
  (bindings/bind_qpgraph.cpp:41-95, src/core/qpadm/qpgraph_search.cpp).
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _core

_SIG_DONE = False


class SearchOptions(C.Structure):
    """stp_qpgraph_search_options (include/steppe_hip.h)."""
    _fields_ = [("max_nadmix", C.c_int32), ("numstart", C.c_int32), ("maxit", C.c_int32), ("tol", C.c_double),
                ("fudge", C.c_double), ("diag_f3", C.c_double), ("constrained", C.c_int32),
                ("run_heuristic", C.c_int32), ("heuristic_seeds", C.c_int32)]


class SearchSummary(C.Structure):
    """stp_qpgraph_search_summary (include/steppe_hip.h)."""
    _fields_ = [("status", C.c_int32), ("n_trees", C.c_int32), ("n_admix1", C.c_int32),
                ("n_candidates", C.c_int32), ("n_fitted", C.c_int32), ("best", C.c_int32),
                ("best_candidate", C.c_int32), ("best_nadmix", C.c_int32), ("best_hash", C.c_uint64),
                ("best_score", C.c_double), ("second_best_score", C.c_double), ("heuristic_recovered", C.c_int32),
                ("n_seeds", C.c_int32), ("fit_all_wall_ms", C.c_double), ("topologies_per_s", C.c_double)]


class FitOptions(C.Structure):
    """stp_qpgraph_fit_options (include/steppe_hip.h)."""
    _fields_ = [("numstart", C.c_int32), ("maxit", C.c_int32), ("tol", C.c_double), ("fudge", C.c_double),
                ("constrained", C.c_int32)]


def _lib():
    global _SIG_DONE
    L = _core.lib()
    if not _SIG_DONE:
        vp, i32, dp = C.c_void_p, C.c_int32, C.c_double
        sig = {
            "stp_qpgraph_parse": (C.c_int, [i32, vp, vp, i32, vp, C.POINTER(vp)]),
            "stp_qpgraph_topology_free": (C.c_int, [vp]),
            "stp_qpgraph_topology_dims": (C.c_int, [vp, vp]),
            "stp_qpgraph_topology_label": (C.c_int, [vp, i32, i32, C.c_char_p, i32]),
            "stp_qpgraph_set_create": (C.c_int, [vp, i32, vp, vp, dp, C.POINTER(vp)]),
            "stp_qpgraph_set_free": (C.c_int, [vp]),
            "stp_qpgraph_set_status": (C.c_int, [vp, vp]),
            "stp_qpgraph_set_strides": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
            "stp_qpgraph_set_basis": (C.c_int, [vp, vp, i32, vp, vp, vp]),
            "stp_qpgraph_eval": (C.c_int, [vp, vp, i32, vp, vp, vp, i32, dp, i32, vp, vp, vp, vp, vp]),
            "stp_qpgraph_set_create_basis": (C.c_int, [vp, i32, vp, vp, vp, C.POINTER(vp)]),
            "stp_qpgraph_fit": (C.c_int, [vp, vp, i32, vp, vp, C.POINTER(FitOptions), vp, vp, vp]),
            "stp_qpgraph_fit_fleet": (C.c_int, [vp, vp, C.POINTER(FitOptions)] + [vp] * 9),
            "stp_qpgraph_enumerate": (C.c_int, [i32, vp, i32, C.POINTER(vp)]),
            "stp_qpgraph_neighbors": (C.c_int, [i32, vp, vp, i32, C.c_uint64, i32, vp, i32, C.POINTER(vp)]),
            "stp_qpgraph_graph_hash": (C.c_int, [i32, vp, vp, C.POINTER(C.c_uint64)]),
            "stp_qpgraph_list_size": (C.c_int, [vp, C.POINTER(i32)]),
            "stp_qpgraph_list_get": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_uint64),
                                              C.POINTER(i32)]),
            "stp_qpgraph_list_edge": (C.c_int, [vp, i32, i32, C.c_char_p, C.c_char_p, i32]),
            "stp_qpgraph_list_free": (C.c_int, [vp]),
            "stp_qpgraph_search_run": (C.c_int, [vp, vp, i32, vp, i32, vp, C.POINTER(SearchOptions), C.POINTER(vp)]),
            "stp_qpgraph_search_get_summary": (C.c_int, [vp, C.POINTER(SearchSummary)]),
            "stp_qpgraph_search_candidates": (vp, [vp]),
            "stp_qpgraph_search_scores": (C.c_int, [vp, vp, vp, vp]),
            "stp_qpgraph_search_seeds": (C.c_int, [vp, vp]),
            "stp_qpgraph_search_best_fit": (C.c_int, [vp] + [vp] * 8),
            "stp_qpgraph_search_best_label": (C.c_int, [vp, i32, i32, C.c_char_p, i32]),
            "stp_qpgraph_search_free": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _SIG_DONE = True
    return L


def _cstrs(items):
    return (C.c_char_p * max(len(items), 1))(*[str(x).encode() for x in items])


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Topology:
    """A parsed graph; labels follow the reference's QpGraphResult fields."""

    def __init__(self, edges, names):
        edges = [(str(a), str(b)) for a, b in edges]
        if not edges:
            raise ValueError("qpgraph: the graph edge list is empty")
        fr, to, nm = _cstrs([a for a, _ in edges]), _cstrs([b for _, b in edges]), _cstrs(list(names))
        h = C.c_void_p()
        _core.check(_lib().stp_qpgraph_parse(len(edges), fr, to, len(names), nm, C.byref(h)), "qpgraph")
        self._h = h
        self.edges = edges
        d = np.zeros(6, np.int32)
        _core.check(_lib().stp_qpgraph_topology_dims(h, _p(d)))
        self.npop, self.nedge, self.nadmix, self.npair, self.npath, self.base = (int(x) for x in d)
        self._names = {}

    # label lists are read from the engine on first use (a fleet of graphs needs none)
    leaves = property(lambda self: self._label_list(0, self.npop))
    admix_from = property(lambda self: self._label_list(1, self.nadmix))
    admix_to = property(lambda self: self._label_list(2, self.nadmix))
    edge_from = property(lambda self: self._label_list(3, self.nedge))
    edge_to = property(lambda self: self._label_list(4, self.nedge))

    def _label_list(self, kind, n):
        if kind not in self._names:
            self._names[kind] = self._labels(kind, n)
        return list(self._names[kind])

    def _labels(self, kind, n):
        out = []
        buf = C.create_string_buffer(256)
        for i in range(n):
            _core.check(_lib().stp_qpgraph_topology_label(self._h, kind, i, buf, 256))
            out.append(buf.value.decode())
        return out

    @property
    def raw(self):
        return self._h

    def centred_pair(self, k):
        """Leaves (a, b) of f3 basis entry k (base; a <= b over the non-base leaves)."""
        cols = [c for c in range(self.npop - 1) for _ in range(c, self.npop - 1)]
        cols2 = [b for a in range(self.npop - 1) for b in range(a, self.npop - 1)]
        leaf = lambda c: c if c < self.base else c + 1  # noqa: E731
        return self.leaves[leaf(cols[k])], self.leaves[leaf(cols2[k])]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _core._lib is not None:
            _core._lib.stp_qpgraph_topology_free(h)
            self._h = None


class GraphSet:
    """Graphs resident on one device, graph g scored against f2 handle f2s[g]."""

    def __init__(self, topologies, f2s=None, diag_f3=1e-5, basis=None, device=0):
        """Graph g against f2 handle f2s[g] (basis and covariance computed on the device),
        or, with ``basis`` = [(f_obs, Qinv), ...], against a given f3 basis and inverse
        covariance per graph (the reference fleet fit's inputs)."""
        self.topologies = list(topologies)
        if not self.topologies:
            raise ValueError("qpgraph: no graphs")
        n = len(self.topologies)
        tps = (C.c_void_p * n)(*[t.raw for t in self.topologies])
        s = C.c_void_p()
        if basis is None:
            if f2s is None or len(f2s) != n:
                raise ValueError("qpgraph: one f2 handle per topology is required")
            self.f2s = list(f2s)
            dev = f2s[0].device
            if any(h.device != dev for h in f2s):
                raise ValueError("qpgraph: every f2 handle must be on the same device")
            self.ctx = _core.Context.get(dev)
            hs = (C.c_void_p * n)(*[h.raw for h in f2s])
            _core.check(_lib().stp_qpgraph_set_create(self.ctx.h, n, tps, hs, float(diag_f3), C.byref(s)),
                        "qpgraph set")
        else:
            if len(basis) != n:
                raise ValueError("qpgraph: one (f_obs, Qinv) per topology is required")
            fo, qi = [], []
            for t, (f, q) in zip(self.topologies, basis):
                f = np.asarray(f, np.float64).reshape(-1)
                q = np.asarray(q, np.float64)
                if f.size != t.npair or q.shape != (t.npair, t.npair):
                    raise ValueError("qpgraph: basis shapes do not match the graph's npair")
                fo.append(f)
                qi.append(q.reshape(-1))
            fo = np.ascontiguousarray(np.concatenate(fo))
            qi = np.ascontiguousarray(np.concatenate(qi))
            self.f2s = []
            self.ctx = _core.Context.get(device)
            _core.check(_lib().stp_qpgraph_set_create_basis(self.ctx.h, n, tps, _p(fo), _p(qi), C.byref(s)),
                        "qpgraph set")
        self._h = s
        self.n = n
        dm, em, pm = C.c_int32(), C.c_int32(), C.c_int32()
        _core.check(_lib().stp_qpgraph_set_strides(s, C.byref(dm), C.byref(em), C.byref(pm)))
        self.dmax, self.emax, self.pmax = int(dm.value), int(em.value), int(pm.value)
        st = np.zeros(n, np.int32)
        _core.check(_lib().stp_qpgraph_set_status(s, _p(st)))
        self.status = st

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _core._lib is not None:
            _core._lib.stp_qpgraph_set_free(h)
            self._h = None

    def basis(self, g):
        """(f_obs, unridged Q diagonal, Qinv) of graph g, computed on the device."""
        np_ = self.topologies[g].npair
        f, qd, qi = np.zeros(np_), np.zeros(np_), np.zeros(np_ * np_)
        _core.check(_lib().stp_qpgraph_set_basis(self.ctx.h, self._h, int(g), _p(f), _p(qd), _p(qi)))
        return f, qd, qi.reshape(np_, np_)

    def eval(self, inst_graph, theta, bl=None, mode=0, fudge=1e-4, constrained=True, want_fit=False):
        """Score a batch of instances.  theta: (n, dmax); bl: (n, emax) for mode 2.
        Returns a dict with score (n,), grad_theta (n, dmax) for modes 1/2, grad_bl (n, emax)
        for mode 2, bl (n, emax) and, when asked, fit (n, pmax)."""
        ig = np.ascontiguousarray(inst_graph, np.int32)
        n = len(ig)
        th = np.zeros((n, self.dmax))
        if theta is not None:
            t = np.asarray(theta, np.float64).reshape(n, -1)
            th[:, :t.shape[1]] = t
        th = np.ascontiguousarray(th)
        blin = None
        if mode == 2:
            b = np.asarray(bl, np.float64).reshape(n, -1)
            blin = np.zeros((n, self.emax))
            blin[:, :b.shape[1]] = b
            blin = np.ascontiguousarray(blin)
        score = np.zeros(n)
        gth = np.zeros((n, self.dmax)) if mode else None
        gbl = np.zeros((n, self.emax)) if mode == 2 else None
        blo = np.zeros((n, self.emax))
        fit = np.zeros((n, self.pmax)) if want_fit else None
        _core.check(_lib().stp_qpgraph_eval(self.ctx.h, self._h, n, _p(ig), _p(th), _p(blin), int(mode), float(fudge),
                                            1 if constrained else 0, _p(score), _p(gth), _p(gbl), _p(blo), _p(fit)),
                    "qpgraph eval")
        out = {"score": score, "bl": blo}
        if gth is not None:
            out["grad_theta"] = gth
        if gbl is not None:
            out["grad_bl"] = gbl
        if fit is not None:
            out["fit"] = fit
        return out


    @staticmethod
    def _options(numstart=10, maxit=200, tol=1e-9, fudge=1e-4, constrained=True):
        return FitOptions(int(numstart), int(maxit), float(tol), float(fudge), 1 if constrained else 0)

    def fit(self, inst_graph, theta0, maxit=200, tol=1e-9, fudge=1e-4, constrained=True):
        """Device L-BFGS of each instance from its start point.  Returns (theta (n, dmax),
        score (n,), iterations (n,); -1 where the start's edge-length system is singular)."""
        ig = np.ascontiguousarray(inst_graph, np.int32)
        n = len(ig)
        th = np.zeros((n, self.dmax))
        t = np.asarray(theta0, np.float64).reshape(n, -1)
        th[:, :t.shape[1]] = t
        th = np.ascontiguousarray(th)
        out = np.zeros((n, self.dmax))
        score = np.zeros(n)
        iters = np.zeros(n, np.int32)
        opt = self._options(1, maxit, tol, fudge, constrained)
        _core.check(_lib().stp_qpgraph_fit(self.ctx.h, self._h, n, _p(ig), _p(th), C.byref(opt), _p(out), _p(score),
                                           _p(iters)), "qpgraph fit")
        return out, score, iters

    def fit_fleet(self, numstart=10, maxit=200, tol=1e-9, fudge=1e-4, constrained=True):
        """Multistart fit of every graph on the device (stp_qpgraph_fit_fleet)."""
        n = self.n
        r = {"score": np.zeros(n), "restart_spread": np.zeros(n), "theta": np.zeros((n, self.dmax)),
             "theta_lo": np.zeros((n, self.dmax)), "theta_hi": np.zeros((n, self.dmax)),
             "edge_length": np.zeros((n, self.emax)), "f3_fit": np.zeros((n, self.pmax)),
             "iterations": np.zeros(n, np.int32), "status": np.zeros(n, np.int32)}
        opt = self._options(numstart, maxit, tol, fudge, constrained)
        _core.check(_lib().stp_qpgraph_fit_fleet(
            self.ctx.h, self._h, C.byref(opt), *[_p(r[k]) for k in ("score", "restart_spread", "theta", "theta_lo",
                                                                     "theta_hi", "edge_length", "f3_fit",
                                                                     "iterations", "status")]), "qpgraph fleet")
        return r


# ---------------------------------------------------------------------------------------
# torch.autograd wrappers
# ---------------------------------------------------------------------------------------
def _functions():
    import torch

    class QpGraphScore(torch.autograd.Function):
        """score(theta) with the edge lengths re-solved inside (the reference's score);
        backward is the device adjoint through the NNLS active set."""

        @staticmethod
        def forward(ctx, theta, gset, inst_graph, fudge=1e-4, constrained=True):
            r = gset.eval(inst_graph, theta.detach().cpu().double().numpy(), mode=1, fudge=fudge,
                          constrained=constrained)
            d = theta.shape[1]
            ctx.save_for_backward(torch.from_numpy(r["grad_theta"][:, :d].copy()))
            return torch.from_numpy(r["score"].copy()).to(theta.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout.reshape(-1, 1).double().cpu() * g, None, None, None, None

    class QpGraphJointScore(torch.autograd.Function):
        """score(theta, edge lengths) for given edge lengths; backward returns both
        adjoints."""

        @staticmethod
        def forward(ctx, theta, bl, gset, inst_graph, fudge=1e-4):
            r = gset.eval(inst_graph, theta.detach().cpu().double().numpy(), bl=bl.detach().cpu().double().numpy(),
                          mode=2, fudge=fudge)
            ctx.save_for_backward(torch.from_numpy(r["grad_theta"][:, :theta.shape[1]].copy()),
                                  torch.from_numpy(r["grad_bl"][:, :bl.shape[1]].copy()))
            return torch.from_numpy(r["score"].copy())

        @staticmethod
        def backward(ctx, gout):
            gt, gb = ctx.saved_tensors
            g = gout.reshape(-1, 1).double().cpu()
            return g * gt, g * gb, None, None, None

    return QpGraphScore, QpGraphJointScore


def score_function():
    return _functions()[0]


def joint_score_function():
    return _functions()[1]


# ---------------------------------------------------------------------------------------
# multistart fits and the single-graph fit
# ---------------------------------------------------------------------------------------
def fit_sets(gset: GraphSet, numstart=10, fudge=1e-4, constrained=True, maxit=200, tol=1e-9):
    """Fits every graph of the set from ``numstart`` restarts each, all restarts of all
    graphs in one device launch.  Returns per graph (None when its covariance is
    singular): best score, restart spread, best theta, per-coordinate min / max over
    restarts, edge lengths and fitted f3 at the best theta, and the most L-BFGS
    iterations any restart used."""
    r = gset.fit_fleet(numstart=numstart, maxit=maxit, tol=tol, fudge=fudge, constrained=constrained)
    out = []
    for g, t in enumerate(gset.topologies):
        if gset.status[g] != 0:
            out.append(None)
            continue
        d = t.nadmix
        out.append({"score": float(r["score"][g]), "restart_spread": float(r["restart_spread"][g]),
                    "theta": r["theta"][g, :d].copy(), "theta_lo": r["theta_lo"][g, :d].copy(),
                    "theta_hi": r["theta_hi"][g, :d].copy(), "edge_length": r["edge_length"][g, :t.nedge].copy(),
                    "f3_fit": r["f3_fit"][g, :t.npair].copy(), "iterations": int(r["iterations"][g]),
                    "status": int(r["status"][g])})
    return out


def run_qpgraph(f2, edges, numstart=10, fudge=1e-4, diag_f3=1e-5, constrained=True, maxit=200, tol=1e-9) -> dict:
    """Single-graph fit (bindings/bind_qpgraph.cpp:14-36); the reference's result dict
    (bindings/internal/bind_common.hpp:305-323)."""
    if not edges:
        raise ValueError("qpgraph: the graph edge list is empty")
    topo = Topology(edges, f2.pops)
    gset = GraphSet([topo], [f2], diag_f3=diag_f3)
    base = {"leaves": topo.leaves, "admix_from": topo.admix_from, "admix_to": topo.admix_to,
            "edge_from": topo.edge_from, "edge_to": topo.edge_to, "precision": "fp64"}
    if gset.status[0] != 0:
        return dict(base, status="non_spd_covariance", score=0.0, restart_spread=0.0, weight=[], weight_lo=[],
                    weight_hi=[], edge_length=[], worst_residual_z=0.0, worst_pop2="", worst_pop3="")
    fit = fit_sets(gset, numstart=numstart, fudge=fudge, constrained=constrained, maxit=maxit, tol=tol)[0]
    score = fit["score"]
    if fit["status"] != 0 or not np.isfinite(score):
        return dict(base, status="non_spd_covariance", score=score, restart_spread=0.0, weight=[], weight_lo=[],
                    weight_hi=[], edge_length=[], worst_residual_z=0.0, worst_pop2="", worst_pop3="")
    f_obs, qdiag, _ = gset.basis(0)
    fitted = fit["f3_fit"]
    with np.errstate(invalid="ignore", divide="ignore"):
        z = (f_obs - fitted) / np.where(qdiag > 0, np.sqrt(qdiag), np.nan)
    worst, wk = 0.0, -1
    for k, zk in enumerate(z):
        if np.isfinite(zk) and abs(zk) > abs(worst):
            worst, wk = float(zk), k
    p2, p3 = topo.centred_pair(wk) if wk >= 0 else ("", "")
    return dict(base, status="ok", score=score, restart_spread=fit["restart_spread"],
                weight=fit["theta"].tolist(), weight_lo=fit["theta_lo"].tolist(), weight_hi=fit["theta_hi"].tolist(),
                edge_length=fit["edge_length"].tolist(), worst_residual_z=worst, worst_pop2=p2, worst_pop3=p3)


# ---------------------------------------------------------------------------------------
# topology enumeration and search
# ---------------------------------------------------------------------------------------
def _read_list(h):
    """stp_qpgraph_list -> [{"nadmix", "id", "hash", "edges"}], in list order."""
    L = _lib()
    n = C.c_int32()
    _core.check(L.stp_qpgraph_list_size(h, C.byref(n)))
    out = []
    na, gid, ne = C.c_int32(), C.c_int32(), C.c_int32()
    hv = C.c_uint64()
    fb, tb = C.create_string_buffer(256), C.create_string_buffer(256)
    for i in range(n.value):
        _core.check(L.stp_qpgraph_list_get(h, i, C.byref(na), C.byref(gid), C.byref(hv), C.byref(ne)))
        edges = []
        for k in range(ne.value):
            _core.check(L.stp_qpgraph_list_edge(h, i, k, fb, tb, 256))
            edges.append((fb.value.decode(), tb.value.decode()))
        out.append({"nadmix": int(na.value), "id": int(gid.value), "hash": int(hv.value), "edges": edges})
    return out


def _taking_list(fn, *args):
    L = _lib()
    h = C.c_void_p()
    _core.check(fn(*args, C.byref(h)), "qpgraph")
    try:
        return _read_list(h)
    finally:
        L.stp_qpgraph_list_free(h)


def enumerate_graphs(leaves, max_nadmix=1):
    """Every rooted tree on ``leaves`` then every distinct one-admixture graph
    (enumerate_bounded_space, src/core/qpadm/qpgraph_enumerate.cpp:229-236): dicts with
    nadmix, id, hash and edges, in the reference's order."""
    leaves = [str(x) for x in leaves]
    return _taking_list(_lib().stp_qpgraph_enumerate, len(leaves), _cstrs(leaves), int(max_nadmix))


def graph_hash(edges):
    """Shape hash of an edge list (graph_hash, src/core/qpadm/qpgraph_enumerate.cpp:183-202)."""
    edges = [(str(a), str(b)) for a, b in edges]
    h = C.c_uint64()
    _core.check(_lib().stp_qpgraph_graph_hash(len(edges), _cstrs([a for a, _ in edges]), _cstrs([b for _, b in edges]),
                                              C.byref(h)), "graph_hash")
    return int(h.value)


def topology_neighbors(graph, leaves, max_nadmix=1):
    """Hill-climb moves from ``graph`` (a dict from enumerate_graphs, or one with edges,
    nadmix and hash) (topology_neighbors, src/core/qpadm/qpgraph_enumerate.cpp:383-407)."""
    edges = [(str(a), str(b)) for a, b in graph["edges"]]
    leaves = [str(x) for x in leaves]
    hv = graph.get("hash")
    if hv is None:
        hv = graph_hash(edges)
    return _taking_list(_lib().stp_qpgraph_neighbors, len(edges), _cstrs([a for a, _ in edges]),
                        _cstrs([b for _, b in edges]), int(graph["nadmix"]), C.c_uint64(int(hv)), len(leaves),
                        _cstrs(leaves), int(max_nadmix))


def run_qpgraph_search(f2, pops, max_nadmix=1, numstart=10, fudge=1e-4, diag_f3=1e-5, constrained=True,
                       run_heuristic=True, *, heuristic_seeds=8, maxit=200, tol=1e-9, details=False) -> dict:
    """The topology search (bindings/bind_qpgraph.cpp:41-95): every candidate on ``pops``
    fitted in one device fleet against one f3 basis f3(pops[0]; a, b) shared by all
    candidates.  Returns the reference's dict (n_trees, n_admix1, n_candidates,
    best_score, second_best_score, best_nadmix, best_hash, best_edges,
    heuristic_recovered, fit_all_wall_ms, topologies_per_s, cand_nadmix, cand_hash,
    cand_score, cand_restart_spread) plus cand_id and seed_hashes; ``details=True`` adds
    cand_edges and best_fit (the best graph's run_qpgraph-style result)."""
    pops = [str(p) for p in pops]
    if len(pops) < 3:
        raise ValueError("qpgraph-search: need >= 3 population labels")
    L = _lib()
    names = list(f2.pops)
    opt = SearchOptions(int(max_nadmix), int(numstart), int(maxit), float(tol), float(fudge), float(diag_f3),
                        1 if constrained else 0, 1 if run_heuristic else 0, int(heuristic_seeds))
    ctx = _core.Context.get(f2.device)
    h = C.c_void_p()
    _core.check(L.stp_qpgraph_search_run(ctx.h, f2.raw, len(names), _cstrs(names), len(pops), _cstrs(pops),
                                         C.byref(opt), C.byref(h)), "qpgraph-search")
    try:
        sm = SearchSummary()
        _core.check(L.stp_qpgraph_search_get_summary(h, C.byref(sm)))
        cands = _read_list(L.stp_qpgraph_search_candidates(h))
        nf = int(sm.n_fitted)
        idx = np.zeros(nf, np.int32)
        score = np.zeros(nf)
        spread = np.zeros(nf)
        _core.check(L.stp_qpgraph_search_scores(h, _p(idx), _p(score), _p(spread)))
        seeds = np.zeros(max(int(sm.n_seeds), 1), np.uint64)
        _core.check(L.stp_qpgraph_search_seeds(h, _p(seeds)))
        d = {"status": "ok" if sm.status == 0 else "non_spd_covariance",
             "n_trees": int(sm.n_trees), "n_admix1": int(sm.n_admix1), "n_candidates": int(sm.n_candidates),
             "best_score": float(sm.best_score), "second_best_score": float(sm.second_best_score),
             "best_nadmix": int(sm.best_nadmix), "best_hash": int(sm.best_hash),
             "best_edges": list(cands[sm.best_candidate]["edges"]) if sm.best_candidate >= 0 else [],
             "heuristic_recovered": bool(sm.heuristic_recovered), "fit_all_wall_ms": float(sm.fit_all_wall_ms),
             "topologies_per_s": float(sm.topologies_per_s),
             "cand_nadmix": [cands[i]["nadmix"] for i in idx], "cand_hash": [cands[i]["hash"] for i in idx],
             "cand_id": [cands[i]["id"] for i in idx], "cand_score": score.tolist(),
             "cand_restart_spread": spread.tolist(), "seed_hashes": [int(x) for x in seeds[:int(sm.n_seeds)]]}
        if details:
            d["cand_edges"] = [cands[i]["edges"] for i in idx]
            if sm.best >= 0:
                dims = np.zeros(6, np.int32)
                _core.check(L.stp_qpgraph_search_best_fit(h, _p(dims), None, None, None, None, None, None, None))
                npop, ne, na, npair = (int(x) for x in dims[:4])
                w, lo, hi = np.zeros(max(na, 1)), np.zeros(max(na, 1)), np.zeros(max(na, 1))
                bl, fit, fo, qd = np.zeros(ne), np.zeros(npair), np.zeros(npair), np.zeros(npair)
                _core.check(L.stp_qpgraph_search_best_fit(h, _p(dims), _p(w), _p(lo), _p(hi), _p(bl), _p(fit),
                                                          _p(fo), _p(qd)))
                buf = C.create_string_buffer(256)

                def labels(kind, n):
                    out = []
                    for i in range(n):
                        _core.check(L.stp_qpgraph_search_best_label(h, kind, i, buf, 256))
                        out.append(buf.value.decode())
                    return out
                d["best_fit"] = {"score": float(sm.best_score), "weight": w[:na].tolist(),
                                 "weight_lo": lo[:na].tolist(), "weight_hi": hi[:na].tolist(),
                                 "edge_length": bl.tolist(), "f3_fit": fit.tolist(), "f3_obs": fo.tolist(),
                                 "leaves": labels(0, npop), "admix_from": labels(1, na), "admix_to": labels(2, na),
                                 "edge_from": labels(3, ne), "edge_to": labels(4, ne)}
        return d
    finally:
        L.stp_qpgraph_search_free(h)


class _CallableModule(type(sys.modules[__name__])):
    """This module, callable as the reference facade's ``qpgraph(f2, edges, ...)``
    (bindings/steppe/__init__.py:696): returns a QpGraphResult."""

    def __call__(self, f2, edges, **kw):
        from . import _qpgraph_fit
        return _qpgraph_fit(f2, edges, **kw)


sys.modules[__name__].__class__ = _CallableModule
