"""qpGraph on the MI355X engine.

* ``Topology``: a parsed admixture graph (the reference's parse_qpgraph,
  src/core/qpadm/qpgraph_model.cpp:41-243) held by the engine.
* ``GraphSet``: graphs resident on the device with their f3 basis and jackknife
  covariance (computed there once); ``GraphSet.eval`` scores batches of instances and,
  on request, returns the hand-written adjoint (d score / d theta through the
  edge-length NNLS, or d score / d (theta, edge lengths) in the joint form).
* ``QpGraphScore`` / ``QpGraphJointScore``: ``torch.autograd.Function`` wrappers of the
  device score whose backward is the device adjoint.
* ``lbfgs_box``: a batched projected L-BFGS over [0, 1]^D driven by one device call per
  evaluation round for all instances.
* ``run_qpgraph``: the reference's single-graph fit (bindings/bind_qpgraph.cpp:14-36) with
  the finite-difference coordinate search replaced by L-BFGS on the adjoint gradient,
  started from the reference's splitmix multistart points.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _core

_SIG_DONE = False


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
        self.leaves = self._labels(0, self.npop)
        self.admix_from = self._labels(1, self.nadmix)
        self.admix_to = self._labels(2, self.nadmix)
        self.edge_from = self._labels(3, self.nedge)
        self.edge_to = self._labels(4, self.nedge)

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

    def __init__(self, topologies, f2s, diag_f3=1e-5):
        if len(topologies) != len(f2s) or not topologies:
            raise ValueError("qpgraph: one f2 handle per topology is required")
        self.topologies = list(topologies)
        self.f2s = list(f2s)
        dev = f2s[0].device
        if any(h.device != dev for h in f2s):
            raise ValueError("qpgraph: every f2 handle must be on the same device")
        self.ctx = _core.Context.get(dev)
        n = len(topologies)
        tps = (C.c_void_p * n)(*[t.raw for t in topologies])
        hs = (C.c_void_p * n)(*[h.raw for h in f2s])
        s = C.c_void_p()
        _core.check(_lib().stp_qpgraph_set_create(self.ctx.h, n, tps, hs, float(diag_f3), C.byref(s)), "qpgraph set")
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
# batched projected L-BFGS on [0, 1]^D
# ---------------------------------------------------------------------------------------
def lbfgs_box(fun, x0, lo=0.0, hi=1.0, m=8, max_iter=200, gtol=1e-9, xtol=1e-12, ftol=1e-14,
              max_backtrack=30, c1=1e-4):
    """Minimises n independent box-constrained problems at once.  ``fun(x, idx)`` returns
    (f, g) for the rows idx of x (one device call per round).  Per instance: the search
    direction is the two-loop L-BFGS direction on the free variables (a variable at a
    bound whose gradient points outward is held), the step is projected onto the box and
    accepted by Armijo backtracking.  Stops an instance when its projected gradient is
    below gtol, its step below xtol, or its relative decrease below ftol.  Returns
    (x, f, iterations)."""
    x = np.clip(np.array(x0, dtype=np.float64), lo, hi)
    n, D = x.shape
    f, g = fun(x, np.arange(n))
    S = np.zeros((n, m, D))
    Y = np.zeros((n, m, D))
    rho = np.zeros((n, m))
    hist = np.zeros(n, np.int64)
    active = np.ones(n, bool)
    iters = np.zeros(n, np.int64)
    for _ in range(max_iter):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        xi, gi = x[idx], g[idx]
        held = ((xi <= lo) & (gi > 0)) | ((xi >= hi) & (gi < 0))
        pg = np.where(held, 0.0, gi)
        done = np.abs(pg).max(axis=1) <= gtol
        # two-loop recursion on the free variables
        q = pg.copy()
        k = hist[idx]
        alpha = np.zeros((idx.size, m))
        for j in range(m - 1, -1, -1):           # newest (j = m-1) to oldest
            use = (m - 1 - j) < np.minimum(k, m)  # the entry's age is m-1-j
            slot = (k - 1 - (m - 1 - j)) % m
            s = S[idx, slot] * ~held
            y = Y[idx, slot] * ~held
            r = rho[idx, slot]
            a = np.where(use, r * np.einsum("ij,ij->i", s, q), 0.0)
            alpha[:, j] = a
            q = q - a[:, None] * y
        last = (k - 1) % m
        sy = np.einsum("ij,ij->i", S[idx, last] * ~held, Y[idx, last] * ~held)
        yy = np.einsum("ij,ij->i", Y[idx, last] * ~held, Y[idx, last] * ~held)
        gamma = np.where((k > 0) & (yy > 0), sy / np.where(yy > 0, yy, 1.0), 1.0)
        gamma = np.where(gamma > 0, gamma, 1.0)
        r_ = gamma[:, None] * q
        for j in range(m):                       # oldest to newest
            use = (m - 1 - j) < np.minimum(k, m)
            slot = (k - 1 - (m - 1 - j)) % m
            s = S[idx, slot] * ~held
            y = Y[idx, slot] * ~held
            rr = rho[idx, slot]
            b = np.where(use, rr * np.einsum("ij,ij->i", y, r_), 0.0)
            r_ = r_ + (alpha[:, j] - b)[:, None] * s
        d = -r_ * ~held
        slope = np.einsum("ij,ij->i", d, pg)
        bad = slope >= 0
        d[bad] = -pg[bad]
        hist[idx[bad]] = 0
        # projected Armijo backtracking, all instances of the round together
        step = np.ones(idx.size)
        x_new = np.clip(xi + d, lo, hi)
        todo = ~done
        f_new = f[idx].copy()
        g_new = gi.copy()
        for _bt in range(max_backtrack):
            cur = np.nonzero(todo)[0]
            if cur.size == 0:
                break
            fc, gc = fun(x_new[cur], idx[cur])
            dec = np.einsum("ij,ij->i", gi[cur], x_new[cur] - xi[cur])
            ok = fc <= f[idx[cur]] + c1 * dec
            f_new[cur[ok]] = fc[ok]
            g_new[cur[ok]] = gc[ok]
            todo[cur[ok]] = False
            nok = cur[~ok]
            step[nok] *= 0.5
            x_new[nok] = np.clip(xi[nok] + step[nok, None] * d[nok], lo, hi)
        failed = todo & ~done
        # accept
        acc = ~done & ~failed
        sv = x_new - xi
        yv = g_new - gi
        syv = np.einsum("ij,ij->i", sv, yv)
        upd = acc & (syv > 1e-16 * np.linalg.norm(sv, axis=1) * np.linalg.norm(yv, axis=1))
        ii = idx[upd]
        slot = hist[ii] % m
        S[ii, slot] = sv[upd]
        Y[ii, slot] = yv[upd]
        rho[ii, slot] = 1.0 / syv[upd]
        hist[ii] += 1
        ia = idx[acc]
        df = f[ia] - f_new[acc]
        dx = np.abs(sv[acc]).max(axis=1)
        x[ia] = x_new[acc]
        f[ia] = f_new[acc]
        g[ia] = g_new[acc]
        iters[ia] += 1
        stop = np.zeros(idx.size, bool)
        stop[done | failed] = True
        stop[np.nonzero(acc)[0][(dx <= xtol) | (df <= ftol * (1.0 + np.abs(f[ia])))]] = True
        active[idx[stop]] = False
    return x, f, iters


# ---------------------------------------------------------------------------------------
# the reference's multistart points and the single-graph fit
# ---------------------------------------------------------------------------------------
_M64 = (1 << 64) - 1


def init_theta(inst: int, dim: int) -> float:
    """The reference's splitmix start point for restart ``inst``, coordinate ``dim``
    (src/core/qpadm/qpgraph_opt_constants.hpp:18-26, cpu_backend.cpp:2966-2973)."""
    z = (inst * 0x100000001B3 + dim * 0x9E3779B97F4A7C15 + 0xD1B54A32D192ED03) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    z = z ^ (z >> 31)
    return (z & 0xFFFFFFFFFFFFF) / 4503599627370496.0


def fit_sets(gset: GraphSet, numstart=10, fudge=1e-4, constrained=True, maxit=200, tol=1e-9):
    """Fits every graph of the set from ``numstart`` restarts each, all restarts of all
    graphs in one L-BFGS batch.  Returns per graph: best score, restart spread, best
    theta, per-coordinate min / max over restarts, and the L-BFGS iterations used."""
    D = [t.nadmix for t in gset.topologies]
    inst_graph, x0 = [], []
    for g, d in enumerate(D):
        if gset.status[g] != 0:
            continue
        for s in range(numstart if d > 0 else 1):
            inst_graph.append(g)
            x0.append([init_theta(s, j) for j in range(d)] + [0.0] * (gset.dmax - d))
    inst_graph = np.asarray(inst_graph, np.int32)
    x0 = np.asarray(x0, np.float64).reshape(len(inst_graph), gset.dmax)
    mask = np.zeros_like(x0)
    for i, g in enumerate(inst_graph):
        mask[i, :D[g]] = 1.0

    def fun(x, idx):
        r = gset.eval(inst_graph[idx], x, mode=1, fudge=fudge, constrained=constrained)
        return r["score"], r["grad_theta"] * mask[idx]

    x, f, it = lbfgs_box(fun, x0, max_iter=maxit, gtol=tol * 1e-2)
    out = []
    for g in range(gset.n):
        sel = np.nonzero(inst_graph == g)[0]
        if sel.size == 0:
            out.append(None)
            continue
        fs = f[sel]
        b = sel[int(np.argmin(fs))]
        d = D[g]
        out.append({"score": float(f[b]), "restart_spread": float(fs.max() - fs.min()),
                    "theta": x[b, :d].copy(), "theta_lo": x[sel, :d].min(axis=0), "theta_hi": x[sel, :d].max(axis=0),
                    "iterations": int(it[sel].max())})
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
    r = gset.eval([0], fit["theta"][None, :] if topo.nadmix else None, mode=0, fudge=fudge,
                  constrained=constrained, want_fit=True)
    score = float(r["score"][0])
    if not np.isfinite(score):
        return dict(base, status="non_spd_covariance", score=score, restart_spread=0.0, weight=[], weight_lo=[],
                    weight_hi=[], edge_length=[], worst_residual_z=0.0, worst_pop2="", worst_pop3="")
    f_obs, qdiag, _ = gset.basis(0)
    fitted = r["fit"][0, :topo.npair]
    with np.errstate(invalid="ignore", divide="ignore"):
        z = (f_obs - fitted) / np.where(qdiag > 0, np.sqrt(qdiag), np.nan)
    worst, wk = 0.0, -1
    for k, zk in enumerate(z):
        if np.isfinite(zk) and abs(zk) > abs(worst):
            worst, wk = float(zk), k
    p2, p3 = topo.centred_pair(wk) if wk >= 0 else ("", "")
    return dict(base, status="ok", score=score, restart_spread=fit["restart_spread"],
                weight=fit["theta"].tolist(), weight_lo=fit["theta_lo"].tolist(), weight_hi=fit["theta_hi"].tolist(),
                edge_length=r["bl"][0, :topo.nedge].tolist(), worst_residual_z=worst, worst_pop2=p2, worst_pop3=p3)
