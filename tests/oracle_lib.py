"""ctypes front end to the CPU parity checkers under oracle/ (test-side only).

``cref`` is the restated oracle (oracle/libcpu_ref.so); ``ref`` is the real
reference CpuBackend compiled from its own sources (oracle/_ref/libsteppe_ref.so),
present wherever oracle/Makefile's ``ref`` target could run.  Both expose the same
argument conventions, so tests call them through the same Python wrappers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CREF_SO = os.path.join(ORACLE, "libcpu_ref.so")
REF_SO = os.path.join(ORACLE, "_ref", "libsteppe_ref.so")

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")


def _ptr_or_null(a, dtype):
    if a is None:
        return None
    return np.ascontiguousarray(a, dtype=dtype).ctypes.data_as(C.c_void_p)


def ensure_built():
    if not os.path.exists(CREF_SO):
        subprocess.run(["make", "-s", "-C", ORACLE, "all"], check=True)
    if not os.path.exists(REF_SO) and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", ORACLE, "ref"], check=True)


class _Lib:
    def __init__(self, path: str, prefix: str):
        self.lib = C.CDLL(path)
        self.prefix = prefix
        self.path = path

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    # ---- decode -------------------------------------------------------------
    def decode_af(self, panel, ploidy_mode="sample"):
        P, M = panel.n_pop, panel.n_snp
        q = np.zeros(P * M, np.float64)
        v = np.zeros_like(q)
        n = np.zeros_like(q)
        sp, pl, det = self._ploidy_args(panel, ploidy_mode)
        f = self.fn("decode_af")
        f.restype = C.c_int
        rc = f(panel.packed.ctypes.data_as(C.c_void_p), C.c_size_t(panel.bytes_per_record),
               C.c_size_t(M), C.c_size_t(panel.n_ind),
               panel.pop_offsets.ctypes.data_as(C.c_void_p), C.c_int(P), sp, C.c_int(pl),
               C.c_int(det), q.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
               n.ctypes.data_as(C.c_void_p))
        assert rc == 0
        return (q.reshape(M, P).T, v.reshape(M, P).T, n.reshape(M, P).T)

    def detect_ploidy(self, panel):
        out = np.zeros(panel.n_ind, np.int32)
        f = self.fn("detect_ploidy")
        rc = f(panel.packed.ctypes.data_as(C.c_void_p), C.c_size_t(panel.bytes_per_record),
               C.c_size_t(panel.n_snp), C.c_size_t(panel.n_ind), out.ctypes.data_as(C.c_void_p))
        assert rc == 0
        return out

    @staticmethod
    def _ploidy_args(panel, mode):
        if mode == "sample":
            sp = np.ascontiguousarray(panel.sample_ploidy, np.int32)
            panel._sp_keep = sp
            return sp.ctypes.data_as(C.c_void_p), 2, 0
        if mode == "detect":
            return None, 2, 1
        return None, int(mode), 0

    # ---- blocks -------------------------------------------------------------
    def assign_blocks(self, chrom, genpos, blgsize, physpos=None, bp_window=2.0e6):
        chrom = np.ascontiguousarray(chrom, np.int32)
        genpos = np.ascontiguousarray(genpos, np.float64)
        M = len(chrom)
        out = np.zeros(M, np.int32)
        f = self.fn("assign_blocks")
        f.restype = C.c_int
        ph = None if physpos is None else np.ascontiguousarray(physpos, np.float64)
        nb = f(chrom.ctypes.data_as(C.c_void_p), genpos.ctypes.data_as(C.c_void_p),
               C.c_int64(M), C.c_double(blgsize),
               None if ph is None else ph.ctypes.data_as(C.c_void_p), C.c_double(bp_window),
               out.ctypes.data_as(C.c_void_p))
        return out, nb

    # ---- f2 -----------------------------------------------------------------
    def f2_blocks(self, Q, V, N, block_id, nb, n_threads=1):
        """Q/V/N as [P, M] arrays; returns f2, vpair as (P, P, nb) F-order views and
        block_sizes."""
        P, M = Q.shape
        q = np.ascontiguousarray(Q.T, np.float64)   # memory i + P*s
        v = np.ascontiguousarray(V.T, np.float64)
        n = np.ascontiguousarray(N.T, np.float64)
        bid = np.ascontiguousarray(block_id, np.int32)
        f2 = np.zeros(P * P * nb, np.float64)
        vp = np.zeros_like(f2)
        bs = np.zeros(nb, np.int32)
        args = [q.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
                n.ctypes.data_as(C.c_void_p), C.c_int(P), C.c_int64(M),
                bid.ctypes.data_as(C.c_void_p), C.c_int(nb), f2.ctypes.data_as(C.c_void_p),
                vp.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p)]
        f = self.fn("f2_blocks")
        f.restype = C.c_int
        if self.prefix == "cref_":
            args.append(C.c_int(n_threads))
        rc = f(*args)
        assert rc == 0
        return (f2.reshape(nb, P, P).transpose(2, 1, 0), vp.reshape(nb, P, P).transpose(2, 1, 0),
                bs)


    # ---- qpAdm ----------------------------------------------------------------
    def qpadm(self, f2, vpair, block_sizes, target, left, right, fudge=1e-4, als_iterations=20,
              rank=-1, rank_alpha=0.05, jackknife=2):
        """One model on a (P, P, nb) tensor; returns the reference's result dict."""
        from steppe_amd._core import QpAdmResult
        f2 = np.asarray(f2, np.float64)
        P, _, nb = f2.shape
        fl = np.asfortranarray(f2).ravel(order="F")
        vp = None if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
        bs = np.ascontiguousarray(block_sizes, np.int32)
        le = np.ascontiguousarray(left, np.int32)
        ri = np.ascontiguousarray(right, np.int32)
        out = QpAdmResult()
        f = self.fn("qpadm")
        f.restype = C.c_int
        rc = f(fl.ctypes.data_as(C.c_void_p), None if vp is None else vp.ctypes.data_as(C.c_void_p),
               bs.ctypes.data_as(C.c_void_p), C.c_int(P), C.c_int(nb), C.c_int(int(target)),
               le.ctypes.data_as(C.c_void_p), C.c_int(len(le)), ri.ctypes.data_as(C.c_void_p),
               C.c_int(len(ri)), C.c_double(fudge), C.c_int(als_iterations), C.c_int(rank),
               C.c_double(rank_alpha), C.c_int(jackknife), C.byref(out))
        assert rc == 0
        return out.to_dict()


    # ---- genotype files -> f2 ------------------------------------------------------
    def extract(self, geno, snp, ind, pops=None, top_k=0, min_n=0, maf=0.0, maxmiss=0.0, autosomes_only=True,
                drop_monomorphic=False, transversions_only=False, strand_mode=0, blgsize=0.05, ploidy=0):
        """The extract over a genotype triple; returns a dict (f2, vpair as (P, P, nb),
        block_sizes, labels, info) or raises ValueError with the library's message."""
        mode = 0 if pops else (1 if top_k > 0 else 2)
        labels = [str(x).encode() for x in (pops or [])]
        arr = (C.c_char_p * max(len(labels), 1))(*labels)
        err = C.create_string_buffer(512)
        f = self.fn("extract_open")
        f.restype = C.c_void_p
        f.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, C.c_void_p, C.c_int, C.c_longlong, C.c_longlong,
                      C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_char_p,
                      C.c_int]
        h = f(str(geno).encode(), str(snp).encode(), str(ind).encode(), mode, C.cast(arr, C.c_void_p), len(labels),
              top_k, min_n, maf, maxmiss, int(autosomes_only), int(drop_monomorphic), int(transversions_only),
              strand_mode, blgsize, ploidy, err, 512)
        if not h:
            raise ValueError(err.value.decode())
        try:
            P, nb = C.c_int(), C.c_int()
            info = (C.c_longlong * 4)()
            d = self.fn("extract_dims")
            d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            d(h, C.byref(P), C.byref(nb), info)
            P, nb = P.value, nb.value
            f2 = np.zeros(P * P * nb)
            vp = np.zeros_like(f2)
            bs = np.zeros(nb, np.int32)
            g = self.fn("extract_fetch")
            g.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
            g(h, f2.ctypes.data_as(C.c_void_p), vp.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p))
            lab = self.fn("extract_label")
            lab.restype = C.c_char_p
            lab.argtypes = [C.c_void_p, C.c_int]
            labels = [lab(h, i).decode() for i in range(P)]
        finally:
            fr = self.fn("extract_free")
            fr.argtypes = [C.c_void_p]
            fr(h)
        return {"f2": f2.reshape((P, P, nb), order="F"), "vpair": vp.reshape((P, P, nb), order="F"),
                "block_sizes": bs, "labels": labels,
                "info": {"n_snp_total": info[0], "n_snp_kept": info[1], "n_pseudo_haploid": info[2],
                         "n_diploid": info[3]}}

    # ---- f-statistics -------------------------------------------------------------
    @staticmethod
    def _tensor_args(f2, vpair, block_sizes):
        f2 = np.asarray(f2, np.float64)
        P, _, nb = f2.shape
        fl = np.asfortranarray(f2).ravel(order="F")
        vp = None if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
        bs = np.ascontiguousarray(block_sizes, np.int32)
        return (fl, vp, bs), (fl.ctypes.data_as(C.c_void_p), None if vp is None else vp.ctypes.data_as(C.c_void_p),
                              bs.ctypes.data_as(C.c_void_p), C.c_int(P), C.c_int(nb))

    def _fstat(self, name, f2, vpair, block_sizes, tuples, width, fields):
        keep, targs = self._tensor_args(f2, vpair, block_sizes)
        t = np.ascontiguousarray(np.asarray(tuples, np.int32).reshape(-1, width))
        N = t.shape[0]
        outs = [np.full(max(N, 1), -7.0) for _ in fields]
        f = self.fn(name)
        f.restype = C.c_int
        rc = f(*targs, t.ctypes.data_as(C.c_void_p), C.c_int(N), *[o.ctypes.data_as(C.c_void_p) for o in outs])
        assert rc == 0, rc
        return {k: o[:N] for k, o in zip(fields, outs)}

    def f4(self, f2, vpair, block_sizes, quartets):
        """run_f4 over (N, 4) quartets -> dict of est / se / z / p arrays."""
        return self._fstat("f4", f2, vpair, block_sizes, quartets, 4, ("est", "se", "z", "p"))

    def f3(self, f2, vpair, block_sizes, triples):
        """run_f3 over (N, 3) triples (C, A, B) -> dict of est / se / z / p arrays."""
        return self._fstat("f3", f2, vpair, block_sizes, triples, 3, ("est", "se", "z", "p"))

    def f4ratio(self, f2, vpair, block_sizes, tuples):
        """run_f4ratio over (N, 5) tuples -> dict of alpha / se / z arrays."""
        return self._fstat("f4ratio", f2, vpair, block_sizes, tuples, 5, ("alpha", "se", "z"))

    def fstat_sweep(self, f2, vpair, block_sizes, k, subset=None, filter="min_z", min_z=3.0, top_k=100,
                    sure=False):
        """All-combination f4 (k=4) / f3 (k=3) sweep (restated oracle only)."""
        keep, targs = self._tensor_args(f2, vpair, block_sizes)
        sub = None if subset is None else np.ascontiguousarray(subset, np.int32)
        P = np.asarray(f2).shape[0]
        rng = P if sub is None else len(sub)
        from math import comb
        mode = 1 if filter == "top_k" else 0
        K = (top_k if (mode == 1 and top_k > 0) else 1000000)
        cap = max(1, min(K, comb(rng, k)))
        keys = np.zeros((cap, 4), np.int32)
        est, se, z, p = (np.zeros(cap) for _ in range(4))
        n_out = C.c_int64(0)
        enum = C.c_uint64(0)
        capped = C.c_int(0)
        f = self.fn("fstat_sweep")
        f.restype = C.c_int
        st = f(*targs, C.c_int(k), None if sub is None else sub.ctypes.data_as(C.c_void_p),
               C.c_int(0 if sub is None else len(sub)), C.c_int(mode), C.c_double(min_z), C.c_int64(top_k),
               C.c_int(1 if sure else 0), C.c_int64(cap), keys.ctypes.data_as(C.c_void_p),
               est.ctypes.data_as(C.c_void_p), se.ctypes.data_as(C.c_void_p), z.ctypes.data_as(C.c_void_p),
               p.ctypes.data_as(C.c_void_p), C.byref(n_out), C.byref(enum), C.byref(capped))
        n = min(n_out.value, cap)
        return {"status": st, "keys": keys[:n], "est": est[:n], "se": se[:n], "z": z[:n], "p": p[:n],
                "enumerated": enum.value, "capped": bool(capped.value), "survivors": n_out.value}

    def unrank(self, r, k):
        c = np.zeros(4, np.int32)
        f = self.fn("unrank")
        f.restype = None
        f(C.c_int64(r), C.c_int(k), c.ctypes.data_as(C.c_void_p))
        return tuple(int(x) for x in c[:k])

    # ---- qpGraph ----------------------------------------------------------------
    @staticmethod
    def _strs(items):
        arr = (C.c_char_p * max(len(items), 1))(*[x.encode() for x in items])
        return arr

    def _graph_args(self, edges, names):
        fr = self._strs([a for a, _ in edges])
        to = self._strs([b for _, b in edges])
        nm = self._strs(list(names))
        self._keep = (fr, to, nm)
        return [C.c_int(len(edges)), fr, to, C.c_int(len(names)), nm]

    def qpgraph_fit(self, f2, vpair, block_sizes, edges, names, fudge=1e-4, diag_f3=1e-5, numstart=10,
                    constrained=True, maxit=200, tol=1e-9):
        """run_qpgraph on a host tensor: dict with score, restart_spread, weight, weight_lo,
        weight_hi, edge_length, worst_residual_z, status."""
        f2 = np.asarray(f2, np.float64)
        P, _, nb = f2.shape
        fl = np.asfortranarray(f2).ravel(order="F")
        vp = None if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
        bs = np.ascontiguousarray(block_sizes, np.int32)
        scal = np.zeros(5)
        th, lo, hi = np.zeros(64), np.zeros(64), np.zeros(64)
        bl, fit = np.zeros(256), np.zeros(4096)
        f = self.fn("qpgraph_fit")
        f.restype = C.c_int
        rc = f(fl.ctypes.data_as(C.c_void_p), None if vp is None else vp.ctypes.data_as(C.c_void_p),
               bs.ctypes.data_as(C.c_void_p), C.c_int(P), C.c_int(nb), *self._graph_args(edges, names),
               C.c_double(fudge), C.c_double(diag_f3), C.c_int(numstart), C.c_int(1 if constrained else 0),
               C.c_int(maxit), C.c_double(tol), scal.ctypes.data_as(C.c_void_p), th.ctypes.data_as(C.c_void_p),
               lo.ctypes.data_as(C.c_void_p), hi.ctypes.data_as(C.c_void_p), bl.ctypes.data_as(C.c_void_p),
               fit.ctypes.data_as(C.c_void_p))
        assert rc == 0
        dims = cref().qpgraph_model(edges, names)
        D, ne, npair = dims["nadmix"], dims["nedge_norm"], dims["npair"]
        return {"score": scal[0], "restart_spread": scal[1], "worst_residual_z": scal[2],
                "status": int(scal[4]), "weight": th[:D].copy(), "weight_lo": lo[:D].copy(),
                "weight_hi": hi[:D].copy(), "edge_length": bl[:ne].copy(), "f3_fit": fit[:npair].copy()}

    def qpgraph_score(self, edges, names, theta, f_obs, qinv, fudge=1e-4, constrained=True):
        """(score, edge lengths, fitted f3) at mixture weights theta."""
        dims = cref().qpgraph_model(edges, names)
        bl = np.zeros(max(dims["nedge_norm"], 1))
        fit = np.zeros(max(dims["npair"], 1))
        th = np.ascontiguousarray(theta if theta is not None and len(theta) else np.zeros(1), np.float64)
        fo = np.ascontiguousarray(f_obs, np.float64)
        qi = np.asfortranarray(np.asarray(qinv, np.float64)).ravel(order="F")
        f = self.fn("qpgraph_score")
        f.restype = C.c_double
        s = f(*self._graph_args(edges, names), th.ctypes.data_as(C.c_void_p), fo.ctypes.data_as(C.c_void_p),
              qi.ctypes.data_as(C.c_void_p), C.c_double(fudge), C.c_int(1 if constrained else 0),
              bl.ctypes.data_as(C.c_void_p), fit.ctypes.data_as(C.c_void_p))
        return s, bl[:dims["nedge_norm"]].copy(), fit[:dims["npair"]].copy()

    def qpgraph_model(self, edges, names):
        d = np.zeros(8, np.int32)
        f = self.fn("qpgraph_model")
        f.restype = C.c_int
        rc = f(*self._graph_args(edges, names), d.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise ValueError("invalid graph")
        keys = ("npop", "nedge_norm", "nadmix", "npair", "npath", "base_leaf", "n_pe", "n_pae")
        return {k: int(v) for k, v in zip(keys, d)}

    def qpgraph_tables(self, edges, names):
        d = self.qpgraph_model(edges, names)
        ne, npop, npair = d["nedge_norm"], d["npop"], d["npair"]
        t = {"pwts0": np.zeros(ne * npop), "pe_edge": np.zeros(d["n_pe"], np.int32),
             "pe_leaf": np.zeros(d["n_pe"], np.int32), "pe_path": np.zeros(d["n_pe"], np.int32),
             "pae_path": np.zeros(d["n_pae"], np.int32), "pae_id": np.zeros(d["n_pae"], np.int32),
             "cmb1": np.zeros(npair, np.int32), "cmb2": np.zeros(npair, np.int32),
             "leaf_to_f2": np.zeros(npop, np.int32)}
        f = self.fn("qpgraph_tables")
        f.restype = C.c_int
        args = [t[k].ctypes.data_as(C.c_void_p) if t[k].size else None for k in
                ("pwts0", "pe_edge", "pe_leaf", "pe_path", "pae_path", "pae_id", "cmb1", "cmb2", "leaf_to_f2")]
        assert f(*self._graph_args(edges, names), *args) == 0
        t["pwts0"] = t["pwts0"].reshape((ne, npop), order="F")
        t.update(d)
        return t

    def qpgraph_basis(self, f2, vpair, block_sizes, edges, names, diag_f3=1e-5):
        """(f_obs, Q, Qinv) of the f3 basis; Q/Qinv as (npair, npair) arrays."""
        d = self.qpgraph_model(edges, names)
        npair = d["npair"]
        f2 = np.asarray(f2, np.float64)
        P, _, nb = f2.shape
        fl = np.asfortranarray(f2).ravel(order="F")
        vp = None if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
        bs = np.ascontiguousarray(block_sizes, np.int32)
        fo, Q, Qi = np.zeros(npair), np.zeros(npair * npair), np.zeros(npair * npair)
        f = self.fn("qpgraph_basis")
        f.restype = C.c_int
        rc = f(fl.ctypes.data_as(C.c_void_p), None if vp is None else vp.ctypes.data_as(C.c_void_p),
               bs.ctypes.data_as(C.c_void_p), C.c_int(P), C.c_int(nb), *self._graph_args(edges, names),
               C.c_double(diag_f3), fo.ctypes.data_as(C.c_void_p), Q.ctypes.data_as(C.c_void_p),
               Qi.ctypes.data_as(C.c_void_p))
        assert rc == 0, rc
        return fo, Q.reshape((npair, npair), order="F"), Qi.reshape((npair, npair), order="F")


    # ---- qpGraph topology enumeration and search (oracle/_ref only) ------------
    def _qg_list(self, h):
        L = self.lib
        L.ref_qg_list_size.argtypes = [C.c_void_p]
        L.ref_qg_list_get.argtypes = [C.c_void_p, C.c_int] + [C.c_void_p] * 5
        L.ref_qg_list_edges.argtypes = [C.c_void_p, C.c_int]
        L.ref_qg_list_edges.restype = C.c_char_p
        L.ref_qg_list_free.argtypes = [C.c_void_p]
        out = []
        na, gid = C.c_int(), C.c_int()
        hv = C.c_ulonglong()
        sc, sp = C.c_double(), C.c_double()
        for i in range(L.ref_qg_list_size(h)):
            L.ref_qg_list_get(h, i, C.byref(na), C.byref(gid), C.byref(hv), C.byref(sc), C.byref(sp))
            es = L.ref_qg_list_edges(h, i).decode()
            edges = [tuple(t.split(">", 1)) for t in es.split(";") if t]
            out.append({"nadmix": na.value, "id": gid.value, "hash": hv.value, "edges": edges,
                        "score": sc.value, "restart_spread": sp.value})
        L.ref_qg_list_free(h)
        return out

    def qg_enumerate(self, leaves, max_nadmix=1):
        f = self.fn("qg_enumerate")
        f.restype = C.c_void_p
        return self._qg_list(f(C.c_int(len(leaves)), self._strs(leaves), C.c_int(max_nadmix)))

    def qg_neighbors(self, graph, leaves, max_nadmix=1):
        f = self.fn("qg_neighbors")
        f.restype = C.c_void_p
        f.argtypes = [C.c_char_p, C.c_int, C.c_ulonglong, C.c_int, C.c_void_p, C.c_int]
        es = ";".join(f"{a}>{b}" for a, b in graph["edges"]).encode()
        return self._qg_list(f(es, graph["nadmix"], graph["hash"], len(leaves), self._strs(leaves), max_nadmix))

    def qg_graph_hash(self, edges):
        f = self.fn("qg_graph_hash")
        f.restype = C.c_ulonglong
        f.argtypes = [C.c_char_p]
        return int(f(";".join(f"{a}>{b}" for a, b in edges).encode()))

    def qg_search(self, f2, vpair, block_sizes, names, pops, max_nadmix=1, numstart=10, fudge=1e-4,
                  diag_f3=1e-5, constrained=True, run_heuristic=False):
        """run_qpgraph_search on a host tensor: (summary dict, candidate list with scores)."""
        f2 = np.asarray(f2, np.float64)
        P, _, nb = f2.shape
        fl = np.asfortranarray(f2).ravel(order="F")
        vp = np.zeros_like(fl) if vpair is None else np.asfortranarray(np.asarray(vpair, np.float64)).ravel(order="F")
        bs = np.ascontiguousarray(block_sizes, np.int32)
        scal = np.zeros(9)
        bh = (C.c_ulonglong * 1)()
        f = self.fn("qg_search")
        f.restype = C.c_void_p
        nm, pp = self._strs(list(names)), self._strs(list(pops))
        h = f(fl.ctypes.data_as(C.c_void_p), vp.ctypes.data_as(C.c_void_p), bs.ctypes.data_as(C.c_void_p),
              C.c_int(P), C.c_int(nb), C.c_int(len(names)), nm, C.c_int(len(pops)), pp, C.c_int(max_nadmix),
              C.c_int(numstart), C.c_double(fudge), C.c_double(diag_f3), C.c_int(1 if constrained else 0),
              C.c_int(1 if run_heuristic else 0), scal.ctypes.data_as(C.c_void_p), bh)
        assert h, "reference search failed"
        summary = {"best_score": scal[0], "second_best_score": scal[1], "status": int(scal[2]),
                   "heuristic_recovered": bool(scal[3]), "best_nadmix": int(scal[4]), "n_trees": int(scal[5]),
                   "n_admix1": int(scal[6]), "n_candidates": int(scal[7]), "fit_all_wall_ms": scal[8],
                   "best_hash": int(bh[0])}
        return summary, self._qg_list(h)


_cache = {}


def cref():
    ensure_built()
    if "cref" not in _cache:
        _cache["cref"] = _Lib(CREF_SO, "cref_")
    return _cache["cref"]


def ref():
    """The real reference library, or None when it could not be built here."""
    ensure_built()
    if not os.path.exists(REF_SO):
        return None
    if "ref" not in _cache:
        _cache["ref"] = _Lib(REF_SO, "ref_")
    return _cache["ref"]
