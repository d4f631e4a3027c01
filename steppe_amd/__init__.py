"""steppe_amd — an MI355X-native engine for Steppe's f-statistics hot path.

Packed 2-bit genotypes -> per-population allele counts -> per-jackknife-block f2
tensor resident in HBM -> fits, on hand-written HIP kernels for gfx950 behind a
C ABI (include/steppe_hip.h).  The Python surface follows the reference's
``steppe`` package (bindings/steppe/__init__.py:444-718): ``read_f2``,
``extract_f2`` and the fit functions return the same shapes and names.
"""
from ._core import (  # noqa: F401
    F2Handle,
    PRECISION_NAMES,
    STP_PREC_FP64,
    STP_PREC_SPLIT_BF16,
    qpadm_models,
    read_f2,
    read_f2_host,
    run_extract_f2,
    run_f3,
    run_f4,
    run_f4ratio,
    run_fstat_sweep,
    run_qpadm,
    run_qpadm_search,
    run_qpdstat,
    run_qpwave,
    upload_f2,
    write_f2_dir,
)

import enum

__version__ = "0.1.0"


def extract_f2(prefix, *args, pops=None, out=None, device=0, blgsize=0.05, maf=0.0, maxmiss=0.0,
               autosomes_only=True, drop_monomorphic=False, transversions_only=False, ploidy="auto",
               strand_mode="drop", precision=None, **kw):
    """Facade over ``_core.run_extract_f2`` (bindings/steppe/__init__.py:493-559):
    ``extract_f2(prefix, pops=[...], out=None, ...)`` reads ``<prefix>.{geno,snp,ind}`` (or
    the PLINK ``.bed/.bim/.fam``) and returns a resident F2Handle, or the directory path
    when ``out`` is given.  Called with an in-memory packed tile instead of a prefix
    (``extract_f2(packed, n_snp, pop_offsets, block_id, n_block, pops=..., ...)``) it runs
    the tile form."""
    import os
    if isinstance(prefix, (str, os.PathLike)):
        if args:
            raise TypeError("extract_f2(prefix, ...): pass the options as keywords")
        return run_extract_f2(str(prefix), pops, out or "", device, blgsize, maf, maxmiss, autosomes_only,
                              drop_monomorphic, transversions_only, ploidy, strand_mode, precision)
    return run_extract_f2(prefix, *args, pops=pops, out=out or "", device=device, ploidy=ploidy,
                          precision=precision, **kw)


class Status(enum.Enum):
    """Result status values of the fits (bindings/steppe/__init__.py:146-160)."""
    OK = "ok"
    DEVICE_OOM = "device_oom"
    RANK_DEFICIENT = "rank_deficient"
    NON_SPD_COVARIANCE = "non_spd_covariance"
    CHISQ_UNDEFINED = "chisq_undefined"
    INVALID_CONFIG = "invalid_config"


def _frame(cols: dict):
    import pandas as pd
    return pd.DataFrame({k: list(v) for k, v in cols.items()})


class QpAdmResult:
    """One qpAdm fit with the reference facade's attributes (bindings/steppe/__init__.py:
    193-270): weight / se / z lists, p, chisq, dof, est_rank, f4rank, status, precision,
    model_index, feasible, and the weights / rankdrop / popdrop DataFrames."""

    def __init__(self, d: dict, *, target, left):
        self._d = d
        self.target, self.left = target, list(left)
        self.weight, self.se, self.z = list(d["weight"]), list(d["se"]), list(d["z"])
        self.p, self.chisq, self.dof = d["p"], d["chisq"], d["dof"]
        self.est_rank, self.f4rank = d["est_rank"], d["f4rank"]
        self.status = Status(d["status"])
        self.precision, self.model_index = d["precision"], d["model_index"]
        pf = d.get("popdrop_feasible") or []
        w = self.weight
        self.feasible = bool(pf[0]) if pf else (bool(w) and all(0.0 <= x <= 1.0 for x in w))

    @property
    def weights(self):
        n = len(self.weight)
        nan = float("nan")
        return _frame({"target": [self.target] * n, "left": self.left, "weight": self.weight,
                       "se": self.se if len(self.se) == n else [nan] * n,
                       "z": self.z if len(self.z) == n else [nan] * n})

    @property
    def rankdrop(self):
        d = self._d
        return _frame({c: d["rankdrop_" + c] for c in ("f4rank", "dof", "chisq", "p", "dofdiff", "chisqdiff",
                                                       "p_nested")})

    @property
    def popdrop(self):
        d = self._d
        return _frame({c: d["popdrop_" + c] for c in ("pat", "wt", "dof", "f4rank", "chisq", "p", "feasible")})

    def __repr__(self) -> str:
        return (f"QpAdmResult(target={self.target!r}, left={self.left!r}, weight={self.weight!r}, p={self.p!r}, "
                f"f4rank={self.f4rank}, feasible={self.feasible}, status={self.status.name})")


def qpadm(f2, target=None, left=None, right=None, *, rank=-1, fudge=1e-4, als_iterations=20, rank_alpha=0.05,
          allow_negative_weights=True):
    """One qpAdm model (bindings/steppe/__init__.py:604-640): a QpAdmResult."""
    d = run_qpadm(f2, target, left, right, rank=rank, fudge=fudge, als_iterations=als_iterations,
                  rank_alpha=rank_alpha, allow_negative_weights=allow_negative_weights)
    return QpAdmResult(d, target=target, left=left)


def qpadm_search(f2, target=None, models=None, right=None, *, rank=-1, fudge=1e-4, als_iterations=20,
                 rank_alpha=0.05, allow_negative_weights=True, jackknife="all", as_dataframe=False, **kw):
    """Batched qpAdm over candidate source sets (bindings/steppe/__init__.py:1103-1160):
    ``models`` lists source sets (a list of labels, or {"left": [...]}); a QpAdmResult
    per model in input order, or one DataFrame with ``as_dataframe``."""
    if models is None:
        models = kw.pop("lefts", None)
    lefts = [list(m["left"]) if isinstance(m, dict) else list(m) for m in models]
    ds = run_qpadm_search(f2, target, lefts, right, rank=rank, fudge=fudge, als_iterations=als_iterations,
                          rank_alpha=rank_alpha, allow_negative_weights=allow_negative_weights,
                          jackknife=jackknife, **kw)
    res = [QpAdmResult(d, target=target, left=left) for d, left in zip(ds, lefts)]
    if not as_dataframe:
        return res
    return _frame({"model": list(range(len(res))), "left": [",".join(r.left) for r in res],
                   "p": [r.p for r in res], "chisq": [r.chisq for r in res], "f4rank": [r.f4rank for r in res],
                   "feasible": [r.feasible for r in res], "status": [r.status.value for r in res],
                   "weight": [r.weight for r in res], "se": [r.se for r in res]})


class QpGraphResult:
    """One qpGraph fit with the reference facade's attributes (bindings/steppe/__init__.py:
    646-693): score, restart_spread, status, worst_residual_z, and the weights / edges
    DataFrames."""

    def __init__(self, d: dict):
        self._d = d

    score = property(lambda self: float(self._d["score"]))
    restart_spread = property(lambda self: float(self._d["restart_spread"]))
    status = property(lambda self: Status(self._d["status"]))
    worst_residual_z = property(lambda self: float(self._d["worst_residual_z"]))

    @property
    def weights(self):
        d = self._d
        return _frame({"from": d["admix_from"], "to": d["admix_to"], "weight": d["weight"], "low": d["weight_lo"],
                       "high": d["weight_hi"]})

    @property
    def edges(self):
        d = self._d
        return _frame({"from": d["edge_from"], "to": d["edge_to"], "length": d["edge_length"]})

    def __repr__(self) -> str:
        return (f"QpGraphResult(score={self.score:.6f}, nadmix={len(self._d['weight'])}, "
                f"nedge={len(self._d['edge_length'])}, status={self.status.name})")


def _qpgraph_fit(f2, edges, *, numstart=10, fudge=1e-4, diag_f3=1e-5, constrained=True):
    """One admixture-graph fit on the device (bindings/steppe/__init__.py:696-718): a
    QpGraphResult.  Reached as ``steppe_amd.qpgraph(f2, edges, ...)``: the qpgraph
    submodule is callable, so the reference's ``steppe.qpgraph(...)`` call and
    ``from steppe_amd import qpgraph`` (the module) both work."""
    from .qpgraph import run_qpgraph
    return QpGraphResult(run_qpgraph(f2, edges, numstart=numstart, fudge=fudge, diag_f3=diag_f3,
                                     constrained=constrained))


from . import qpgraph  # noqa: E402  (the callable submodule, after the facade it calls)


class FstatTable:
    """A table of per-row statistics with list-valued columns (pop1.. then the values),
    the shape of the reference's F4Result / F3Result / F4RatioResult
    (bindings/steppe/__init__.py:321-441): attribute access per column, len(), and
    ``.table`` as a pandas DataFrame."""

    def __init__(self, d: dict, columns):
        self._d = d
        self.columns = list(columns)
        for c in self.columns:
            setattr(self, c, list(d[c]))
        self.status = d["status"]
        self.precision = d["precision"]

    @property
    def table(self):
        import pandas as pd
        return pd.DataFrame({c: list(getattr(self, c)) for c in self.columns})

    def __len__(self) -> int:
        return len(getattr(self, self.columns[-1]))

    def __repr__(self) -> str:
        return f"FstatTable(rows={len(self)}, status={self.status})"


_F4_COLS = ("pop1", "pop2", "pop3", "pop4", "est", "se", "z", "p")
_F3_COLS = ("pop1", "pop2", "pop3", "est", "se", "z", "p")
_F4R_COLS = ("pop1", "pop2", "pop3", "pop4", "pop5", "alpha", "se", "z")


def f4(f2, quartets, *, as_dataframe=False):
    """f4(p1,p2;p3,p4) per (p1, p2, p3, p4) row: est / se / z / p (bindings/steppe/__init__.py:865)."""
    res = FstatTable(run_f4(f2, quartets), _F4_COLS)
    return res.table if as_dataframe else res


def qpdstat(f2, quartets, *, as_dataframe=False):
    """qpDstat over f2 data: the f4 statistic per quadruple (bindings/steppe/__init__.py:891)."""
    res = FstatTable(run_qpdstat(f2, quartets), _F4_COLS)
    return res.table if as_dataframe else res


def f3(f2, triples, *, as_dataframe=False):
    """f3(C;A,B) per (C, A, B) row: est / se / z / p (bindings/steppe/__init__.py:1045)."""
    res = FstatTable(run_f3(f2, triples), _F3_COLS)
    return res.table if as_dataframe else res


def f4ratio(f2, tuples, *, as_dataframe=False):
    """alpha = f4(p1,p2;p3,p4) / f4(p1,p2;p5,p4) per 5-tuple (bindings/steppe/__init__.py:1073)."""
    res = FstatTable(run_f4ratio(f2, tuples), _F4R_COLS)
    return res.table if as_dataframe else res


class QpWaveResult:
    """A qpWave rank sweep with the reference facade's attributes
    (bindings/steppe/__init__.py:273-318): f4rank, est_rank, status, precision, and the
    rankdrop / per_rank DataFrames."""

    def __init__(self, d: dict, *, left):
        self._d = d
        self.left = list(left)
        self.f4rank, self.est_rank = d["f4rank"], d["est_rank"]
        self.status = Status(d["status"])
        self.precision = d["precision"]

    @property
    def rankdrop(self):
        d = self._d
        return _frame({c: d["rankdrop_" + c] for c in ("f4rank", "dof", "chisq", "p", "dofdiff", "chisqdiff",
                                                       "p_nested")})

    @property
    def per_rank(self):
        d = self._d
        return _frame({"rank": range(len(d["rank_chisq"])), "chisq": d["rank_chisq"], "dof": d["rank_dof"],
                       "p": d["rank_p"]})

    def __repr__(self) -> str:
        return (f"QpWaveResult(left={self.left!r}, f4rank={self.f4rank}, est_rank={self.est_rank}, "
                f"status={self.status.name})")


def qpwave(f2, *, left, right, fudge=1e-4, rank_alpha=0.05):
    """qpWave rank sweep, left[0] the reference row (bindings/steppe/__init__.py:632-643):
    a QpWaveResult."""
    return QpWaveResult(run_qpwave(f2, left, right, fudge, rank_alpha), left=left)


def f4_sweep(f2, pops=None, *, min_z=3.0, top_k=0, sure=False, all_combinations=False, method="auto"):
    """Every f4 over the 4-subsets of pops (all populations by default), filtered by |z|
    (src/core/qpadm/fstat_sweep.cpp run_f4_sweep; the CLI's f4-sweep)."""
    return run_fstat_sweep(f2, 4, pops, min_z=min_z, top_k=top_k, sure=sure, all_combinations=all_combinations,
                           method=method)


def f3_sweep(f2, pops=None, *, min_z=3.0, top_k=0, sure=False, all_combinations=False):
    """Every f3 over the 3-subsets of pops, filtered by |z| (run_f3_sweep)."""
    return run_fstat_sweep(f2, 3, pops, min_z=min_z, top_k=top_k, sure=sure, all_combinations=all_combinations)


class QpGraphSearchResult:
    """Topology-search result with the reference facade's attributes
    (bindings/steppe/__init__.py:721-815): n_trees, n_admix1, n_candidates, best_score,
    second_best_score, best_nadmix, best_hash, best_edges, heuristic_recovered,
    fit_all_wall_ms, topologies_per_s, argmin and candidates (a DataFrame of nadmix,
    hash, score, restart_spread in enumeration order)."""

    def __init__(self, d: dict):
        self._d = d

    def __getattr__(self, name):
        d = self.__dict__.get("_d", {})
        if name in ("n_trees", "n_admix1", "n_candidates", "best_nadmix", "best_hash"):
            return int(d[name])
        if name in ("best_score", "second_best_score", "fit_all_wall_ms", "topologies_per_s"):
            return float(d[name])
        if name == "heuristic_recovered":
            return bool(d[name])
        if name == "best_edges":
            return [(str(p), str(c)) for p, c in d["best_edges"]]
        raise AttributeError(name)

    @property
    def argmin(self) -> int:
        scores = self._d.get("cand_score") or []
        if not scores:
            return -1
        best, idx = scores[0], 0
        for i, s in enumerate(scores):
            if s < best:
                best, idx = s, i
        return idx

    @property
    def candidates(self):
        import pandas as pd
        return pd.DataFrame({"nadmix": list(self._d.get("cand_nadmix") or []),
                             "hash": list(self._d.get("cand_hash") or []),
                             "score": list(self._d.get("cand_score") or []),
                             "restart_spread": list(self._d.get("cand_restart_spread") or [])})

    def __repr__(self) -> str:
        return (f"QpGraphSearchResult(n_candidates={self.n_candidates}, best_score={self.best_score:.6f}, "
                f"best_nadmix={self.best_nadmix}, heuristic_recovered={self.heuristic_recovered})")


def qpgraph_search(f2, *, pops, max_nadmix=1, numstart=10, fudge=1e-4, diag_f3=1e-5, constrained=True,
                   run_heuristic=True):
    """qpGraph topology search (bindings/steppe/__init__.py:820-857): every rooted topology
    on ``pops`` with at most ``max_nadmix`` (<= 1) admixture events, fitted on the device."""
    from .qpgraph import run_qpgraph_search
    return QpGraphSearchResult(run_qpgraph_search(f2, pops, max_nadmix, numstart, fudge, diag_f3, constrained,
                                                  run_heuristic))
