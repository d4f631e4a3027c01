"""The package facade in the reference's shape (bindings/steppe/__init__.py): extract /
qpadm / qpadm_search / qpgraph return the same values as the _core calls, wrapped in the
reference's result objects and DataFrames."""
import os

import numpy as np
import pytest
import torch  # noqa: F401  (first import on a fresh box is slow; keep it out of test timeouts)

import steppe_amd as st
from steppe_amd import _core, qpgraph
from fixtures import read_raw_fixture

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "at2")
NAMES = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
         "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]
RIGHT = ["Mbuti", "Israel_Natufian", "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]


def test_qpadm_facade():
    f2, bs = read_raw_fixture(os.path.join(GOLD, "f2_fit0_9pop.bin"))
    h = _core.upload_f2(f2, None, bs, NAMES)
    d = _core.run_qpadm(h, "England_BellBeaker", ["Czechia_EBA_CordedWare", "Turkey_N"], RIGHT)
    r = st.qpadm(h, target="England_BellBeaker", left=["Czechia_EBA_CordedWare", "Turkey_N"], right=RIGHT)
    assert r.weight == d["weight"] and r.p == d["p"] and r.status == st.Status.OK
    w = r.weights
    assert list(w.columns) == ["target", "left", "weight", "se", "z"] and len(w) == 2
    assert list(r.rankdrop.columns)[:4] == ["f4rank", "dof", "chisq", "p"]
    assert "feasible" in r.popdrop.columns
    res = st.qpadm_search(h, target="England_BellBeaker", right=RIGHT,
                          models=[["Czechia_EBA_CordedWare", "Turkey_N"], {"left": ["Turkey_N"]}])
    assert len(res) == 2 and res[0].weight == d["weight"]
    df = st.qpadm_search(h, target="England_BellBeaker", right=RIGHT, models=[["Turkey_N"]], as_dataframe=True)
    assert len(df) == 1
    qw = st.qpwave(h, left=["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N"], right=RIGHT)
    dw = _core.run_qpwave(h, ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N"], RIGHT)
    assert qw.f4rank == dw["f4rank"] and list(qw.per_rank["chisq"]) == dw["rank_chisq"]


def test_qpgraph_facade():
    f2, bs = read_raw_fixture(os.path.join(GOLD, "f2_qpgraph_9pop.bin"))
    h = _core.upload_f2(f2, None, bs, NAMES)
    c = qpgraph.enumerate_graphs(["Mbuti", "Han", "Iran_GanjDareh_N", "Israel_Natufian"], 1)[20]
    d = qpgraph.run_qpgraph(h, c["edges"])
    r = st.qpgraph(h, c["edges"])
    assert r.score == d["score"] and r.status == st.Status.OK
    assert list(r.weights.columns) == ["from", "to", "weight", "low", "high"]
    assert list(r.edges.columns) == ["from", "to", "length"] and len(r.edges) == len(d["edge_length"])
    assert np.isfinite(r.worst_residual_z)
