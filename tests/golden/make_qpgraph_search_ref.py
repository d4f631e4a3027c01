"""Write the reference CPU topology search's per-candidate scores on the AT2 qpGraph
fixture (at2/f2_qpgraph_9pop.bin, search pops Mbuti, Han, Iran_GanjDareh_N,
Israel_Natufian, Czechia_EBA_CordedWare, default fit options, no hill-climb) to
qpgraph_search_ref_5pop.csv, using oracle/_ref/libsteppe_ref.so (the reference's
run_qpgraph_search, src/core/qpadm/qpgraph_search.cpp:290, compiled from its sources by
oracle/Makefile).  Run from the repository root: python tests/golden/make_qpgraph_search_ref.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from fixtures import read_raw_fixture  # noqa: E402
import oracle_lib  # noqa: E402

NAMES = ["England_BellBeaker", "Czechia_EBA_CordedWare", "Turkey_N", "Mbuti", "Israel_Natufian",
         "Iran_GanjDareh_N", "Han", "Papuan", "Karitiana"]
POPS = ["Mbuti", "Han", "Iran_GanjDareh_N", "Israel_Natufian", "Czechia_EBA_CordedWare"]


def main():
    ref = oracle_lib.ref()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (needs the reference sources)")
    f2, bs = read_raw_fixture(os.path.join(HERE, "at2", "f2_qpgraph_9pop.bin"))
    summary, cands = ref.qg_search(f2, None, bs, NAMES, POPS)
    out = os.path.join(HERE, "qpgraph_search_ref_5pop.csv")
    with open(out, "w") as f:
        summary.pop("fit_all_wall_ms")
        f.write("# " + json.dumps(summary, sort_keys=True) + "\n")
        f.write("nadmix,id,hash,score,restart_spread\n")
        for c in cands:
            f.write(f"{c['nadmix']},{c['id']},{c['hash']},{c['score']!r},{c['restart_spread']!r}\n")
    print(f"wrote {out}: {len(cands)} candidates, best {summary['best_score']!r} ({summary['best_hash']})")


if __name__ == "__main__":
    main()
