"""Writes tests/golden/flop_counts.json: algorithmic FP64 op counts per node of
one RHS evaluation, by category, from the oracle's op-counting instantiation
(+,-,*,/, sqrt, log, exp, pow count one each) on each workload's initial
state.  Counts depend on the state only through the log-mean branch taken.
The N=7 Taylor-Green count is taken on an 8^3-element mesh of the same box
(the 32^3 mesh is too large for the counting scalar); per-node counts differ
only through log-mean branch choices.

"ensemble64" holds, for the 64-case ensemble of 2-D Euler cases (seed
20261202), the per-RHS totals summed over all cases and the node count.

Usage: python tests/golden/make_flop_counts.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.dirname(HERE)]

import oracle_lib as O  # noqa: E402
import bench  # noqa: E402
from prandtl_amd import cases  # noqa: E402


def main():
    out = {}
    for name, cfg in bench.WORKLOADS.items():
        cfg = bench.resolve(dict(cfg))
        if name == "tgv3d_n7":
            cfg["nelem"] = (8, 8, 8)
        for quirks in (True, False):
            orc = bench.make_oracle(cfg, quirks, 1)
            u = bench.initial_state(name, cfg)
            c = orc.count_flops(u)
            nodes = orc.NE * orc.K
            out["%s_%s" % (name, "faithful" if quirks else "corrected")] = {
                k: v / nodes for k, v in c.items()}
    tot, nodes = {}, 0
    for c in cases.ensemble_cases(64, 20261202):
        kw = cases.operator_kwargs(c)
        kw.pop("precision")
        orc = O.Oracle(**kw)
        for k, v in orc.count_flops(cases.ensemble_state(c)).items():
            tot[k] = tot.get(k, 0.0) + v
        nodes += orc.NE * orc.K
    out["ensemble64"] = dict(tot, nodes=nodes)
    with open(os.path.join(HERE, "flop_counts.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
