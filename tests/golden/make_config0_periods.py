"""Writes config0_periods.json (run from the repository root; about 2.5 minutes on 8 cores):

    python tests/golden/make_config0_periods.py

Config 0 without its Mach-1-singular corner system (7), solved by the oracle in canonical
arithmetic over the collapse-strength protocol's shape (transient periods, accessory reset,
recording periods); per-system digests of the results (digests.py). The GPU test runs the same
protocol and must reproduce every digest.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from bubbleensemble_implicitextension_amd import workloads as wl  # noqa: E402
from oracle import pyoracle  # noqa: E402
from digests import system_digests  # noqa: E402

EXCLUDED = [7]
TRANSIENT, RECORD = 4, 4


def main():
    pyoracle.build()
    ens = wl.config0()
    keep = [s for s in range(ens.NS) if s not in EXCLUDED]
    sub = ens.subset(keep)
    p = pyoracle.OracleProblem.from_ensemble(sub)
    for k in range(TRANSIENT + RECORD):
        if k == TRANSIENT:
            p.reset_accessories()
        p.solve(mode=pyoracle.CANONICAL, nthreads=0)
    d = system_digests(sub.NS, sub.N, p.ActualState, p.ActualTime, p.UnitAccessories,
                       p.IntegerUnitAccessories, p.stats_table())
    out = {"systems": keep, "transient": TRANSIENT, "record": RECORD, "digests": d,
           "events_total": int(p.event_counts().sum())}
    with open(os.path.join(HERE, "config0_periods.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(keep)} systems, {TRANSIENT}+{RECORD} periods, {out['events_total']} events")


if __name__ == "__main__":
    main()
