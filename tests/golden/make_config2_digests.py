"""Writes config2_digests.json (run from the repository root; under a minute on 8 cores):

    python tests/golden/make_config2_digests.py

Sixteen systems of each config-2 size class (N = 8, 16, 32, 64, 128 bubbles at random positions
in a cube, per-system coupling, 0.3-2.5 bar at 20-500 kHz), each solved by the oracle in
canonical arithmetic over the first quarter of a period from rest; per-system digests of the
results (digests.py). The GPU test runs the five classes as one pool and must reproduce every
digest.
"""
import json
import os
import sys
from concurrent.futures import ProcessPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

NS, SEED, END = 16, 12, 0.25
UNITS = (8, 16, 32, 64, 128)


def solve_one(job):
    n, sid = job
    from bubbleensemble_implicitextension_amd import workloads as wl
    from digests import system_digests
    from oracle import pyoracle

    ens = wl.config2_class(n, NS, SEED)
    sub = ens.subset([sid])
    sub.coupling = sub.coupling[0]
    p = pyoracle.OracleProblem.from_ensemble(sub)
    p.TimeDomain[1:] = END
    p.solve(mode=pyoracle.CANONICAL, nthreads=1)
    d = system_digests(1, sub.N, p.ActualState, p.ActualTime, p.UnitAccessories,
                       p.IntegerUnitAccessories, p.stats_table())
    return d[0], int(p.event_counts().sum())


def main():
    from oracle import pyoracle

    pyoracle.build()
    jobs = [(n, sid) for n in UNITS for sid in range(NS)]
    with ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        res = list(ex.map(solve_one, jobs))
    out = {"ns": NS, "seed": SEED, "end_time": END, "units": list(UNITS),
           "digests": {str(n): [res[i * NS + k][0] for k in range(NS)]
                       for i, n in enumerate(UNITS)},
           "events_total": sum(r[1] for r in res)}
    with open(os.path.join(HERE, "config2_digests.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(UNITS)} classes x {NS} systems over [0, {END}], {out['events_total']} events")


if __name__ == "__main__":
    main()
