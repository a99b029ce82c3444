"""Writes config1_digests.json (run from the repository root; about ten minutes on 8 cores):

    python tests/golden/make_config1_digests.py

The headline workload as the benchmark runs it — config 1's 16384 systems of 32 bubbles, one
period from rest, 64 Hermite dense-output points — solved by the oracle in canonical
arithmetic; per-system digests (digests.py, with the system's dense-output times, states and
index appended), condensed to one digest per block of 64 consecutive systems. The GPU test runs
the same launch and must reproduce every block digest.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

from bubbleensemble_implicitextension_amd import workloads as wl  # noqa: E402
from digests import dense_digests, system_digests  # noqa: E402
from oracle import pyoracle  # noqa: E402

NDO = 64
BLOCK = 64


def block_digests(per_system):
    return [hashlib.sha256("".join(per_system[b:b + BLOCK]).encode()).hexdigest()
            for b in range(0, len(per_system), BLOCK)]


def main():
    pyoracle.build()
    ens = wl.config1()
    p = pyoracle.OracleProblem.from_ensemble(ens, NDO=NDO)
    p.opts.update({"DenseOutputInterpolation": 1, "DenseOutputMinimumTimeStep": 1.0 / NDO})
    t0 = time.perf_counter()
    p.solve(mode=pyoracle.CANONICAL, nthreads=0)
    d = system_digests(ens.NS, ens.N, p.ActualState, p.ActualTime, p.UnitAccessories,
                       p.IntegerUnitAccessories, p.stats_table())
    dd = dense_digests(ens.NS, ens.N, NDO, p.DenseOutputTimeInstances, p.DenseOutputStates,
                       p.DenseOutputIndex)
    per = [a + b for a, b in zip(d, dd)]
    out = {"ns": ens.NS, "ndo": NDO, "block": BLOCK, "blocks": block_digests(per),
           "events_total": int(p.event_counts().sum()),
           "accepted_total": int(p.stats_table()[:, 0].sum())}
    with open(os.path.join(HERE, "config1_digests.json"), "w") as f:
        json.dump(out, f, indent=0)
    print(f"{ens.NS} systems, {time.perf_counter() - t0:.0f} s, {out['events_total']} events")


if __name__ == "__main__":
    main()
