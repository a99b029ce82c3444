"""Per-system SHA-256 digests of a solve's results in the reference's SoA buffers: the state,
time, accessories, integer accessories and the counters of each system. Shared by the fixture
generator (oracle) and the GPU test, so that a multi-period run of many systems is compared bit
for bit through a few kilobytes of committed data."""
import hashlib

import numpy as np


def system_digests(NS, N, state, time, acc, iacc, stats):
    state = np.asarray(state, dtype=np.float64).reshape(-1, NS, N)
    acc = np.asarray(acc, dtype=np.float64).reshape(-1, NS, N)
    iacc = np.asarray(iacc, dtype=np.int32).reshape(-1, NS, N)
    time = np.asarray(time, dtype=np.float64).reshape(NS)
    stats = np.asarray(stats, dtype=np.int64).reshape(NS, -1)
    out = []
    for k in range(NS):
        h = hashlib.sha256()
        for part in (state[:, k, :], time[k:k + 1], acc[:, k, :], iacc[:, k, :], stats[k]):
            h.update(np.ascontiguousarray(part).tobytes())
        out.append(h.hexdigest())
    return out
