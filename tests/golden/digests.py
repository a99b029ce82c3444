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


def dense_digests(NS, N, NDO, times, states, index, UD=2):
    """Per-system digests of dense output: the system's NDO output times, its [NDO][UD][N]
    states (reference layout [idx][comp][NS*N]) and its index."""
    times = np.asarray(times, dtype=np.float64).reshape(NDO, NS)
    states = np.asarray(states, dtype=np.float64).reshape(NDO, UD, NS, N)
    index = np.asarray(index, dtype=np.int32).reshape(NS)
    out = []
    for k in range(NS):
        h = hashlib.sha256()
        for part in (times[:, k], states[:, :, k, :], index[k:k + 1]):
            h.update(np.ascontiguousarray(part).tobytes())
        out.append(h.hexdigest())
    return out
