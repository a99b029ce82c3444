"""Host logic of the split solver (CPU only): the cost measured from a solve's counters, the
costliest-first hand-out order that keeps the systems of one wave together, and the three-part
partition of an ensemble."""
import numpy as np
import pytest

from bubbleensemble_implicitextension_amd import batched
from bubbleensemble_implicitextension_amd import workloads as wl


def test_measured_cost_counts_stages_iterations_and_attempts():
    st = np.zeros((3, 32), dtype=np.int64)
    st[0, 2], st[0, 3], st[0, 10] = 6, 0, 1     # one attempt, six stages, no BiCG iteration
    st[1, 2], st[1, 3], st[1, 10] = 0, 1, 0     # one BiCG iteration alone
    st[2, 2], st[2, 3], st[2, 10] = 12, 30, 2
    n = 32
    c = batched.measured_cost(st, n)
    assert c[0] == 6 * (4 * n * n + 47 * n) + 140 * n
    assert c[1] == 4 * n * n + 21 * n
    assert c[2] == 2 * c[0] + 30 * c[1]


def test_hand_out_order_pairs_stay_together_costliest_first():
    cost = np.array([1.0, 5.0, 2.0, 2.0, 9.0, 0.5, 3.0])
    o = batched.hand_out_order(cost, 2)
    # pairs (0,1) max 5, (2,3) max 2, (4,5) max 9; system 6 has no partner and stays last
    assert o.tolist() == [4, 5, 0, 1, 2, 3, 6]
    assert sorted(o.tolist()) == list(range(7))
    o1 = batched.hand_out_order(cost, 1)
    assert o1.tolist() == [4, 1, 6, 2, 3, 0, 5]
    assert batched.hand_out_order([7.0], 2).tolist() == [0]
    assert batched.hand_out_order(np.zeros(0), 2).tolist() == []


def test_three_part_partition():
    """heavy = largest cost proxy, medium = next by the given cost, light = the rest; together
    every system exactly once, each part sorted."""
    ens = wl.config0()
    proxy = batched.cost_proxy(ens)
    cost = np.random.default_rng(3).permutation(ens.NS).astype(np.float64)
    h, m, l = batched.split_systems(ens, 5, 11, cost)
    assert len(h) == 5 and len(m) == 11 and len(l) == ens.NS - 16
    assert sorted(np.concatenate([h, m, l]).tolist()) == list(range(ens.NS))
    for part in (h, m, l):
        assert np.all(np.diff(part) > 0)
    # heavy: the five largest proxies (the strongest, slowest drives)
    assert np.min(proxy[h]) >= np.max(proxy[np.concatenate([m, l])])
    # medium: the largest costs among the rest
    assert np.min(cost[m]) > np.max(cost[l])
    # sizes are clamped so that the light part keeps at least one system
    h2, m2, l2 = batched.split_systems(ens, 60, 60, cost)
    assert (len(h2), len(m2), len(l2)) == (60, 3, 1)


def test_split_rejects_a_wrong_cost_shape():
    ens = wl.config0()
    with pytest.raises(ValueError):
        batched.split_systems(ens, 4, 4, np.ones(ens.NS + 1))
    with pytest.raises(ValueError):
        batched.split_systems(ens.subset([0]), 1, 0, np.ones(1))
