"""Learned + batched solving of one ensemble (no reference counterpart).

The costliest systems (by a cost proxy that needs no run) take the learned BiCG warm start, the
others the plain start; both parts are ordinary solver objects on their own streams, launched
together. This module needs no torch: the network is a trained parameter file.
"""
from __future__ import annotations

import os

import numpy as np

from . import _native
from . import collapse
from . import workloads as wl
from .problem_solver import LAYOUT_AUTO, LAYOUT_SYSTEM_PAIR
from .problem_solver import ListOfSolverOptions as O
from .problem_solver import ListOfVariables as V

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_NETWORK = os.path.join(HERE, "data", "warm_start_n32.npy")


def cost_proxy(ens):
    """Per-system cost estimate of the collapse workloads without running them: forcing
    amplitude over frequency (a stronger, slower drive gives larger expansions, more violent
    collapses and stiffer coupling solves)."""
    return np.asarray(ens.pa_bar, dtype=np.float64) / np.asarray(ens.f_khz, dtype=np.float64)


def measured_cost(stats, n):
    """Per-system cost measured by a previous solve: the FP64 FLOPs of its RK stages, BiCG
    iterations and step attempts (per stage 4N^2 + 47N, per iteration 4N^2 + 21N, per attempt
    140N) from the solver's counters ([NS][32], stats())."""
    st = np.asarray(stats, dtype=np.float64)
    return (st[:, 2] * (4 * n * n + 47 * n) + st[:, 3] * (4 * n * n + 21 * n) +
            st[:, 10] * 140 * n)


def hand_out_order(cost, group=1):
    """Launch order of a part's systems (indices into it), costliest first: systems are taken in
    index order in groups of `group` (the systems that share a wave; neighbouring grid points
    cost about the same), and the whole groups are sorted by their costliest member; an
    incomplete last group stays last, so that every wave keeps its group."""
    cost = np.asarray(cost, dtype=np.float64)
    n = cost.shape[0]
    ng = n // group
    gcost = cost[: ng * group].reshape(ng, group).max(axis=1)
    gorder = np.argsort(-gcost, kind="stable")
    idx = (gorder[:, None] * group + np.arange(group)[None, :]).reshape(-1)
    return np.concatenate([idx, np.arange(ng * group, n)]).astype(np.int64)


class LearnedBatchedSolver:
    """Learned + batched: the `heavy` systems with the largest cost_proxy run with the learned
    BiCG start (one system per wave, role pair), the others with the plain start on the
    system-pair layout (two systems per wave: the right-hand side, reductions and controller run
    once for both), as two solver objects (each on its own stream) launched together, the heavy
    ones first. The bulk pays no network evaluation where the start saves nothing and runs at
    the system-pair layout's throughput; the learned start shortens the systems that bound the
    launch. With network=None the heavy systems take the plain start too (a plain batched split,
    bit-identical to the plain solve system by system).

    Solve() is asynchronous like ProblemSolver.Solve(); synchronise() waits for both parts;
    results() returns state, accessories, event counts and counters in the ensemble's order.
    """

    def __init__(self, ens, network, heavy=64, end_time=1.0, ndo=0, options=None,
                 light_layout=LAYOUT_SYSTEM_PAIR, cost=None):
        """cost: per-system cost estimate that ranks the systems (default cost_proxy(ens); the
        counters of a previous solve give measured_cost). Each part's persistent launch hands
        its systems out costliest first, in pairs of neighbouring grid points on the
        system-pair layout."""
        if ens.NS < 2:
            raise ValueError("a learned + batched split needs at least two systems")
        cost = cost_proxy(ens) if cost is None else np.asarray(cost, dtype=np.float64)
        if cost.shape != (ens.NS,):
            raise ValueError("cost must give one value per system")
        order = np.argsort(-cost, kind="stable")
        heavy = max(1, min(int(heavy), ens.NS - 1))
        self.heavy = np.sort(order[:heavy])
        self.light = np.sort(order[heavy:])
        self.NS, self.N = ens.NS, ens.N
        self.parts = []
        for sids, net, layout in ((self.heavy, network, LAYOUT_AUTO),
                                  (self.light, None, light_layout)):
            sub = ens.subset(sids)
            s = collapse.make_solver(sub, NDO=ndo)
            collapse.fill(s, sub)
            s.host_view(V.TimeDomain)[sub.NS:] = end_time
            s.SolverOption(O.KernelLayout, layout)
            if options:
                collapse.apply_options(s, options)
            if net is not None:
                s.set_warm_start_network(net)
                s.SolverOption(O.WarmStart, 1)
            s.set_launch_order(hand_out_order(cost[sids], 2 if layout == LAYOUT_SYSTEM_PAIR else 1))
            s.SynchroniseFromHostToDevice(V.All)
            self.parts.append((sids, s))

    def Solve(self):
        for _, s in self.parts:
            s.Solve()

    def synchronise(self):
        for _, s in self.parts:
            s.SynchroniseDevice()

    def reset_stats(self):
        for _, s in self.parts:
            s.reset_stats()

    def checkpoint_save(self):
        for _, s in self.parts:
            s.checkpoint_save()

    def checkpoint_restore(self):
        for _, s in self.parts:
            s.checkpoint_restore()

    def results(self, dense=False):
        """State, accessories, event counts and counters in the ensemble's order; with dense,
        also the dense-output states [points, 2, NS, N] (points stored by every system)."""
        out = {"state": np.zeros((self.NS, self.N, 2)), "acc": np.zeros((self.NS, self.N, wl.NUA)),
               "events": np.zeros((self.NS, self.N), dtype=np.int64),
               "stats": np.zeros((self.NS, _native.NSTAT), dtype=np.int64)}
        dense_parts = []
        for sids, s in self.parts:
            s.SynchroniseFromDeviceToHost(V.All)
            if dense:
                s.SynchroniseFromDeviceToHost(V.DenseOutput)
            s.InsertSynchronisationPoint()  # the copies are asynchronous on the solver's stream
            s.SynchroniseSolver()
            out["state"][sids] = s.state()
            out["acc"][sids] = s.accessories()
            out["events"][sids] = s.event_counts()
            out["stats"][sids] = s.stats()
            if dense:
                idx = s.host_view(V.DenseIndex)
                dense_parts.append((sids, s, int(idx.min()) if idx.size else 0))
        if dense and dense_parts:
            n = min(k for _, _, k in dense_parts)
            ds = np.zeros((n, 2, self.NS, self.N))
            for sids, s, _ in dense_parts:
                ds[:, :, sids, :] = s.host_view(V.DenseState)[: n * 2 * len(sids) * self.N].reshape(
                    n, 2, len(sids), self.N)
            out["dense_state"] = ds
        return out

    def close(self):
        for _, s in self.parts:
            s.close()
        self.parts = []
