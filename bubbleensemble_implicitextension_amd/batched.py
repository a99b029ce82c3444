"""Learned + batched solving of one ensemble (no reference counterpart).

The stiffest systems (by a cost proxy that needs no run) take the learned BiCG warm start, the
others the plain start, split by cost into a tier of one system per wave and a bulk of two per
wave; the parts are ordinary solver objects on their own streams, launched together. This
module needs no torch: the network is a trained parameter file.
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


def split_systems(ens, heavy, medium, cost):
    """(heavy, medium, light) system indices, each sorted: the `heavy` largest cost_proxy, then
    the `medium` largest `cost` among the others, then the rest (at least one system)."""
    if ens.NS < 2:
        raise ValueError("a learned + batched split needs at least two systems")
    cost = np.asarray(cost, dtype=np.float64)
    if cost.shape != (ens.NS,):
        raise ValueError("cost must give one value per system")
    heavy = max(1, min(int(heavy), ens.NS - 1))
    medium = max(0, min(int(medium), ens.NS - 1 - heavy))
    by_proxy = np.argsort(-cost_proxy(ens), kind="stable")
    rest = np.sort(by_proxy[heavy:])
    rest = rest[np.argsort(-cost[rest], kind="stable")]
    return np.sort(by_proxy[:heavy]), np.sort(rest[:medium]), np.sort(rest[medium:])


class LearnedBatchedSolver:
    """Learned + batched: the ensemble split into three tiers that run as two solver objects
    (each on its own stream), launched together, costliest first:

    * heavy: the `heavy` systems with the largest cost_proxy (the stiff corner of the forcing
      grid, near the Keller-Miksis Mach-1 singularity) on the learned BiCG start, one system per
      wave (role pair). There the learned start cuts the BiCG iterations and shortens the
      systems that bound the launch; elsewhere it saves none and costs its network evaluation.
    * medium: the next `medium` systems by `cost` on the plain start, one system per wave (role
      pair: the shortest time per system), in the same launch as the heavy ones (the learned
      kernel with the learned start switched off per system).
    * light: the rest on the plain start on the system-pair layout (two systems per wave: the
      right-hand side, reductions and controller run once for both), the highest throughput.

    cost ranks the systems for the medium tier (default cost_proxy(ens); the counters of a
    previous solve give measured_cost). Each launch is persistent and hands its systems out
    costliest first; the heavy and medium waves issue at the top priority. With network=None
    the heavy systems take the plain start too (a plain batched split, bit-identical to the
    plain solve system by system).

    Solve() is asynchronous like ProblemSolver.Solve(); synchronise() waits for every part;
    results() returns state, accessories, event counts and counters in the ensemble's order.
    """

    def __init__(self, ens, network, heavy=64, end_time=1.0, ndo=0, options=None,
                 light_layout=LAYOUT_SYSTEM_PAIR, cost=None, medium=0):
        cost = cost_proxy(ens) if cost is None else np.asarray(cost, dtype=np.float64)
        self.heavy, self.medium, self.light = split_systems(ens, heavy, medium, cost)
        self.NS, self.N = ens.NS, ens.N
        self.parts = []
        # The heavy and medium systems share one launch (one wave each, all started at once, so
        # no two of these top-priority waves meet on a SIMD); the light systems are the other.
        one_per_wave = np.sort(np.concatenate([self.heavy, self.medium]))
        plan = [(one_per_wave, network, LAYOUT_AUTO, True),
                (self.light, None, light_layout, False)]
        for sids, net, layout, urgent in plan:
            sub = ens.subset(sids)
            s = collapse.make_solver(sub, NDO=ndo)
            collapse.fill(s, sub)
            s.host_view(V.TimeDomain)[sub.NS:] = end_time
            s.SolverOption(O.KernelLayout, layout)
            if urgent:
                # The heavy and medium systems bound the launch: their waves take the top issue
                # priority from the first accepted step on (the light part keeps the default
                # step threshold).
                s.SolverOption(O.PriorityStepThreshold, 1)
            if options:
                collapse.apply_options(s, options)
                if "ActiveSystems" in options:
                    # Systems 0 .. n-1 of the ensemble are the part's first local ones (each
                    # part keeps its systems in ascending order).
                    s.SolverOption(O.ActiveSystems,
                                   int(np.count_nonzero(sids < int(options["ActiveSystems"]))))
            if net is not None:
                s.set_warm_start_network(net)
                s.SolverOption(O.WarmStart, 1)
                if sids.size != self.heavy.size:
                    # only the heavy systems take the learned start, the medium ones the plain
                    # start (bit-identical to the plain solve)
                    s.set_warm_start_systems(np.isin(sids, self.heavy))
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
