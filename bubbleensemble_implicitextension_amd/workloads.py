"""Seeded synthetic workloads for the coupled Keller-Miksis bubble-ensemble solver.

The physical parameter construction follows the reference's collapse-strength driver
(CollapseStrength.cu:279-371 for the per-unit constants C0..C20, :398-408 for the coupling
matrix) and its file fixtures (M_ij = R_Ej^3 / (R_Ei^2 d_ij), zero diagonal, radii and
distances in micrometres).

Everything here is host-side numpy that produces the structure-of-arrays buffers a
``ProblemSolver`` (or a ``ProblemPool``) is filled with.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

# Water at 25 C, ambient pressure and polytropic exponent (CollapseStrength.cu:288-295).
PV = 3.166775638952003e03
RHO = 9.970639504998557e02
SIGMA = 0.071977583160056
MU = 8.902125058209557e-04
CL = 1.497251785455527e03
P_AMBIENT_BAR = 1.0
GAMMA = 1.4
PI = 3.14159265358979323846

NUP = 21          # unit parameters C0..C20 (CollapseStrength.cu:43)
NCOLL = 32        # recorded collapses per bubble (CollapseStrength.cu:22)
NUA = 2 * NCOLL   # unit accessories: NCOLL maxima, then NCOLL minima
NIUA = 1          # event counter (starts at -1, CollapseStrength.cu:359)


def unit_parameters(pa1_bar, f1_khz, re_um, pa2_bar=0.0, f2_khz=0.0, theta=0.0):
    """Per-unit constants C0..C20 of the nondimensional Keller-Miksis equation.

    Restates CollapseStrength.cu:321-357 elementwise (broadcasting over numpy arrays).
    P1 = pa1 [bar], P2 = f1 [kHz], P3 = pa2 [bar], P4 = f2 [kHz], P5 = theta, P6 = R_E [um].
    """
    P1 = np.asarray(pa1_bar, dtype=np.float64)
    P2 = np.asarray(f1_khz, dtype=np.float64)
    P3 = np.asarray(pa2_bar, dtype=np.float64)
    P4 = np.asarray(f2_khz, dtype=np.float64)
    P5 = np.asarray(theta, dtype=np.float64)
    P6 = np.asarray(re_um, dtype=np.float64)
    P1, P2, P3, P4, P5, P6 = np.broadcast_arrays(P1, P2, P3, P4, P5, P6)
    Pinf = P_AMBIENT_BAR * 1e5
    PA1 = P1 * 1e5
    PA2 = P3 * 1e5
    RE = P6 / 1e6
    f1 = 2.0 * PI * (P2 * 1000)
    f2 = 2.0 * PI * (P4 * 1000)
    L = 2.0 * PI / RE / f1
    C = np.empty(P1.shape + (NUP,), dtype=np.float64)
    C[..., 0] = (2.0 * SIGMA / RE + Pinf - PV) * L ** 2.0 / RHO
    C[..., 1] = (1.0 - 3.0 * GAMMA) * (2 * SIGMA / RE + Pinf - PV) * L / CL / RHO
    C[..., 2] = (Pinf - PV) * L ** 2.0 / RHO
    C[..., 3] = (2.0 * SIGMA / RE / RHO) * L ** 2.0
    C[..., 4] = (4.0 * MU / RHO / RE ** 2.0) * (2.0 * PI / f1)
    C[..., 5] = PA1 * L ** 2.0 / RHO
    C[..., 6] = PA2 * L ** 2.0 / RHO
    C[..., 7] = (RE * f1 * PA1 / RHO / CL) * L ** 2.0
    C[..., 8] = (RE * f2 * PA2 / RHO / CL) * L ** 2.0
    C[..., 9] = RE * f1 / (2.0 * PI) / CL
    C[..., 10] = 3.0 * GAMMA
    C[..., 11] = np.where(P2 != 0, P4 / np.where(P2 != 0, P2, 1.0), 0.0)
    C[..., 12] = P5
    C[..., 13] = 2.0 * PI / f1
    C[..., 14] = RE
    C[..., 15] = P1
    C[..., 16] = P2
    C[..., 17] = P3
    C[..., 18] = P4
    C[..., 19] = P5
    C[..., 20] = P6
    return C


def coupling_from_geometry(re_um, dist_um):
    """M_ij = R_Ej^3 / (R_Ei^2 d_ij), M_ii = 0 (the reference fixture's convention)."""
    re = np.asarray(re_um, dtype=np.float64)
    d = np.asarray(dist_um, dtype=np.float64)
    n = re.shape[-1]
    with np.errstate(divide="ignore"):
        M = re[..., None, :] ** 3 / (re[..., :, None] ** 2 * d)
    idx = np.arange(n)
    M[..., idx, idx] = 0.0
    return M


def line_geometry(n, spacing_um=150.0, jitter=0.10, rmin_um=2.0, rmax_um=8.0, seed=0):
    """Bubbles on a line with spacing*(1+U[-jitter, jitter]) gaps and log-uniform radii."""
    rng = np.random.default_rng(seed)
    gaps = spacing_um * (1.0 + rng.uniform(-jitter, jitter, size=n - 1))
    pos = np.concatenate([[0.0], np.cumsum(gaps)])
    re = np.exp(rng.uniform(math.log(rmin_um), math.log(rmax_um), size=n))
    d = np.abs(pos[:, None] - pos[None, :])
    return pos, re, d


def random_positions(n, min_sep_um, rng, box_um=None, sphere_radius_um=None, max_tries=200000):
    """Rejection-sampled 3D positions with a minimum pairwise separation."""
    pts = np.empty((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place bubbles with the requested separation")
        if sphere_radius_um is not None:
            p = rng.uniform(-sphere_radius_um, sphere_radius_um, size=3)
            if p @ p > sphere_radius_um ** 2:
                continue
        else:
            p = rng.uniform(0.0, box_um, size=3)
        if k and np.min(np.sum((pts[:k] - p) ** 2, axis=1)) < min_sep_um ** 2:
            continue
        pts[k] = p
        k += 1
    d = np.sqrt(np.sum((pts[:, None, :] - pts[None, :, :]) ** 2, axis=-1))
    return pts, d


@dataclasses.dataclass
class Ensemble:
    """Host buffers for a set of systems that share one unit count N."""

    name: str
    N: int
    pa_bar: np.ndarray        # [NS]
    f_khz: np.ndarray         # [NS]
    re_um: np.ndarray         # [N] or [NS, N]
    params: np.ndarray        # [NS, N, NUP]
    coupling: np.ndarray      # [N, N] (one matrix for all systems) or [NS, N, N]
    state0: np.ndarray        # [NS, N, 2]
    options: dict
    periods: int

    @property
    def NS(self):
        return self.pa_bar.shape[0]

    def unit_soa(self, arr):
        """[NS, N, K] -> reference unit-scope SoA [K][NS*N]."""
        ns, n, k = arr.shape
        return np.ascontiguousarray(arr.reshape(ns * n, k).T)

    def coupling_colmajor(self):
        """Column-major M[row + col*N] (CoupledSystems_PerBlock_Interface.cuh:1007)."""
        return np.ascontiguousarray(np.swapaxes(self.coupling, -1, -2)).reshape(-1)

    def subset(self, sids):
        sids = np.asarray(sids)
        re = self.re_um if self.re_um.ndim == 1 else self.re_um[sids]
        cpl = self.coupling if self.coupling.ndim == 2 else self.coupling[sids]
        return Ensemble(self.name + "-subset", self.N, self.pa_bar[sids], self.f_khz[sids], re,
                        self.params[sids], cpl, self.state0[sids], dict(self.options),
                        self.periods)


def default_options(rtol=1e-10, atol=1e-10):
    """Solver options of the collapse-strength workloads (CollapseStrength.cu:86-92, with the
    tolerances applied to both state components)."""
    return {
        "InitialTimeStep": 1e-3,
        "MinimumTimeStep": 1e-13,
        "MaximumTimeStep": 1.0e6,
        "TimeStepGrowLimit": 5.0,
        "TimeStepShrinkLimit": 0.1,
        "RelativeTolerance": [rtol, rtol],
        "AbsoluteTolerance": [atol, atol],
        "EventDirection": [-1],
        "EventTolerance": [1e-4],
    }


def _grid_ensemble(name, n, pa_vals, f_vals, seed, periods, dual, spacing=150.0):
    pos, re, d = line_geometry(n, spacing_um=spacing, seed=seed)
    # Frequency outer, amplitude inner (CollapseStrength.cu:306-308).
    F, PA = np.meshgrid(f_vals, pa_vals, indexing="ij")
    f = F.reshape(-1)
    pa = PA.reshape(-1)
    if dual:
        params = unit_parameters(pa[:, None], f[:, None], re[None, :], pa[:, None],
                                 2.0 * f[:, None], 0.0)
    else:
        params = unit_parameters(pa[:, None], f[:, None], re[None, :])
    M = coupling_from_geometry(re, d)
    ns = pa.shape[0]
    state0 = np.zeros((ns, n, 2))
    state0[..., 0] = 1.0  # CollapseStrength.cu:325-326
    return Ensemble(name, n, pa, f, re, params, M, state0, default_options(), periods)


GRID1_PA = np.linspace(0.5, 2.0, 128)
GRID1_F = np.linspace(20.0, 200.0, 128)
GRID0_INDEX = np.array([round(j * 127 / 7) for j in range(8)])


def config1(seed=2024, periods=1024, dual=False, n=32):
    """16384 systems on a 128x128 (amplitude x frequency) grid, N=32 bubbles on a line."""
    return _grid_ensemble("cfg1", n, GRID1_PA, GRID1_F, seed, periods, dual)


def config0(seed=2024, periods=64, dual=False, n=32):
    """64 systems: the 8x8 subset of config 1's grid at indices round(j*127/7)."""
    return _grid_ensemble("cfg0", n, GRID1_PA[GRID0_INDEX], GRID1_F[GRID0_INDEX], seed,
                          periods, dual)


def config0_in_config1_sids():
    """System ids of config 0's systems inside config 1 (frequency-major ordering)."""
    fi, ai = np.meshgrid(GRID0_INDEX, GRID0_INDEX, indexing="ij")
    return (fi * 128 + ai).reshape(-1)


def small_ensemble(n=8, ns=4, seed=7, pa=(0.8, 1.2), f=(40.0, 120.0), periods=2):
    """Tiny grid for quick parity checks."""
    pa_vals = np.linspace(pa[0], pa[1], max(1, int(round(math.sqrt(ns)))))
    f_vals = np.linspace(f[0], f[1], max(1, ns // max(1, len(pa_vals))))
    return _grid_ensemble("small", n, pa_vals, f_vals, seed, periods, False)


def random_positions_batch(ns, n, min_sep_um, rng, box_um=None, sphere_radius_um=None,
                           max_rounds=100000):
    """Rejection-sampled 3D positions for ns independent systems at once: bubble k of every
    system is redrawn until it keeps min_sep_um from bubbles 0..k-1 of its system.
    Returns positions [ns, n, 3] and distances [ns, n, n]."""
    pts = np.zeros((ns, n, 3))

    def draw(m):
        if sphere_radius_um is None:
            return rng.uniform(0.0, box_um, size=(m, 3))
        out = np.empty((m, 3))
        todo = np.arange(m)
        while todo.size:
            c = rng.uniform(-sphere_radius_um, sphere_radius_um, size=(todo.size, 3))
            ok = np.sum(c * c, axis=1) <= sphere_radius_um ** 2
            out[todo[ok]] = c[ok]
            todo = todo[~ok]
        return out

    for k in range(n):
        todo = np.arange(ns)
        rounds = 0
        while todo.size:
            rounds += 1
            if rounds > max_rounds:
                raise RuntimeError("could not place bubbles with the requested separation")
            cand = draw(todo.size)
            if k:
                d2 = np.sum((pts[todo, :k, :] - cand[:, None, :]) ** 2, axis=-1)
                ok = np.min(d2, axis=1) >= min_sep_um ** 2
            else:
                ok = np.ones(todo.size, dtype=bool)
            pts[todo[ok], k] = cand[ok]
            todo = todo[~ok]
    d = np.sqrt(np.sum((pts[:, :, None, :] - pts[:, None, :, :]) ** 2, axis=-1))
    return pts, d


CONFIG2_UNITS = (8, 16, 32, 64, 128)


def config2_cube_side_um(n, min_sep_um=50.0):
    """Edge of the cube that holds an N-bubble system of config 2 (the configuration leaves it
    open): six times the volume of N cubes of the minimum separation."""
    return min_sep_um * (6.0 * n) ** (1.0 / 3.0)


def config2_class(n, ns=4096, seed=2025, periods=64):
    """One size class of config 2: ns systems of n bubbles at random positions in a cube with
    50 um minimum separation, per-system coupling M_ij = R_j^3/(R_i^2 d_ij), radii log-uniform
    1-10 um, forcing amplitude 0.3-2.5 bar and frequency 20-500 kHz drawn per system."""
    rng = np.random.default_rng([seed, n])
    _, d = random_positions_batch(ns, n, 50.0, rng, box_um=config2_cube_side_um(n))
    re = np.exp(rng.uniform(math.log(1.0), math.log(10.0), size=(ns, n)))
    pa = rng.uniform(0.3, 2.5, size=ns)
    f = rng.uniform(20.0, 500.0, size=ns)
    params = unit_parameters(pa[:, None], f[:, None], re)
    M = coupling_from_geometry(re, d)
    state0 = np.zeros((ns, n, 2))
    state0[..., 0] = 1.0
    return Ensemble(f"cfg2-N{n}", n, pa, f, re, params, M, state0, default_options(), periods)


def config2(ns_per_class=4096, seed=2025, units=CONFIG2_UNITS):
    """Config 2: 5 x 4096 systems with N in {8, 16, 32, 64, 128}, one Ensemble per size."""
    return [config2_class(n, ns_per_class, seed) for n in units]


def config4(ns=4096, seed=2026, periods=256, n=256):
    """Config 4: ns systems of n = 256 bubbles at random 3D positions in a sphere of 1 mm
    diameter with at least 30 um separation (per-system geometry and coupling), radii
    log-uniform 2-8 um, forcing 2-3 bar at 25-50 kHz drawn per system."""
    rng = np.random.default_rng([seed, n])
    _, d = random_positions_batch(ns, n, 30.0, rng, sphere_radius_um=500.0)
    re = np.exp(rng.uniform(math.log(2.0), math.log(8.0), size=(ns, n)))
    pa = rng.uniform(2.0, 3.0, size=ns)
    f = rng.uniform(25.0, 50.0, size=ns)
    params = unit_parameters(pa[:, None], f[:, None], re)
    M = coupling_from_geometry(re, d)
    state0 = np.zeros((ns, n, 2))
    state0[..., 0] = 1.0
    return Ensemble(f"cfg4-N{n}", n, pa, f, re, params, M, state0, default_options(), periods)
