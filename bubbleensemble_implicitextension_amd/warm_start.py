"""Learned BiCG warm start: sample collection, GPU training and evaluation.

    python -m bubbleensemble_implicitextension_amd.warm_start [--samples 1000000] [--out PATH]

Pipeline (no reference counterpart; the reference starts BiCG from the uncoupled acceleration,
Steppers.cuh:494-499):

1. collect: the plain solver in sample-collection mode runs first periods of a held-out
   ensemble (same physics, different geometry/radius seed) and records, every k-th RK stage,
   32 rows per system: 5 features from the system definition's WarmStartFeatures hook, the
   target x/b, x, b and the uncoupled acceleration (device kernel, C ABI bex_enable_sampling).
2. train: Adam on the mean squared error of the network factor against x/b, with the
   library's own device kernels (bex_mlp_loss_grad, bex_mlp_adam); torch provides device
   memory and the stream only.
3. use: ProblemSolver.set_warm_start_network + SolverOption(WarmStart, 1): the solver kernel
   evaluates the network on MFMA in every RK stage and starts BiCG from x0 = b * y.
"""
from __future__ import annotations

import argparse
import ctypes
import os
import time

import numpy as np

# torch supplies device memory and streams for training. Import it before libbex.so is loaded
# so that the process has one HIP runtime (torch's); loading libbex first binds the system
# runtime and torch then reports no device.
try:
    import torch as _torch_module  # noqa: F401
except ImportError:  # training and batch prediction then report that they need torch
    _torch_module = None

from . import _native
from . import collapse
from . import workloads as wl
from .batched import DEFAULT_NETWORK, LearnedBatchedSolver, cost_proxy  # noqa: F401
from .problem_solver import MLP_NPARAM, SAMPLE_FLOATS
from .problem_solver import ListOfSolverOptions as O
from .problem_solver import ListOfVariables as V

TRAIN_SEED = 7001  # held-out geometry/radius seed of the training ensemble


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("warm-start training needs a GPU")
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def collect_samples(ens, capacity, every, end_time=1.0):
    """First period of every system of `ens` (N = 32) with sample collection on."""
    s = collapse.make_solver(ens)
    collapse.fill(s, ens)
    s.host_view(V.TimeDomain)[ens.NS:] = end_time
    s.SynchroniseFromHostToDevice(V.All)
    s.enable_sampling(capacity, every)
    s.Solve()
    s.SynchroniseDevice()
    rows = s.samples()
    s.close()
    return rows


def init_params(seed=0):
    """He-uniform initial parameters (same rule as the host oracle's init_params)."""
    rng = np.random.default_rng(seed)
    H = 64
    p = np.zeros(MLP_NPARAM, dtype=np.float32)
    p[0:6 * H] = rng.uniform(-1.0, 1.0, 6 * H) * 0.5
    p[5 * H:6 * H] = 0.0
    lim2 = np.sqrt(6.0 / H)
    p[6 * H:6 * H + H * H] = rng.uniform(-lim2, lim2, H * H)
    o = 6 * H + H * H + H
    p[o:o + H] = rng.uniform(-1.0, 1.0, H) / np.sqrt(H)
    p[o + H] = 1.0
    return p


class DeviceTrainer:
    """Adam on device-resident samples with the library's training kernels."""

    def __init__(self, rows, params, batch=4096, lr=2e-3, seed=0):
        torch = _torch()
        self.torch = torch
        self.lib = _native.load()
        self.rows = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.float32)).cuda()
        self.n = self.rows.shape[0]
        self.stride = self.rows.shape[1]
        self.p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).cuda()
        self.m = torch.zeros_like(self.p)
        self.v = torch.zeros_like(self.p)
        self.g = torch.zeros_like(self.p)
        self.loss = torch.zeros(1, dtype=torch.float32, device="cuda")
        self.batch = min(batch, self.n)
        self.ws = torch.zeros(int(self.lib.bex_mlp_workspace_floats(self.batch)),
                              dtype=torch.float32, device="cuda")
        self.lr = lr
        self.step_count = 0
        self.rng = np.random.default_rng(seed)

    def loss_grad(self, start, count):
        view = self.rows[start:start + count]
        _native.check(self.lib.bex_mlp_loss_grad(_ptr(view), self.stride, count, _ptr(self.p),
                                                 _ptr(self.g), _ptr(self.loss), _ptr(self.ws),
                                                 _stream(self.torch)))
        return self.loss

    def full_loss(self, chunk=65536):
        """Mean squared error over all rows (chunked; sample-weighted mean of chunk losses)."""
        ws = self.torch.zeros(int(self.lib.bex_mlp_workspace_floats(min(chunk, self.n))),
                              dtype=self.torch.float32, device="cuda")
        g = self.torch.zeros_like(self.p)
        loss = self.torch.zeros(1, dtype=self.torch.float32, device="cuda")
        total = 0.0
        for start in range(0, self.n, chunk):
            count = min(chunk, self.n - start)
            view = self.rows[start:start + count]
            _native.check(self.lib.bex_mlp_loss_grad(_ptr(view), self.stride, count,
                                                     _ptr(self.p), _ptr(g), _ptr(loss), _ptr(ws),
                                                     _stream(self.torch)))
            total += float(loss.item()) * count
        return total / self.n

    def epoch(self):
        # Shuffle rows on the device once per epoch (gather by a host-made permutation).
        perm = self.torch.from_numpy(self.rng.permutation(self.n)).cuda()
        self.rows = self.rows[perm].contiguous()
        for start in range(0, self.n - self.batch + 1, self.batch):
            self.loss_grad(start, self.batch)
            self.step_count += 1
            _native.check(self.lib.bex_mlp_adam(_ptr(self.p), _ptr(self.g), _ptr(self.m),
                                                _ptr(self.v), self.step_count, self.lr, 0.9,
                                                0.999, 1e-8, _stream(self.torch)))
        return float(self.loss.item())

    def params(self):
        return self.p.cpu().numpy()


def predict(params, rows, mfma=False):
    """Network output for host rows through the FP32 training forward or the MFMA path."""
    torch = _torch()
    lib = _native.load()
    r = torch.from_numpy(np.ascontiguousarray(rows, dtype=np.float32)).cuda()
    out = torch.zeros(r.shape[0], dtype=torch.float32, device="cuda")
    p = np.ascontiguousarray(params, dtype=np.float32)
    if mfma:
        _native.check(lib.bex_mlp_predict_mfma(_ptr(r), r.shape[1], r.shape[0],
                                               p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               _ptr(out), _stream(torch)))
    else:
        pd = torch.from_numpy(p).cuda()
        _native.check(lib.bex_mlp_predict(_ptr(r), r.shape[1], r.shape[0], _ptr(pd), _ptr(out),
                                          _stream(torch)))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def guess_errors(rows, y):
    """Median relative error of the learned guess b*y and of the plain guess (uncoupled
    acceleration) against the converged solution x."""
    x, b, xu = rows[:, 6].astype(np.float64), rows[:, 7].astype(np.float64), rows[:, 8].astype(np.float64)
    ok = np.isfinite(x) & (np.abs(x) > 0)
    yl = np.clip(y.astype(np.float64), -5.0, 5.0)
    learned = np.abs(b * yl - x)[ok] / np.abs(x[ok])
    plain = np.abs(xu - x)[ok] / np.abs(x[ok])
    return float(np.median(learned)), float(np.median(plain))


def train_network(samples=1_000_000, epochs=8, batch=4096, lr=2e-3, seed=0, systems=4096,
                  every=None, log=print):
    """Collect samples from a held-out ensemble and train; returns (params, report)."""
    ens = wl.config1(seed=TRAIN_SEED)
    sids = np.sort(np.random.default_rng(seed).choice(ens.NS, size=min(systems, ens.NS),
                                                      replace=False))
    sub = ens.subset(sids)
    if every is None:
        # ~15k RK stages per system-period: spread the samples over the whole period.
        per_system = max(1, samples // (32 * sub.NS))
        every = max(1, 15000 // per_system)
    t0 = time.perf_counter()
    rows = collect_samples(sub, capacity=samples, every=every)
    t1 = time.perf_counter()
    rows = rows[np.all(np.isfinite(rows[:, :6]), axis=1)]
    rng = np.random.default_rng(seed + 1)
    idx = rng.permutation(len(rows))
    nval = max(32, len(rows) // 20)
    val, tr = rows[idx[:nval]], rows[idx[nval:]]
    trainer = DeviceTrainer(tr, init_params(seed), batch=batch, lr=lr, seed=seed)
    losses = [trainer.full_loss()]
    log(f"# initial training-set loss {losses[0]:.5f}")
    for ep in range(epochs):
        trainer.epoch()
        losses.append(trainer.full_loss())
        log(f"# epoch {ep}: training-set loss {losses[-1]:.5f}")
    t2 = time.perf_counter()
    p = trainer.params()
    y = predict(p, val)
    le, pe = guess_errors(val, y)
    report = {"samples": int(len(rows)), "every": int(every), "collect_s": t1 - t0,
              "train_s": t2 - t1, "epochs": epochs, "initial_loss": losses[0],
              "final_loss": losses[-1],
              "val_median_rel_err_learned": le, "val_median_rel_err_plain": pe}
    return p, report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--systems", type=int, default=4096)
    ap.add_argument("--out", default=DEFAULT_NETWORK)
    args = ap.parse_args()
    p, report = train_network(samples=args.samples, epochs=args.epochs, systems=args.systems)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    np.save(args.out, p.astype(np.float32))
    print(report, flush=True)


if __name__ == "__main__":
    main()
