"""Training data and training of the pressure solve's warm-start network (natray_sim_set_warm_start,
natray_net_*). No counterpart in the reference, whose solve starts from the particle-carried
pressure (shaders/fluid.wgsl:450-489).

The simulator makes its own training set: with the warm start in mode 1 a step keeps, for each
16^3 tile with water, the network's inputs (initial residual / tile scale, cell kind), and after
the solve the target (solved - carried pressure) / tile scale on the water cells. Training is
Adam on random mini-batches of tiles, with the network's own GPU kernels.
"""
from __future__ import annotations

import time

import numpy as np

from .net import Net
from .sim import Sim


def collect(sim: Sim, steps: int, every: int = 1, iters: int = 0, min_water: int = 1, log=None):
    """Runs `steps` steps of `sim` with the warm start in mode 1 and returns the samples of every
    `every`-th step as (inputs [T][4096][2], targets [T][4096], masks [T][4096]); tiles with
    fewer than `min_water` water cells are dropped. Leaves the warm start off."""
    sim.set_warm_start(None, 1)
    xs, ys, ms = [], [], []
    total = 0
    for k in range(steps):
        sim.run(1, iters=iters)
        if k % every:
            continue
        x, y, m = sim.ws_samples()
        keep = m.sum(axis=1) >= min_water
        xs.append(x[keep])
        ys.append(y[keep])
        ms.append(m[keep])
        total += int(keep.sum())
        if log:
            log(f"collect: step {k + 1}/{steps}, {total} tiles")
    sim.set_warm_start(None, 0)
    if not xs:
        return (np.zeros((0, 4096, 2), np.float32), np.zeros((0, 4096), np.float32), np.zeros((0, 4096), np.float32))
    return np.concatenate(xs), np.concatenate(ys), np.concatenate(ms)


def train(net: Net, X: np.ndarray, Y: np.ndarray, M: np.ndarray, steps: int, batch: int = 32, lr: float = 1e-3,
          seed: int = 0, lr_final: float | None = None, log=None, log_every: int = 100) -> list[float]:
    """`steps` Adam steps on mini-batches of `batch` tiles drawn without replacement per epoch;
    the learning rate decays by cosine from lr to lr_final (lr / 20 by default). Returns the
    losses (masked MSE before each step)."""
    rng = np.random.default_rng(seed)
    T = X.shape[0]
    batch = min(batch, T, net.max_tiles)
    lr_final = lr / 20.0 if lr_final is None else lr_final
    order = rng.permutation(T)
    pos = 0
    losses = []
    t0 = time.perf_counter()
    for k in range(steps):
        if pos + batch > T:
            order = rng.permutation(T)
            pos = 0
        idx = np.sort(order[pos:pos + batch])
        pos += batch
        a = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * k / max(steps - 1, 1)))
        losses.append(net.train_step(X[idx], Y[idx], M[idx], lr=float(a)))
        if log and (k + 1) % log_every == 0:
            log(f"train: step {k + 1}/{steps} loss {np.mean(losses[-log_every:]):.4g} ({time.perf_counter() - t0:.1f} s)")
    return losses


def masked_loss(net: Net, X: np.ndarray, Y: np.ndarray, M: np.ndarray, batch: int | None = None) -> float:
    """The masked MSE of the network on a set of tiles (forward passes only)."""
    batch = batch or net.max_tiles
    num = den = 0.0
    for lo in range(0, X.shape[0], batch):
        out = net.forward(X[lo:lo + batch])
        m = M[lo:lo + batch].astype(np.float64)
        num += float((((out - Y[lo:lo + batch]) ** 2) * m).sum())
        den += float(m.sum())
    return num / max(den, 1.0)
