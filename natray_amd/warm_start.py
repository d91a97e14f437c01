"""Training data and training of the pressure solve's warm-start network (natray_sim_set_warm_start,
natray_net_*). No counterpart in the reference, whose solve starts from the particle-carried
pressure (shaders/fluid.wgsl:450-489).

The simulator makes its own training set: with the warm start in mode 1 a step keeps, for each
16^3 tile with water, the network's inputs (x = initial residual / tile scale, cell kind), the
solve's operator on the tile (diagonal, water-neighbour bits) and, after the solve, the target
(solved - carried pressure) / tile scale. The network is trained with its own GPU kernels on
the residual its guess leaves, x - A o: the quantity by which the step decides on a guess.
"""
from __future__ import annotations

import time

import numpy as np

from .net import Net
from .sim import Sim

KEYS = ("inputs", "targets", "masks", "diag", "nbm")


def collect(sim: Sim, steps: int, every: int = 1, iters: int = 0, min_water: int = 1, log=None) -> dict:
    """Runs `steps` steps of `sim` with the warm start in mode 1 and returns the samples of every
    `every`-th step (natray_sim_ws_samples, concatenated over steps); tiles with fewer than
    `min_water` water cells are dropped. Leaves the warm start off."""
    sim.set_warm_start(None, 1)
    parts = {k: [] for k in KEYS}
    total = 0
    for k in range(steps):
        sim.run(1, iters=iters)
        if k % every:
            continue
        smp = sim.ws_samples()
        keep = smp["masks"].sum(axis=1) >= min_water
        for key in KEYS:
            parts[key].append(smp[key][keep])
        total += int(keep.sum())
        if log:
            log(f"collect: step {k + 1}/{steps}, {total} tiles")
    sim.set_warm_start(None, 0)
    return {k: np.concatenate(v) if v else np.zeros((0,) + ((4096, 2) if k == "inputs" else (4096,)), np.float32)
            for k, v in parts.items()}


def subset(data: dict, idx) -> dict:
    return {k: v[idx] for k, v in data.items()}


def apply_operator(diag: np.ndarray, nbm: np.ndarray, o: np.ndarray) -> np.ndarray:
    """(A o) per tile, [T][4096]: diag o minus the water neighbours' o inside the tile (bits
    2 axis for the - side, 2 axis + 1 for the + side; voxel v = 256 x + 16 y + z); 0 off water."""
    T = o.shape[0]
    o3 = o.reshape(T, 16, 16, 16).astype(np.float64)
    b = nbm.reshape(T, 16, 16, 16)
    s = np.zeros_like(o3)
    for axis in range(3):
        lo = [slice(None)] * 4
        hi = [slice(None)] * 4
        lo[axis + 1] = slice(1, None)
        hi[axis + 1] = slice(None, -1)
        lo, hi = tuple(lo), tuple(hi)
        s[lo] += np.where(b[lo] & (1 << (2 * axis)), o3[hi], 0.0)
        s[hi] += np.where(b[hi] & (1 << (2 * axis + 1)), o3[lo], 0.0)
    out = diag.astype(np.float64) * o.astype(np.float64) - s.reshape(T, 4096)
    return np.where(diag > 0, out, 0.0)


def residual_ratio(net: Net, data: dict, batch: int | None = None) -> float:
    """|x - A o|^2 / |x|^2 over the water cells of a set of tiles: the fraction of the tiles'
    initial residual the network's guess leaves (1 for a zero guess)."""
    batch = batch or net.max_tiles
    num = den = 0.0
    X = data["inputs"]
    for lo in range(0, X.shape[0], batch):
        sl = slice(lo, lo + batch)
        o = net.forward(X[sl])
        x = X[sl, :, 0].astype(np.float64)
        w = data["diag"][sl] > 0
        e = x - apply_operator(data["diag"][sl], data["nbm"][sl], o)
        num += float((e[w] ** 2).sum())
        den += float((x[w] ** 2).sum())
    return num / max(den, 1e-300)


def train(net: Net, data: dict, steps: int, batch: int = 32, lr: float = 1e-3, seed: int = 0,
          lr_final: float | None = None, loss: str = "residual", log=None, log_every: int = 100) -> list[float]:
    """`steps` Adam steps on mini-batches of `batch` tiles drawn without replacement per epoch,
    on the residual loss (default) or the masked MSE against the targets ("mse"); the learning
    rate decays by cosine from lr to lr_final (lr / 20 by default). Returns the losses (before
    each step)."""
    assert loss in ("residual", "mse")
    rng = np.random.default_rng(seed)
    X = data["inputs"]
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
        a = float(lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * k / max(steps - 1, 1))))
        if loss == "residual":
            losses.append(net.train_step_residual(X[idx], data["diag"][idx], data["nbm"][idx], lr=a))
        else:
            losses.append(net.train_step(X[idx], data["targets"][idx], data["masks"][idx], lr=a))
        if log and (k + 1) % log_every == 0:
            log(f"train: step {k + 1}/{steps} loss {np.mean(losses[-log_every:]):.4g} ({time.perf_counter() - t0:.1f} s)")
    return losses
