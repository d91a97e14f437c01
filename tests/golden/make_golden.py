"""Writes the golden vectors under tests/golden/ from an independent float64 formulation.

The reference ships no fixtures for its water substep, so these are computed here from the
mathematics the kernels implement, written differently from the oracle:

  p2g_*.npz   particle-to-grid transfer sums (shaders/fluid.wgsl:229-303) of a few seeded
              particles, evaluated in float64 with the textbook piecewise quadratic B-spline
              N(r) = 3/4 - r^2 (|r| < 1/2), (3/2 - |r|)^2 / 2 (1/2 <= |r| < 3/2), 0 otherwise,
              summed over every lattice node of each particle (no tiles, no per-axis weights).
              Stored per global cell and field as the real-valued sum, the number of
              contributions and the sum of their magnitudes, which bound the fixed-point
              rounding of the oracle and the kernels.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# Lattice offsets of u, v, w faces and cell centres inside a cell (shaders/sim_defs.wgsl:227-231).
LATTICES = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0), (0.5, 0.5, 0.5)]


def bspline(r: np.ndarray) -> np.ndarray:
    a = np.abs(r)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def p2g_float64(pos, vel, C, pres, origin, h):
    """Real-valued transfer sums keyed by (field, cx, cy, cz)."""
    sums, count, mag = {}, {}, {}

    def add(key, val):
        sums[key] = sums.get(key, 0.0) + val
        count[key] = count.get(key, 0) + 1
        mag[key] = mag.get(key, 0.0) + abs(val)

    for p in range(len(pos)):
        x = (pos[p] - origin) / h  # in cells
        for lat, o in enumerate(LATTICES):
            o = np.asarray(o)
            lo = np.floor(x - o - 1.5).astype(int)
            for i in range(lo[0], lo[0] + 4):
                for j in range(lo[1], lo[1] + 4):
                    for k in range(lo[2], lo[2] + 4):
                        node = np.array([i, j, k], float) + o  # node position in cells
                        w = bspline(node[0] - x[0]) * bspline(node[1] - x[1]) * bspline(node[2] - x[2])
                        if w == 0.0:
                            continue
                        if lat < 3:
                            val = vel[p][lat] + np.dot(C[p][lat], (node - x) * h)
                            add((2 * lat, i, j, k), w * val)
                            add((2 * lat + 1, i, j, k), w)
                        else:
                            add((7, i, j, k), w * pres[p] * 1e-6)
                            add((6, i, j, k), w)
        c = np.floor(x).astype(int)
        add((8, c[0], c[1], c[2]), 1.0)
    keys = sorted(sums)
    arr = np.array([[k[0], k[1], k[2], k[3]] for k in keys], dtype=np.int32)
    return arr, np.array([sums[k] for k in keys]), np.array([count[k] for k in keys]), np.array([mag[k] for k in keys])


def make_p2g(name: str, n: int, seed: int, side: int = 16) -> None:
    rng = np.random.default_rng(seed)
    h = np.float32(1.0 / side)
    origin = np.zeros(3)
    # Particles well inside the box, so no stencil reaches the domain edge.
    pos = rng.uniform(4.0, side - 4.0, size=(n, 3)).astype(np.float32) * h
    vel = rng.normal(0.0, 1.0, size=(n, 3)).astype(np.float32)
    C = rng.normal(0.0, 20.0, size=(n, 3, 3)).astype(np.float32)
    pres = rng.normal(0.0, 5e4, size=n).astype(np.float32)
    key, val, cnt, mag = p2g_float64(pos.astype(np.float64), vel.astype(np.float64), C.astype(np.float64),
                                     pres.astype(np.float64), origin, float(h))
    np.savez_compressed(os.path.join(HERE, name), side=side, h=h, pos=pos, vel=vel, C=C, pres=pres, key=key, val=val,
                        cnt=cnt, mag=mag)
    print(f"{name}: {n} particles, {len(key)} cell sums")


if __name__ == "__main__":
    make_p2g("p2g_one.npz", 1, 11)
    make_p2g("p2g_seven.npz", 7, 12)
    make_p2g("p2g_fifty.npz", 50, 13)
