"""Shared helpers of the tests: small scenes, grid views keyed by global cell coordinates."""
from __future__ import annotations

import numpy as np

from natray_amd import abi, scenes


def box_scene(P: np.ndarray, side: int = 16, fixed_dt: float = 1e-4, gravity: float = 9.81, flags: int = abi.SCENE_WALLS,
              precond: int = abi.PRECOND_BLOCK_SGS, precision: int = abi.CG_FP32, iters: int = 200, tol: float = 1e-6,
              capacity: int | None = None, cfl: float = 0.5) -> scenes.Scene:
    """A unit box of side^3 cells holding the particle records P (N x 20)."""
    P = np.ascontiguousarray(P, np.float32)
    cap = capacity or (side // 4) ** 3
    U = scenes.box_params(side, P.shape[0], cap, gravity=gravity, cfl=cfl, tol=tol, iters=iters)
    sc = abi.Scene()
    sc.flags = flags
    sc.precond = precond
    sc.precision = precision
    sc.fixed_dt = fixed_dt
    return scenes.Scene(U, P, sc, scenes.no_limits_state(), {"workload": "test box", "dx": 1.0 / side})


def records(pos, vel=None, C=None, pres=None) -> np.ndarray:
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    P = np.zeros((n, abi.PARTICLE_FLOATS), np.float32)
    P[:, :3] = pos
    P[:, 3] = 1.0
    P[:, 7] = 1.0
    if vel is not None:
        P[:, 4:7] = np.asarray(vel, np.float32).reshape(n, 3)
    if C is not None:
        C = np.asarray(C, np.float32).reshape(n, 3, 3)
        P[:, 8:11] = C[:, 0]
        P[:, 12:15] = C[:, 1]
        P[:, 16:19] = C[:, 2]
    if pres is not None:
        P[:, 11] = np.asarray(pres, np.float32).reshape(n)
    return P


def cell_coords(grid: dict, bdims) -> np.ndarray:
    """(64n, 3) global cell coordinates of the compacted grid's cells."""
    bl = grid["block_list"].astype(np.int64)
    by, bz = int(bdims[1]), int(bdims[2])
    b = np.stack([bl // (bz * by), (bl // bz) % by, bl % bz], axis=1)
    l = np.arange(64)
    loc = np.stack([l // 16, (l // 4) % 4, l % 4], axis=1)
    return (b[:, None, :] * 4 + loc[None, :, :]).reshape(-1, 3)


def keyed(grid: dict, bdims, field_rows: np.ndarray) -> dict:
    """{(cx, cy, cz): index} of the compacted grid."""
    cc = cell_coords(grid, bdims)
    return {tuple(int(v) for v in c): i for i, c in enumerate(cc)}


def run_phases(sim, phases) -> None:
    for ph in phases:
        sim.phase(ph)
