"""Batches of independent scenes stepped together (natray_batch_*): a ctypes view of the C ABI.

No counterpart in the reference, which steps one Sim per scene (src/sim.rs). Every scene is a
natray_amd.Sim of its own; the batch launches each kernel once for all of them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .sim import Sim


class Shape(C.Structure):
    """natray_batch_shape: workgroups per CU for P2G/G2P, the per-block kernels and clearing."""
    _fields_ = [("blocks_per_cu", C.c_uint32), ("waves_per_cu", C.c_uint32), ("cells_per_cu", C.c_uint32),
                ("pad", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"blocks_per_cu": self.blocks_per_cu, "waves_per_cu": self.waves_per_cu, "cells_per_cu": self.cells_per_cu}


def plan(weights, budget: int, min_each: int = 1, max_each=None) -> np.ndarray:
    """The workgroup shares a batch uses (natray_batch_plan)."""
    w = np.ascontiguousarray(weights, dtype=np.uint64)
    hi = None if max_each is None else np.ascontiguousarray(max_each, dtype=np.uint32)
    out = np.zeros(w.size, np.uint32)
    abi.check(abi.lib().natray_batch_plan(abi.ptr(w), abi.ptr(hi), w.size, budget, min_each, abi.ptr(out)))
    return out


class Batch:
    def __init__(self, sims: list[Sim]):
        self.sims = list(sims)  # the batch needs them alive
        handles = (C.c_void_p * len(self.sims))(*[s._h.value if hasattr(s._h, "value") else s._h for s in self.sims])
        h = C.c_void_p()
        abi.check(abi.lib().natray_batch_create(handles, len(self.sims), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().natray_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n_steps: int, iters: int = 0) -> None:
        abi.check(abi.lib().natray_batch_run(self._h, n_steps, iters))

    def set_graph(self, on: bool) -> None:
        abi.check(abi.lib().natray_batch_set_graph(self._h, 1 if on else 0))

    def set_solve_launch(self, mode: int) -> None:
        """abi.SOLVE_PLAIN (the default) or abi.SOLVE_COOPERATIVE (natray_batch_set_solve_launch)."""
        abi.check(abi.lib().natray_batch_set_solve_launch(self._h, int(mode)))

    def set_profiling(self, on: bool) -> None:
        abi.check(abi.lib().natray_batch_set_profiling(self._h, 1 if on else 0))

    def phase_times(self) -> dict:
        """{phase: (ms, launches)} of the kernel-by-kernel steps since profiling was switched on."""
        ms = np.zeros(abi.N_PHASES, np.float64)
        cnt = np.zeros(abi.N_PHASES, np.uint32)
        abi.check(abi.lib().natray_batch_phase_times(self._h, abi.ptr(ms), abi.ptr(cnt)))
        return {abi.PHASE_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(abi.N_PHASES)}

    def shape(self) -> dict:
        s = Shape()
        abi.check(abi.lib().natray_batch_get_shape(self._h, C.byref(s)))
        return s.as_dict()

    def set_shape(self, blocks_per_cu: int, waves_per_cu: int, cells_per_cu: int = 4) -> None:
        s = Shape(blocks_per_cu, waves_per_cu, cells_per_cu, 0)
        abi.check(abi.lib().natray_batch_set_shape(self._h, C.byref(s)))

    def tune(self, steps: int, iters: int = 0) -> dict:
        """The autotuner (natray_batch_tune): the chosen shape and each candidate's ms per step."""
        best = Shape()
        ms = np.zeros(16, np.float64)
        n = C.c_uint32()
        abi.check(abi.lib().natray_batch_tune(self._h, steps, iters, C.byref(best), abi.ptr(ms), ms.size, C.byref(n)))
        cands = [(b, w) for b in (4, 8, 16) for w in (4, 8, 16)]
        return {"best": best.as_dict(),
                "ms_per_step": {f"blocks{b}_waves{w}": float(ms[i]) for i, (b, w) in enumerate(cands[: n.value])}}
