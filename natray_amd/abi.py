"""ctypes view of include/natray.h and the loader of libnatray.so.

The structures are byte-identical to the reference's Params (src/sim.rs:49-64), State
(src/sim.rs:66-94) and Particle (shaders/sim_defs.wgsl:61) layouts. Loading fails loudly when
the library has not been built: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnatray.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "natray.h")

# ---- constants (shaders/sim_defs.wgsl:98-157, shaders/fluid.wgsl) ---------------------------
NONE = 0xFFFFFFFF
PENDING = 0xFFFFFFFE
ACC_FIELDS = 19
CELL_FIELDS = 11
A_WC, A_PC, A_N, A_MF, A_MA, A_BU, A_SRC, A_MW, A_MV = 6, 7, 8, 9, 10, 11, 14, 15, 16
F_U, F_P, F_R, F_Z, F_D, F_Q, F_DIAG, F_TYPE = 0, 3, 4, 5, 6, 8, 9, 10
F_EXT, F_EXT_SET = 4, 7
GAS, WATER, SOLID, COVERED, RUBBER = 0.0, 1.0, 2.0, 3.0, 4.0
FP_W, FP_M, FP_P = float(1 << 24), float(1 << 16), float(1 << 16)

SOUND, CFL, DT_MAX, SOLVE_TOL = 1480.0, 0.8, 1e-3, 1e-5
MAX_ITERS, MAX_BATCH, MAX_BLOCKS = 4096, 256, 65535
RHO_WATER = 1000.0

SCENE_GROUND, SCENE_KNOT, SCENE_WALLS, SCENE_CAPSULE, SCENE_MASK, SCENE_AIR, SCENE_SPRAY = 1, 2, 4, 8, 16, 32, 64
SCENE_REFERENCE = SCENE_GROUND | SCENE_KNOT | SCENE_AIR | SCENE_SPRAY
PRECOND_BLOCK_SGS, PRECOND_JACOBI = 0, 1
CG_FP32, CG_MIXED, CG_PIPELINED = 0, 1, 2
SOLVE_PLAIN, SOLVE_COOPERATIVE, SOLVE_PLAIN_TEST_CHECK = 0, 1, 2
MAX_BULLETS, BULLET_PATCHES, BULLET_SECTIONS, BULLET_LOG = 4, 576, 32, 256

OK, E_OVERFLOW, E_NOMEM, E_HIP, E_ARG, E_SOLVER = 0, -1, -2, -3, -4, -5

PHASE_BEGIN, PHASE_BIN, PHASE_P2G, PHASE_GRID, PHASE_SOLVE, PHASE_PROJECT, PHASE_EXTEND, PHASE_G2P = range(8)
N_PHASES = 8
PHASE_NAMES = ["begin_step", "bin", "p2g", "grid_prep", "solve", "project", "extend", "g2p"]

BUF_PARTICLES, BUF_STATE, BUF_ACC, BUF_CF, BUF_BLOCK_LIST, BUF_STREAM = range(6)


def a_mom(d: int) -> int:
    return 2 * d


def a_wt(d: int) -> int:
    return 2 * d + 1


class Params(C.Structure):
    """== Params (src/sim.rs:49-64), 192 bytes."""

    _fields_ = [
        ("origin", C.c_float * 4),
        ("dims", C.c_uint32 * 4),
        ("bdims", C.c_uint32 * 4),
        ("gravity", C.c_float * 4),
        ("fluid", C.c_float * 4),
        ("wet", C.c_float * 4),
        ("limits", C.c_uint32 * 4),
        ("solve", C.c_float * 4),
        ("mem", C.c_uint32 * 4),
        ("rubber", C.c_float * 4),
        ("rubber2", C.c_float * 4),
        ("knot", C.c_float * 4),
    ]


class State(C.Structure):
    """== State (src/sim.rs:66-94), 96 bytes."""

    _fields_ = [
        ("t_ns", C.c_uint32),
        ("dt_ns", C.c_uint32),
        ("step", C.c_uint32),
        ("n_blocks", C.c_uint32),
        ("max_speed", C.c_float),
        ("n_blocks_prev", C.c_uint32),
        ("overflow", C.c_uint32),
        ("solve_iters", C.c_uint32),
        ("rz", C.c_float),
        ("alpha", C.c_float),
        ("b_norm", C.c_float),
        ("r_norm", C.c_float),
        ("free_dt_ns", C.c_uint32),
        ("iters_max", C.c_uint32),
        ("unconverged", C.c_uint32),
        ("beta", C.c_float),
        ("spray_count", C.c_uint32),
        ("mem_dt", C.c_float),
        ("cap_dt", C.c_float),
        ("mem_dt_seen", C.c_float),
        ("dt_flow", C.c_float),
        ("air_dt", C.c_float),
        ("air_dt_seen", C.c_float),
        ("mem_ticket", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Bullet(C.Structure):
    """== Bullet (bullet.wgsl:16-34, GpuBullet src/bullet.rs:19-28), 112 bytes."""

    _fields_ = [
        ("x", C.c_float * 4),
        ("v", C.c_float * 4),
        ("q", C.c_float * 4),
        ("l", C.c_float * 4),
        ("shape", C.c_float * 4),
        ("reach", C.c_float * 4),
        ("info", C.c_uint32 * 4),
    ]

    def as_array(self) -> np.ndarray:
        """The 28 words as float32 (info reinterpreted)."""
        return np.frombuffer(bytes(self), np.float32).copy()


class Scene(C.Structure):
    """natray_scene: scene semantics beyond Params (solids, preconditioner, precision)."""

    _fields_ = [
        ("flags", C.c_uint32),
        ("precond", C.c_uint32),
        ("precision", C.c_uint32),
        ("refresh_every", C.c_uint32),
        ("fixed_dt", C.c_float),
        ("mask_p", C.c_float),
        ("solve_groups", C.c_uint32),
        ("reserved0", C.c_uint32),
        ("capsule_a", C.c_float * 4),
        ("capsule_b", C.c_float * 4),
        ("capsule_v", C.c_float * 4),
        ("mask", C.c_float * 4),
        ("reserved1", C.c_uint32 * 4),
    ]


class WaterStats(C.Structure):
    """natray_water_stats: the water line of debug_stats (src/sim.rs:691-716)."""

    _fields_ = [
        ("mean", C.c_double * 3),
        ("vmean", C.c_double * 3),
        ("ke", C.c_double),
        ("volume_l", C.c_double),
        ("lo", C.c_float * 3),
        ("hi", C.c_float * 3),
        ("fastest", C.c_float * 3),
        ("vmax", C.c_float),
        ("alive", C.c_uint32),
        ("full_cells", C.c_uint32),
    ]


PARTICLE_FLOATS = 20  # 5 x vec4f, 80 bytes (shaders/sim_defs.wgsl:61)


class NatrayError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"natray error {code}: {msg}")
        self.code = code


_lib = None

_P = C.c_void_p
_SIGNATURES = {
    "natray_abi_sizes": (None, [C.POINTER(C.c_uint32)]),
    "natray_sim_create": (C.c_int, [C.POINTER(Params), _P, C.c_uint32, C.POINTER(Scene), C.POINTER(State), C.POINTER(_P)]),
    "natray_sim_destroy": (None, [_P]),
    "natray_sim_run": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(State)]),
    "natray_sim_advance": (C.c_int, [_P, C.c_double, C.c_uint32, C.POINTER(C.c_uint32)]),
    "natray_sim_step": (C.c_int, [_P, C.POINTER(State)]),
    "natray_sim_phase": (C.c_int, [_P, C.c_int]),
    "natray_sim_read_state": (C.c_int, [_P, C.POINTER(State)]),
    "natray_sim_write_state": (C.c_int, [_P, C.POINTER(State)]),
    "natray_sim_read_particles": (C.c_int, [_P, _P]),
    "natray_sim_write_particles": (C.c_int, [_P, _P]),
    "natray_sim_write_params": (C.c_int, [_P, C.POINTER(Params)]),
    "natray_sim_read_params": (C.c_int, [_P, C.POINTER(Params)]),
    "natray_sim_read_grid": (C.c_int, [_P, C.POINTER(C.c_uint32), _P, _P, _P, _P, _P]),
    "natray_sim_write_grid": (C.c_int, [_P, _P, _P]),
    "natray_sim_read_bins": (C.c_int, [_P, _P, _P]),
    "natray_sim_device_ptr": (C.c_void_p, [_P, C.c_int]),
    "natray_sim_set_profiling": (C.c_int, [_P, C.c_int]),
    "natray_sim_phase_times": (C.c_int, [_P, _P, _P]),
    "natray_sim_set_graph": (C.c_int, [_P, C.c_int]),
    "natray_sim_counters": (C.c_int, [_P, _P]),
    "natray_sim_run_log": (C.c_int, [_P, C.c_uint32, _P, C.POINTER(C.c_uint32)]),
    "natray_sim_water_stats": (C.c_int, [_P, _P]),
    "natray_sim_set_solve_launch": (C.c_int, [_P, C.c_int]),
    "natray_sim_set_warm_start": (C.c_int, [_P, _P, C.c_int]),
    "natray_sim_set_membrane": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P]),
    "natray_sim_pack_water": (C.c_int, [_P, _P]),
    "natray_sim_read_air": (C.c_int, [_P, _P]),
    "natray_sim_read_spray": (C.c_int, [_P, C.c_uint32, _P, _P, _P]),
    "natray_sim_read_wet": (C.c_int, [_P, _P]),
    "natray_sim_pack_keyframe": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, _P, _P]),
    "natray_sim_pack_keyframe_static": (C.c_int, [_P, C.c_uint32, _P, _P, C.c_uint32, _P, _P]),
    "natray_sim_set_bullets": (C.c_int, [_P, C.c_uint32, _P, _P]),
    "natray_sim_read_bullets": (C.c_int, [_P, _P, _P]),
    "natray_sim_read_pose_log": (C.c_int, [_P, _P]),
    "natray_sim_read_membrane": (C.c_int, [_P, _P]),
    "natray_sim_ws_samples": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "natray_sim_ws_stats": (C.c_int, [_P, _P]),
    "natray_sim_ws_guess": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "natray_batch_create": (C.c_int, [C.POINTER(_P), C.c_uint32, C.POINTER(_P)]),
    "natray_batch_destroy": (None, [_P]),
    "natray_batch_run": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "natray_batch_set_graph": (C.c_int, [_P, C.c_int]),
    "natray_batch_set_solve_launch": (C.c_int, [_P, C.c_int]),
    "natray_batch_set_shape": (C.c_int, [_P, _P]),
    "natray_batch_set_profiling": (C.c_int, [_P, C.c_int]),
    "natray_batch_phase_times": (C.c_int, [_P, _P, _P]),
    "natray_batch_get_shape": (C.c_int, [_P, _P]),
    "natray_batch_tune": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "natray_batch_plan": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "natray_plan_batch": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_uint32]),
    "natray_next_iters": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "natray_last_error": (C.c_char_p, []),
}


def exported_symbols() -> list[str]:
    """Every entry point include/natray.h declares, parsed from the header."""
    import re

    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(natray_[a-z_0-9]+)\s*\(", text)))


def lib() -> C.CDLL:
    """Loads libnatray.so (built by natray_amd.build); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m natray_amd.build` (there is no fallback path)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != OK:
        raise NatrayError(code, lib().natray_last_error().decode())


def abi_sizes() -> tuple[int, int, int, int]:
    out = (C.c_uint32 * 4)()
    lib().natray_abi_sizes(out)
    return tuple(int(x) for x in out)


def ptr(a: np.ndarray | None) -> C.c_void_p | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)
