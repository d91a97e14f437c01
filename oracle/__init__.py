"""CPU oracle of the natray water substep — TEST INFRASTRUCTURE ONLY.

Python view of oracle/liboracle.so (natray_oracle.c, built by oracle/Makefile). Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker. `Oracle` has
the same methods as natray_amd.Sim so parity tests read the same on both sides.

    o = Oracle.from_scene(scene, precision="f32")   # FP32-faithful restatement
    o = Oracle.from_scene(scene, precision="f64")   # FP64 oracle of the error metrics
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from natray_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        for pfx in ("o32_", "o64_"):
            sig = {
                "create": (P, [C.POINTER(abi.Params), P, C.c_uint32, C.POINTER(abi.Scene), C.POINTER(abi.State)]),
                "destroy": (None, [P]),
                "phase": (C.c_int, [P, C.c_int]),
                "run": (C.c_int, [P, C.c_uint32, C.c_uint32, C.c_uint32]),
                "read_state": (None, [P, C.POINTER(abi.State)]),
                "write_state": (None, [P, C.POINTER(abi.State)]),
                "read_particles": (None, [P, P]),
                "read_particles_f64": (None, [P, P]),
                "write_particles": (None, [P, P]),
                "n_blocks": (C.c_uint32, [P]),
                "read_grid": (None, [P, P, P, P, P, P]),
                "read_cf_f64": (None, [P, P]),
                "write_grid": (None, [P, P, P]),
                "read_bins": (None, [P, P, P]),
                "solver_scalars": (None, [P, P]),
                "precondition": (None, [P, P, P, P, C.c_uint32]),
                "set_threads": (None, [C.c_int]),
                "set_guess": (None, [P, P]),
                "residual": (None, [P, P, P]),
                "set_membrane": (C.c_int, [P, C.c_uint32, P, P, P, P]),
                "read_membrane": (None, [P, P]),
                "pack_water": (None, [P, P]),
                "read_air": (None, [P, P]),
                "read_spray": (C.c_uint32, [P, C.c_uint32, P, P]),
                "read_wet": (None, [P, P]),
                "elementary": (None, [C.c_int, P, P, P, C.c_uint32]),
                "pack_keyframe": (C.c_uint32, [P, C.c_uint32, P, P, C.c_uint32, P]),
                "set_bullets": (C.c_int, [P, C.c_uint32, P, P]),
                "read_bullets": (None, [P, P, P]),
                "read_pose_log": (None, [P, P]),
            }
            for name, (res, args) in sig.items():
                fn = getattr(L, pfx + name)
                fn.restype = res
                fn.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().o32_set_threads(n)
    lib().o64_set_threads(n)


def precondition(r: np.ndarray, diag: np.ndarray, water: np.ndarray, mode: int, precision: str = "f32") -> np.ndarray:
    """The oracle's preconditioner applied to one block's residual r[64]."""
    r = np.ascontiguousarray(r, np.float32)
    diag = np.ascontiguousarray(diag, np.float32)
    water = np.ascontiguousarray(water, np.int32)
    z = np.empty(64, np.float32)
    fn = lib().o32_precondition if precision == "f32" else lib().o64_precondition
    fn(abi.ptr(r), abi.ptr(diag), abi.ptr(water), abi.ptr(z), mode)
    return z


def elementary(which: str, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """The spray's exp, log, cos or pow(a, b), or the bullet's sin or asin, in float32, as the
    kernels evaluate them."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(a if b is None else b, np.float32)
    out = np.empty_like(a)
    lib().o32_elementary({"exp": 0, "log": 1, "cos": 2, "pow": 3, "sin": 4, "asin": 5}[which], abi.ptr(a), abi.ptr(b), abi.ptr(out), a.size)
    return out


def water_stats(particles: np.ndarray, origin, h: float) -> dict:
    """The water line of debug_stats (src/sim.rs:691-716) over particle records [n][20]: the
    alive ones (float 7 > 0.5, :693); their box (:702-703); mean position and velocity (:704-705,
    here summed in float64); the sum of 0.5 |v|^2 in float64 of the f32 |v|^2 (:706); the first
    particle whose f32 speed sqrt((x x + y y) + z z) is the strict maximum above 0 (:707-710);
    and the cells floor((x - origin) / h), computed in f32, that hold at least 4 (:711-716), with
    their volume in litres (:735)."""
    P = np.asarray(particles, np.float32).reshape(-1, 20)
    A = P[P[:, 7] > np.float32(0.5)]
    x, v = A[:, 0:3], A[:, 4:7]
    n = len(A)
    out = {"alive": n}
    if n:
        out["lo"], out["hi"] = [float(t) for t in x.min(0)], [float(t) for t in x.max(0)]
        out["mean"] = [float(t) for t in x.astype(np.float64).sum(0) / n]
        out["vmean"] = [float(t) for t in v.astype(np.float64).sum(0) / n]
    else:
        big = float(np.finfo(np.float32).max)
        out["lo"], out["hi"] = [big] * 3, [-big] * 3
        out["mean"] = out["vmean"] = [0.0] * 3
    v2 = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    out["ke"] = float((0.5 * v2.astype(np.float64)).sum())
    sp = np.sqrt(v2)
    if n and float(sp.max()) > 0.0:
        k = int(np.argmax(sp))  # the first of the maxima
        out["vmax"], out["fastest"] = float(sp[k]), [float(t) for t in x[k]]
    else:
        out["vmax"], out["fastest"] = 0.0, [0.0, 0.0, 0.0]
    o = np.asarray(origin, np.float32)[:3]
    cells = np.floor((x - o) / np.float32(h)).astype(np.int64)
    full = 0
    if n:
        _, counts = np.unique(cells, axis=0, return_counts=True)
        full = int((counts >= 4).sum())
    out["full_cells"] = full
    out["volume_l"] = full * float(h) ** 3 * 1e3
    return out


class Oracle:
    def __init__(self, params: abi.Params, particles: np.ndarray, scene: abi.Scene | None = None,
                 init: abi.State | None = None, precision: str = "f32"):
        assert precision in ("f32", "f64")
        self.pfx = "o32_" if precision == "f32" else "o64_"
        self._f = lambda name: getattr(lib(), self.pfx + name)
        P = np.ascontiguousarray(particles, dtype=np.float32)
        self.n = int(P.shape[0])
        self.params = params
        self.S = int(params.bdims[0] * params.bdims[1] * params.bdims[2])
        h = self._f("create")(C.byref(params), abi.ptr(P), self.n, C.byref(scene) if scene is not None else None,
                              C.byref(init) if init is not None else None)
        if not h:
            raise MemoryError("oracle allocation failed")
        self._h = C.c_void_p(h)

    @classmethod
    def from_scene(cls, s, precision: str = "f32") -> "Oracle":
        return cls(s.params, s.particles, s.scene, s.init, precision)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._f("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def phase(self, k: int) -> int:
        return int(self._f("phase")(self._h, k))

    def run(self, n_steps: int, iters: int, stop_ns: int = 0xFFFFFFFF) -> abi.State:
        self._f("run")(self._h, n_steps, iters, stop_ns)
        return self.state()

    def state(self) -> abi.State:
        st = abi.State()
        self._f("read_state")(self._h, C.byref(st))
        return st

    def write_state(self, st: abi.State) -> None:
        self._f("write_state")(self._h, C.byref(st))

    def particles(self) -> np.ndarray:
        P = np.empty((self.n, abi.PARTICLE_FLOATS), np.float32)
        self._f("read_particles")(self._h, abi.ptr(P))
        return P

    def particles_f64(self) -> np.ndarray:
        P = np.empty((self.n, abi.PARTICLE_FLOATS), np.float64)
        self._f("read_particles_f64")(self._h, abi.ptr(P))
        return P

    def write_particles(self, P: np.ndarray) -> None:
        P = np.ascontiguousarray(P, dtype=np.float32)
        self._f("write_particles")(self._h, abi.ptr(P))

    def n_blocks(self) -> int:
        return int(self._f("n_blocks")(self._h))

    def grid(self) -> dict:
        n = self.n_blocks()
        bl = np.empty(max(n, 1), np.uint32)
        nb = np.empty(max(n, 1) * 27, np.uint32)
        bs = np.empty(n + 1, np.uint32)
        acc = np.empty((abi.ACC_FIELDS, max(n, 1) * 64), np.int32)
        cf = np.empty((abi.CELL_FIELDS, max(n, 1) * 64), np.float32)
        self._f("read_grid")(self._h, abi.ptr(bl), abi.ptr(nb), abi.ptr(bs), abi.ptr(acc), abi.ptr(cf))
        return {"n_blocks": n, "block_list": bl[:n], "block_nb": nb[: 27 * n].reshape(n, 27), "block_start": bs,
                "acc": acc[:, : 64 * n], "cf": cf[:, : 64 * n]}

    def cf_f64(self) -> np.ndarray:
        n = self.n_blocks()
        cf = np.empty((abi.CELL_FIELDS, max(n, 1) * 64), np.float64)
        self._f("read_cf_f64")(self._h, abi.ptr(cf))
        return cf[:, : 64 * n]

    def write_grid(self, acc: np.ndarray | None, cf: np.ndarray | None) -> None:
        a = None if acc is None else np.ascontiguousarray(acc, dtype=np.int32)
        f = None if cf is None else np.ascontiguousarray(cf, dtype=np.float32)
        self._f("write_grid")(self._h, abi.ptr(a), abi.ptr(f))

    def bins(self) -> tuple[np.ndarray, np.ndarray]:
        bp = np.empty(self.S, np.uint32)
        so = np.empty(max(self.n, 1), np.uint32)
        self._f("read_bins")(self._h, abi.ptr(bp), abi.ptr(so))
        return bp, so[: self.n]

    def set_guess(self, guess: np.ndarray | None) -> None:
        """The next solve starts from `guess` (bdims.w * 64 cells) on the water cells."""
        g = None
        if guess is not None:
            g = np.ascontiguousarray(guess, dtype=np.float32)
            assert g.size == int(self.params.bdims[3]) * 64
        self._f("set_guess")(self._h, abi.ptr(g))

    def residual(self, p: np.ndarray | None = None) -> np.ndarray:
        """b - A p on the cells in use (0 off water) for the grid as it stands before a solve;
        p = None is the particle-carried pressure."""
        n = self.n_blocks()
        r = np.zeros(max(n, 1) * 64, np.float32)
        q = None
        if p is not None:
            q = np.ascontiguousarray(p, dtype=np.float32)
            assert q.size == int(self.params.bdims[3]) * 64
        self._f("residual")(self._h, abi.ptr(q), abi.ptr(r))
        return r[: 64 * n]

    def set_membrane(self, x: np.ndarray, v: np.ndarray, mass: np.ndarray, push: np.ndarray) -> None:
        """The membrane's vertices for the coming steps (positions and velocities [n][4], masses
        [n], push and area [n][2])."""
        x, v = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(v, np.float32)
        mass, push = np.ascontiguousarray(mass, np.float32), np.ascontiguousarray(push, np.float32)
        self._mem_n = int(mass.size)
        assert self._f("set_membrane")(self._h, self._mem_n, abi.ptr(x), abi.ptr(v), abi.ptr(mass), abi.ptr(push)) == 0

    def air(self) -> np.ndarray:
        n = self.n_blocks()
        out = np.zeros((5, 64 * n), np.float32)
        self._f("read_air")(self._h, abi.ptr(out))
        return out

    def spray(self, max_records: int = 1 << 20) -> dict:
        """The spray's records (NATRAY_SCENE_SPRAY): count (State.spray_count), the first
        min(count, cap, max_records) records [k][12] and where each landed [k][4], as doubles."""
        rec = np.zeros((max_records, 12), np.float64)
        landed = np.zeros((max_records, 4), np.float64)
        n = int(self._f("read_spray")(self._h, max_records, abi.ptr(rec), abi.ptr(landed)))
        return {"count": int(self.state().spray_count), "records": rec[:n], "landed": landed[:n]}

    def wet(self) -> np.ndarray:
        """The wet map (limits.x^2 words, f32 bits of the first touch; 0xffffffff untouched)."""
        res = int(self.params.limits[0])
        out = np.zeros((res, res), np.uint32)
        self._f("read_wet")(self._h, abi.ptr(out))
        return out

    def set_bullets(self, bullets: list, patches: np.ndarray) -> None:
        """Bullets for the coming steps: abi.Bullet records and patches [n][576][8]."""
        n = len(bullets)
        arr = (abi.Bullet * max(n, 1))(*bullets)
        pt = np.ascontiguousarray(patches, np.float32)
        assert pt.shape == (n, abi.BULLET_PATCHES, 8)
        assert self._f("set_bullets")(self._h, n, C.cast(arr, C.c_void_p), abi.ptr(pt)) == 0
        self._bullets_n = n

    def bullets(self) -> tuple[list, np.ndarray]:
        """The bullets (as float32 records) and their cavity sections [n][32][12] as doubles."""
        n = getattr(self, "_bullets_n", 0)
        arr = (abi.Bullet * max(n, 1))()
        sec = np.zeros((n, abi.BULLET_SECTIONS, 12), np.float64)
        self._f("read_bullets")(self._h, C.cast(arr, C.c_void_p), abi.ptr(sec))
        return list(arr)[:n], sec

    def pose_log(self) -> np.ndarray:
        """The bullets' pose log [n][256][12] (float32)."""
        out = np.zeros((getattr(self, "_bullets_n", 0), abi.BULLET_LOG, 12), np.float32)
        self._f("read_pose_log")(self._h, abi.ptr(out))
        return out

    def pack_water(self) -> np.ndarray:
        out = np.zeros((self.n, 2), np.uint32)
        self._f("pack_water")(self._h, abi.ptr(out))
        return out

    def pack_keyframe(self, alive: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
        """A whole keyframe (words) and each spray slot's weight and time made [n][2]."""
        a = None if alive is None else np.ascontiguousarray(alive, np.uint32)
        nt = 0 if a is None else int(a.size)
        total = int(self._f("pack_keyframe")(self._h, nt, abi.ptr(a), None, 0, None))
        words = np.zeros(total, np.uint32)
        n = words.size  # at least the slot count
        sw = np.zeros((max(int(self.state().spray_count), 1), 2), np.float32)
        self._f("pack_keyframe")(self._h, nt, abi.ptr(a), abi.ptr(words), total, abi.ptr(sw))
        return words, sw[: int(words[0])] if n else sw[:0]

    def membrane_water(self) -> np.ndarray:
        out = np.zeros((getattr(self, "_mem_n", 0), 4), np.float32)
        self._f("read_membrane")(self._h, abi.ptr(out))
        return out

    def solver_scalars(self) -> dict:
        out = np.zeros(8, np.float64)
        self._f("solver_scalars")(self._h, abi.ptr(out))
        keys = ["rz", "alpha", "beta", "b_norm", "r_norm", "solve_iters", "max_speed", "cg_iters"]
        return dict(zip(keys, out.tolist()))
