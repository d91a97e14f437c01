"""`Sim`: the reference's simulation handle (src/sim.rs:149-205) over the C ABI of libnatray.

    sim = Sim.from_scene(scenes.water_sphere())
    sim.advance(t_target=5e-3, budget=1000)      # == Sim::advance (src/sim.rs:498-522)
    state = sim.run(10, iters=200)                # == Sim::run (src/sim.rs:585-648)
    P = sim.particles()                           # (N, 20) float32 records, the Particle layout

Every call goes through libnatray.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .params import SimParams
from .scenes import Scene, reference_scene


class Sim:
    def __init__(self, params: abi.Params, particles: np.ndarray, scene: abi.Scene | None = None,
                 init: abi.State | None = None):
        L = abi.lib()
        P = np.ascontiguousarray(particles, dtype=np.float32)
        assert P.ndim == 2 and P.shape[1] == abi.PARTICLE_FLOATS
        self.n = int(P.shape[0])
        self.params = params
        self.scene = scene
        h = C.c_void_p()
        abi.check(L.natray_sim_create(C.byref(params), abi.ptr(P), self.n, C.byref(scene) if scene is not None else None,
                                      C.byref(init) if init is not None else None, C.byref(h)))
        self._h = h
        self.S = int(params.bdims[0] * params.bdims[1] * params.bdims[2])
        self.dx = float(params.origin[3])
        self.grid_origin = tuple(float(x) for x in params.origin[:3])
        self.dims = tuple(int(x) for x in params.dims[:3])

    @classmethod
    def from_scene(cls, s: Scene) -> "Sim":
        return cls(s.params, s.particles, s.scene, s.init)

    @classmethod
    def new(cls, p: SimParams) -> "Sim":
        """Sim::new(&SimParams) (src/sim.rs:208) with the round-balloon stand-in profile."""
        return cls.from_scene(reference_scene(p))

    # ---- lifetime ---------------------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_h", None):
            abi.lib().natray_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- stepping ---------------------------------------------------------------------------
    def run(self, n_steps: int, iters: int = 0, stop_ns: int = 0xFFFFFFFF) -> abi.State:
        st = abi.State()
        abi.check(abi.lib().natray_sim_run(self._h, n_steps, iters, stop_ns, C.byref(st)))
        return st

    def advance(self, t_target: float, budget: int) -> int:
        taken = C.c_uint32()
        abi.check(abi.lib().natray_sim_advance(self._h, t_target, budget, C.byref(taken)))
        return int(taken.value)

    def step(self) -> abi.State:
        st = abi.State()
        abi.check(abi.lib().natray_sim_step(self._h, C.byref(st)))
        return st

    def phase(self, k: int) -> None:
        abi.check(abi.lib().natray_sim_phase(self._h, k))

    # ---- data -------------------------------------------------------------------------------
    def state(self) -> abi.State:
        st = abi.State()
        abi.check(abi.lib().natray_sim_read_state(self._h, C.byref(st)))
        return st

    def write_state(self, st: abi.State) -> None:
        abi.check(abi.lib().natray_sim_write_state(self._h, C.byref(st)))

    def write_params(self, U: abi.Params) -> None:
        abi.check(abi.lib().natray_sim_write_params(self._h, C.byref(U)))
        self.params = U

    def particles(self) -> np.ndarray:
        P = np.empty((self.n, abi.PARTICLE_FLOATS), dtype=np.float32)
        abi.check(abi.lib().natray_sim_read_particles(self._h, abi.ptr(P)))
        return P

    def write_particles(self, P: np.ndarray) -> None:
        P = np.ascontiguousarray(P, dtype=np.float32)
        assert P.shape == (self.n, abi.PARTICLE_FLOATS)
        abi.check(abi.lib().natray_sim_write_particles(self._h, abi.ptr(P)))

    def n_blocks(self) -> int:
        n = C.c_uint32()
        abi.check(abi.lib().natray_sim_read_grid(self._h, C.byref(n), None, None, None, None, None))
        return int(n.value)

    def grid(self) -> dict:
        """The blocks in use: block_list, block_nb, block_start, acc (19, 64n), cf (11, 64n)."""
        n = self.n_blocks()
        bl = np.empty(max(n, 1), np.uint32)
        nb = np.empty(max(n, 1) * 27, np.uint32)
        bs = np.empty(n + 1, np.uint32)
        acc = np.empty((abi.ACC_FIELDS, max(n, 1) * 64), np.int32)
        cf = np.empty((abi.CELL_FIELDS, max(n, 1) * 64), np.float32)
        nn = C.c_uint32()
        abi.check(abi.lib().natray_sim_read_grid(self._h, C.byref(nn), abi.ptr(bl), abi.ptr(nb), abi.ptr(bs),
                                                  abi.ptr(acc), abi.ptr(cf)))
        return {"n_blocks": n, "block_list": bl[:n], "block_nb": nb[: 27 * n].reshape(n, 27), "block_start": bs,
                "acc": acc[:, : 64 * n], "cf": cf[:, : 64 * n]}

    def write_grid(self, acc: np.ndarray | None, cf: np.ndarray | None) -> None:
        a = None if acc is None else np.ascontiguousarray(acc, dtype=np.int32)
        f = None if cf is None else np.ascontiguousarray(cf, dtype=np.float32)
        abi.check(abi.lib().natray_sim_write_grid(self._h, abi.ptr(a), abi.ptr(f)))

    def bins(self) -> tuple[np.ndarray, np.ndarray]:
        bp = np.empty(self.S, np.uint32)
        so = np.empty(max(self.n, 1), np.uint32)
        abi.check(abi.lib().natray_sim_read_bins(self._h, abi.ptr(bp), abi.ptr(so)))
        return bp, so[: self.n]

    # ---- measurement ------------------------------------------------------------------------
    def set_graph(self, on: bool) -> None:
        """Steps of run() as two replayed HIP graphs around the solve (on, the default) or kernel
        by kernel (off); the results are the same."""
        abi.check(abi.lib().natray_sim_set_graph(self._h, 1 if on else 0))

    def set_profiling(self, on: bool) -> None:
        abi.check(abi.lib().natray_sim_set_profiling(self._h, 1 if on else 0))

    def phase_times(self) -> dict:
        ms = np.zeros(abi.N_PHASES, np.float64)
        cnt = np.zeros(abi.N_PHASES, np.uint32)
        abi.check(abi.lib().natray_sim_phase_times(self._h, abi.ptr(ms), abi.ptr(cnt)))
        return {abi.PHASE_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(abi.N_PHASES)}

    def counters(self) -> dict:
        """Work counters since creation (natray_sim_counters)."""
        out = np.zeros(8, np.uint64)
        abi.check(abi.lib().natray_sim_counters(self._h, abi.ptr(out)))
        keys = ["steps", "cg_iters", "cg_cell_iters", "blocks", "water_cells", "particles", "blocks_cleared"]
        return {k: int(out[i]) for i, k in enumerate(keys)}

    RUN_LOG_KEYS = ("n_steps", "iters", "stop_ns", "taken", "iters_max", "unconverged", "t_ns")

    def run_log(self) -> list[dict]:
        """The natray_sim_run calls made so far (the last 256), oldest first: the batches that
        advance() issued, with their CG budget and stop time (natray_sim_run_log)."""
        cnt = C.c_uint32()
        abi.check(abi.lib().natray_sim_run_log(self._h, 0, None, C.byref(cnt)))
        rows = np.zeros((max(int(cnt.value), 1), 8), np.uint32)
        abi.check(abi.lib().natray_sim_run_log(self._h, int(cnt.value), abi.ptr(rows), C.byref(cnt)))
        return [dict(zip(self.RUN_LOG_KEYS, (int(v) for v in r[:7]))) for r in rows[: int(cnt.value)]]

    def set_solve_launch(self, mode: int) -> None:
        """How the solve is launched (natray_sim_set_solve_launch): abi.SOLVE_PLAIN (the
        default; its waits are bounded, so a workgroup that never becomes resident fails the
        step), abi.SOLVE_COOPERATIVE, or abi.SOLVE_PLAIN_TEST_CHECK (the plain launch with its
        last workgroup staying out, so the solve aborts)."""
        abi.check(abi.lib().natray_sim_set_solve_launch(self._h, int(mode)))

    def water_stats(self) -> dict:
        """The water line of the reference's debug_stats (src/sim.rs:691-716) for the particles
        as they are now (natray_sim_water_stats): alive count, box, mean position and velocity,
        sum of 0.5 |v|^2 per unit mass, the fastest particle, and the cells with >= 4 particles
        with their volume in litres."""
        w = abi.WaterStats()
        abi.check(abi.lib().natray_sim_water_stats(self._h, C.byref(w)))
        return {"alive": int(w.alive), "lo": list(w.lo), "hi": list(w.hi), "mean": list(w.mean),
                "vmean": list(w.vmean), "ke": float(w.ke), "vmax": float(w.vmax), "fastest": list(w.fastest),
                "full_cells": int(w.full_cells), "volume_l": float(w.volume_l)}

    def debug_stats(self) -> str:
        """The reference's debug_stats water line (src/sim.rs:718-740), from water_stats and
        the State: time, dt, step, water, box, means, fastest, volume, blocks and the solver."""
        w, st = self.water_stats(), self.state()
        f = lambda v, d: "[" + ", ".join(f"{x:.{d}f}" for x in v) + "]"
        return (f"t={st.t_ns * 1e-9:.6f} dt={st.dt_ns}ns steps {st.step} | water {w['alive']}/{self.n} "
                f"bbox {f(w['lo'], 3)}..{f(w['hi'], 3)} mean {f(w['mean'], 4)} v {f(w['vmean'], 3)} "
                f"vmax {w['vmax']:.2f} at {f(w['fastest'], 4)} volume {w['volume_l']:.3f} L | blocks {st.n_blocks} "
                f"solver {st.solve_iters} iters r {st.r_norm:.2e} of {st.b_norm:.2e}, {st.unconverged} unconverged "
                f"steps | spray {st.spray_count}")

    def set_membrane(self, x: np.ndarray, v: np.ndarray, mass: np.ndarray, push: np.ndarray) -> None:
        """The membrane's vertices for the coming steps (natray_sim_set_membrane): positions and
        velocities [n][4], masses [n], push (N) and area (m^2) [n][2]."""
        x, v = np.ascontiguousarray(x, np.float32), np.ascontiguousarray(v, np.float32)
        mass, push = np.ascontiguousarray(mass, np.float32), np.ascontiguousarray(push, np.float32)
        n = int(mass.size)
        assert x.shape == (n, 4) and v.shape == (n, 4) and push.shape == (n, 2)
        abi.check(abi.lib().natray_sim_set_membrane(self._h, n, abi.ptr(x), abi.ptr(v), abi.ptr(mass), abi.ptr(push)))
        self._mem_n = n

    def membrane_water(self) -> np.ndarray:
        """The water sampled at each vertex by the last step: [n][4] velocity and fill."""
        out = np.zeros((getattr(self, "_mem_n", 0), 4), np.float32)
        abi.check(abi.lib().natray_sim_read_membrane(self._h, abi.ptr(out)))
        return out

    def air(self) -> np.ndarray:
        """Each cell's gas (natray_sim_read_air): [5][64 n] for the blocks in use."""
        n = self.n_blocks()
        out = np.zeros((5, 64 * n), np.float32)
        abi.check(abi.lib().natray_sim_read_air(self._h, abi.ptr(out)))
        return out

    def spray(self, max_records: int = 1 << 20) -> dict:
        """The spray (natray_sim_read_spray): count = State.spray_count, the first
        min(count, cap, max_records) records [k][12] and where each landed [k][4]."""
        rec = np.zeros((max_records, 12), np.float32)
        landed = np.zeros((max_records, 4), np.float32)
        cnt = C.c_uint32()
        abi.check(abi.lib().natray_sim_read_spray(self._h, max_records, abi.ptr(rec), abi.ptr(landed), C.byref(cnt)))
        n = min(cnt.value, int(self.params.limits[1]), max_records)
        return {"count": cnt.value, "records": rec[:n], "landed": landed[:n]}

    def wet(self) -> np.ndarray:
        """The wet map (natray_sim_read_wet): [limits.x][limits.x] words."""
        res = int(self.params.limits[0])
        out = np.zeros((res, res), np.uint32)
        abi.check(abi.lib().natray_sim_read_wet(self._h, abi.ptr(out)))
        return out

    def pack_keyframe(self, alive: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
        """A whole keyframe (natray_sim_pack_keyframe): the words, and each spray slot's weight
        and time made [n][2]. `alive` holds the membrane triangles' flags."""
        a = None if alive is None else np.ascontiguousarray(alive, np.uint32)
        nt = 0 if a is None else int(a.size)
        total = C.c_uint32()
        abi.check(abi.lib().natray_sim_pack_keyframe(self._h, nt, abi.ptr(a), None, 0, C.byref(total), None))
        words = np.zeros(total.value, np.uint32)
        sw = np.zeros((max(int(self.params.limits[1]), 1), 2), np.float32)
        abi.check(abi.lib().natray_sim_pack_keyframe(self._h, nt, abi.ptr(a), abi.ptr(words), total.value,
                                                     C.byref(total), abi.ptr(sw)))
        return words, sw[: int(words[0])]

    def set_bullets(self, bullets: list, patches: np.ndarray) -> None:
        """Bullets for the coming steps (natray_sim_set_bullets): abi.Bullet records and their
        patches [n][576][8] (position and area, normal and 0, body frame)."""
        n = len(bullets)
        arr = (abi.Bullet * max(n, 1))(*bullets)
        pt = np.ascontiguousarray(patches, np.float32)
        assert pt.shape == (n, abi.BULLET_PATCHES, 8)
        abi.check(abi.lib().natray_sim_set_bullets(self._h, n, C.cast(arr, C.c_void_p), abi.ptr(pt)))
        self._bullets_n = n

    def bullets(self) -> tuple[list, np.ndarray]:
        """The bullets as the last step left them and their cavity sections [n][32][12]."""
        n = getattr(self, "_bullets_n", 0)
        arr = (abi.Bullet * max(n, 1))()
        sec = np.zeros((n, abi.BULLET_SECTIONS, 12), np.float32)
        abi.check(abi.lib().natray_sim_read_bullets(self._h, C.cast(arr, C.c_void_p), abi.ptr(sec)))
        return list(arr)[:n], sec

    def pose_log(self) -> np.ndarray:
        """The bullets' pose log (natray_sim_read_pose_log): [n][256][12]."""
        out = np.zeros((getattr(self, "_bullets_n", 0), abi.BULLET_LOG, 12), np.float32)
        abi.check(abi.lib().natray_sim_read_pose_log(self._h, abi.ptr(out)))
        return out

    def pack_water(self) -> np.ndarray:
        """The water's keyframe words (natray_sim_pack_water): [N][2] uint32."""
        out = np.zeros((self.n, 2), np.uint32)
        abi.check(abi.lib().natray_sim_pack_water(self._h, abi.ptr(out)))
        return out

    def set_warm_start(self, net=None, mode: int = 2) -> None:
        """The learned warm start of the solve (natray_sim_set_warm_start): mode 0 off, 1 inputs
        only (for ws_samples), 2 the network's guess with fallback to the carried pressure, 3
        every solve from p = 0 (the zero guess; b unchanged)."""
        self._ws_net = net if mode == 2 else None  # the network must outlive its use
        h = net.handle if (net is not None and mode == 2) else None
        abi.check(abi.lib().natray_sim_set_warm_start(self._h, h, mode))

    def ws_samples(self, max_tiles: int = 1 << 16) -> dict:
        """Training samples of the last step (natray_sim_ws_samples): inputs [T][4096][2], targets,
        masks and diag [T][4096] FP32, nbm [T][4096] uint32."""
        n = C.c_uint32()
        abi.check(abi.lib().natray_sim_ws_samples(self._h, None, None, None, None, None, max_tiles, C.byref(n)))
        t = int(n.value)
        out = {"inputs": np.empty((t, 4096, 2), np.float32), "targets": np.empty((t, 4096), np.float32),
               "masks": np.empty((t, 4096), np.float32), "diag": np.empty((t, 4096), np.float32),
               "nbm": np.empty((t, 4096), np.uint32)}
        abi.check(abi.lib().natray_sim_ws_samples(self._h, abi.ptr(out["inputs"]), abi.ptr(out["targets"]),
                                                  abi.ptr(out["masks"]), abi.ptr(out["diag"]), abi.ptr(out["nbm"]), t,
                                                  C.byref(n)))
        assert int(n.value) == t
        return out

    def ws_stats(self) -> dict:
        out = np.zeros(8, np.float64)
        abi.check(abi.lib().natray_sim_ws_stats(self._h, abi.ptr(out)))
        return {"offered": int(out[0]), "accepted": int(out[1]), "guess_rr": float(out[2]), "carried_rr": float(out[3]),
                "tiles": int(out[4])}

    def ws_guess(self):
        """(guess over bdims.w * 64 cells, whether the last solve started from it)."""
        g = np.empty(int(self.params.bdims[3]) * 64, np.float32)
        acc = C.c_uint32()
        abi.check(abi.lib().natray_sim_ws_guess(self._h, abi.ptr(g), C.byref(acc)))
        return g, bool(acc.value)

    def stream_ptr(self) -> int:
        return int(abi.lib().natray_sim_device_ptr(self._h, abi.BUF_STREAM) or 0)
