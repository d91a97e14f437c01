#!/usr/bin/env python3
"""Throughput of the natray water substep on one MI355X.

    python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--workload impact|sphere]
                    [--precision fp32|mixed] [--no-cpu]

A step is one substep of the APIC water (binning, P2G, grid update, pressure CG, projection,
velocity extension, G2P) over the whole scene, from particles already resident in HBM. The
default workload is the balloon impact: 2,097,152 particles (8 per cell) in a water sphere on a
256^3 sparse grid (dx = 1/256) with solid walls, a rigid capsule (radius 0.01 m) entering at
300 m/s as a moving solid, dt from CFL 0.5, block-SGS preconditioned CG to a relative residual
of 1e-6, 200 steps. Data are synthetic (seeded); there are no weights.

Prints one JSON line: value = particle-substeps/s over the whole timed region (CG included);
`roofline` for the dominant kernel from per-kernel HIP events on the library's stream and the
kernel's algorithmic bytes; `cpu_baseline` = the FP32 C restatement (oracle/) on the host's
cores over a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

# Algorithmic bytes (unique bytes of the reference data layout) per alive particle and per
# allocated cell for each phase, and per water cell per CG iteration.
PER_PARTICLE = {"bin": 44, "p2g": 68, "g2p": 104}
PER_CELL = {"bin": 76, "p2g": 36, "grid_prep": 64, "solve": 36, "project": 32, "extend": 76, "g2p": 32}
PER_CG_CELL_ITER = 60


def note(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def phase_bytes(c: dict) -> dict:
    """Algorithmic bytes of each phase summed over the steps in counter deltas `c`."""
    N, C, C_prev, WI = c["particles"], 64 * c["blocks"], 64 * c["blocks_cleared"], c["cg_cell_iters"]
    out = {}
    for ph in ("bin", "p2g", "grid_prep", "solve", "project", "extend", "g2p"):
        b = PER_PARTICLE.get(ph, 0) * N
        if ph == "bin":
            b += 76 * C_prev  # the clear of last step's cells
        else:
            b += PER_CELL.get(ph, 0) * C
        if ph == "solve":
            b += PER_CG_CELL_ITER * WI
        out[ph] = b
    return out


def make_scene(args):
    from natray_amd import abi, scenes

    if args.workload == "impact":
        prec = {"fp32": abi.CG_FP32, "mixed": abi.CG_MIXED, "pipelined": abi.CG_PIPELINED}[args.precision]
        s = scenes.balloon_impact(precision=prec, iters=args.iters, refresh_every=10 if args.precision == "pipelined" else 20)
        cfg = {
            "workload": "balloon impact: 2,097,152 particles, 256^3 sparse grid of 4^3 blocks, dx=1/256, "
                        "capsule r=0.01 at 300 m/s as moving solid, CFL 0.5, block-SGS CG tol 1e-6",
            "n_particles": s.n, "grid": "256^3", "precond": "block-SGS", "cg_precision": args.precision,
            "tol": 1e-6, "cg_iter_cap": args.iters,
        }
    else:
        s = scenes.water_sphere()
        cfg = {"workload": "water sphere: 32,768 particles, 64^3, dt=1e-4, Jacobi CG tol 1e-6 (<=200 iters)",
               "n_particles": s.n, "grid": "64^3", "precond": "Jacobi", "cg_precision": "fp32", "tol": 1e-6,
               "cg_iter_cap": 200}
        args.iters = 200
    return s, cfg


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of `kernel` (fetch + write) from the rocprofv3 counter summary of
    the balloon-impact workload, or None when the summary has no figure for it."""
    path = os.path.join(ROOT, "profiles", "impact_2M_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 4 and parts[0] == "traffic" and parts[1] == kernel:
            kv = dict(p.split("=") for p in parts[2:])
            return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
    return None


# The kernels of each phase, as the counter summary names them (single scene, FP32 CG).
PHASE_KERNELS = {
    "bin": ["k_clear<One>", "k_mark<One,4u>", "k_scan_reduce<One>", "k_scan_tiles<One>", "k_scan_scatter<One>",
            "k_sort<One,4u>"],
    "p2g": ["k_p2g<One>"],
    "grid_prep": ["k_grid_prep<One>"],
    "solve": ["k_solve<false,One>"],
    "project": ["k_project<One>"],
    "extend": ["k_extend1<One>", "k_extend2<One>"],
    "g2p": ["k_g2p<One,false>", "k_vmax<One>"],
}


def measured_phase_traffic() -> dict | None:
    """Counter-measured HBM bytes per step of each phase (its kernels' fetch + write per launch,
    one launch each per step), or None when the summary lacks a kernel."""
    out = {}
    for ph, ks in PHASE_KERNELS.items():
        t = [measured_traffic(k) for k in ks]
        if any(x is None for x in t):
            return None
        out[ph] = float(sum(t))
    return out


def timed_run(sim, warmup: int, steps: int, iters: int, graph: bool = True, events: bool = True):
    """Warmup steps, then `steps` timed steps; wall time, counter deltas, phase times and the
    final State. With `events`, HIP events on the library's stream time the phases: with
    `graph` (the library's default) the steps replay two captured graphs around the solve and
    only the solve is timed; without it every kernel is launched and timed on its own. Without
    `events` (and with `graph`), each step replays one graph that holds the solve too, as a
    caller's steps do."""
    sim.set_graph(graph)
    if warmup:
        sim.run(warmup, iters=iters)
    sim.set_profiling(events)
    c0 = sim.counters()
    t0 = time.perf_counter()
    st = sim.run(steps, iters=iters)  # returns after the stream is idle
    t1 = time.perf_counter()
    c1 = sim.counters()
    times = sim.phase_times()
    sim.set_profiling(False)
    return t1 - t0, {k: c1[k] - c0[k] for k in c0}, times, st


def bullet_workload(args):
    """The 2M impact without the capsule solid, and one bullet (4.2 g, 9 mm across, 20 mm long)
    at 300 m/s along +x with its nose two cells short of the water: (scene, bullet, patches)."""
    from natray_amd import abi, scenes

    s = scenes.balloon_impact(iters=args.iters)
    s.scene.flags &= ~abi.SCENE_CAPSULE
    nose = 0.5 - float(s.meta["r_eff"]) - 2.0 * float(s.meta["dx"])
    b, patches = scenes.bullet_shot(at=(nose - 0.011, 0.5, 0.5), speed=300.0)
    s.meta["workload"] = f"{s.meta['workload']}, one bullet at 300 m/s instead of the capsule"
    return s, b, patches


def bullet_report(args) -> dict:
    """The reference's bullet model in place of the capsule solid: the 2M impact without the
    capsule, with one bullet (4.2 g, 9 mm across, 20 mm long, 300 m/s along +x, its nose two
    cells short of the water), kernel by kernel over `--bullet-steps` steps, against the same
    scene without the bullet. The grid phase holds grid_prep, bullet_step and bullet_apply, so
    the difference of the two grid phases is what the bullet costs a step."""
    from natray_amd import Sim

    s, b, patches = bullet_workload(args)
    out = {"workload": s.meta["workload"], "steps": args.bullet_steps}
    for label, with_bullet in (("with_bullet", True), ("without", False)):
        sim = Sim.from_scene(s)
        if with_bullet:
            sim.set_bullets([b], patches[None])
        wall, c, times, st = timed_run(sim, args.warmup, args.bullet_steps, args.iters, graph=False)
        steps = max(c["steps"], 1)
        out[label] = {"ms_per_step_kernel_by_kernel": wall * 1e3 / max(args.bullet_steps, 1),
                      "grid_phase_us_per_step": times["grid_prep"][0] * 1e3 / steps,
                      "cg_iters_per_step": c["cg_iters"] / steps, "dt_ns_last": st.dt_ns}
        if with_bullet:
            (B,), sec = sim.bullets()
            out[label]["bullet"] = {"x": list(B.x[:3]), "speed": float(np.linalg.norm(list(B.v[:3]))),
                                    "state": int(B.info[2]), "sections_made": int(B.info[3]),
                                    "reach": float(B.reach[3])}
        sim.close()
    out["bullet_us_per_step"] = out["with_bullet"]["grid_phase_us_per_step"] - out["without"]["grid_phase_us_per_step"]
    return out


def compare_precision(scene, args, plain: dict) -> dict:
    """The same workload from the same initial state with the other CG forms, at the same
    tolerance in the reference's convergence form: mixed precision (FP32 SpMV and axpy, FP64
    dots, norms and convergence test, FP64 true residual every 20 iterations) and the pipelined
    CG with that precision (one reduction per iteration, residual replacement every 10
    iterations while the residual is large). Iterations and solve time per step against the
    plain FP32 run."""
    from natray_amd import Sim, abi

    sc = scene.scene
    keep = (sc.precision, sc.refresh_every)
    out = {"tol": scene.params.solve[0], "plain_fp32": plain}
    try:
        for label, prec, refresh, zero in (("plain_fp32_from_zero", abi.CG_FP32, 0, True),
                                           ("mixed_f32_f64", abi.CG_MIXED, 20, False),
                                           ("pipelined_f32_f64", abi.CG_PIPELINED, 10, False)):
            sc.precision, sc.refresh_every = prec, refresh
            sim = Sim.from_scene(scene)
            if zero:
                sim.set_warm_start(None, 3)
            wall, c, times, st = timed_run(sim, args.warmup, args.steps, args.iters)
            sim.close()
            steps = max(c["steps"], 1)
            out[label] = {"cg_iters_per_step": c["cg_iters"] / steps, "solve_ms_per_step": times["solve"][0] / steps,
                          "ms_per_step": wall * 1e3 / max(args.steps, 1), "unconverged_steps": st.unconverged,
                          "refresh_every": refresh, "p0": "0" if zero else "particle-carried"}
    finally:
        sc.precision, sc.refresh_every = keep
    for label, key in (("mixed_f32_f64", "solve_speedup_mixed_over_plain"),
                       ("pipelined_f32_f64", "solve_speedup_pipelined_over_plain")):
        m = out[label]["solve_ms_per_step"]
        out[key] = plain["solve_ms_per_step"] / m if m > 0 else None
        out[key + "_from_zero"] = out["plain_fp32_from_zero"]["solve_ms_per_step"] / m if m > 0 else None
    out["warm_start"] = ("particle-carried pressure (the reference's p0) in every run but plain_fp32_from_zero, "
                         "which starts each solve from p = 0 with the same b")
    return out


def warm_start_report(scene, args) -> dict:
    """The learned warm start on this workload: training tiles logged from seeded runs of the
    same configuration (other seeds), the network trained on the GPU by its own kernels on the
    residual its guess leaves, then the timed steps again from the same initial state with the
    network's guess as the solve's start (plain CG to the same tolerance, the carried pressure
    when the guess does not lower the residual), against the run without it."""
    from natray_amd import Sim, abi, scenes
    from natray_amd import net as netmod
    from natray_amd import warm_start as ws

    t0 = time.perf_counter()
    parts, seeds = [], []
    tiles = 0
    for seed in range(11, 11 + 8):
        if tiles >= args.ws_tiles:
            break
        sc = scenes.balloon_impact(seed=seed, iters=args.iters, precision=scene.scene.precision)
        sim = Sim.from_scene(sc)
        d = ws.collect(sim, args.ws_collect_steps, iters=args.iters)
        sim.close()
        parts.append(d)
        seeds.append(seed)
        tiles += d["inputs"].shape[0]
    data = {k: np.concatenate([p[k] for p in parts]) for k in ws.KEYS}
    del parts
    t_collect = time.perf_counter() - t0
    T = data["inputs"].shape[0]
    rng = np.random.default_rng(0)
    perm = rng.permutation(T)
    nval = max(T // 20, 1)
    val = ws.subset(data, np.sort(perm[:nval]))
    train = ws.subset(data, np.sort(perm[nval:]))
    del data
    net = netmod.Net(max_tiles=1024, seed=1)
    t0 = time.perf_counter()
    losses = ws.train(net, train, args.ws_train_steps, batch=32, lr=2e-3)
    t_train = time.perf_counter() - t0
    del train
    val_ratio = ws.residual_ratio(net, val, 64)
    del val

    def run(with_net: bool, pipelined: bool = False, zero: bool = False) -> dict:
        sc = scene.scene
        keep = (sc.precision, sc.refresh_every)
        if pipelined:
            sc.precision, sc.refresh_every = abi.CG_PIPELINED, 10
        try:
            sim = Sim.from_scene(scene)
        finally:
            sc.precision, sc.refresh_every = keep
        if with_net:
            sim.set_warm_start(net, 2)
        elif zero:
            sim.set_warm_start(None, 3)
        wall, c, times, st = timed_run(sim, args.warmup, args.ws_steps, args.iters)
        stats = sim.ws_stats() if with_net else None
        sim.close()
        steps = max(c["steps"], 1)
        r = {"cg_iters_per_step": c["cg_iters"] / steps, "ms_per_step": wall * 1e3 / max(args.ws_steps, 1),
             "solve_ms_per_step": times["solve"][0] / steps, "unconverged_steps": st.unconverged}
        if stats:
            r["guesses_taken"] = stats["accepted"]
            r["guesses_offered"] = stats["offered"]
            r["tiles_last_step"] = stats["tiles"]
            r["guess_over_carried_residual_sq_last_step"] = (stats["guess_rr"] / stats["carried_rr"]
                                                           if stats["carried_rr"] > 0 else None)
        return r

    plain, learned, combined, zero = run(False), run(True), run(True, pipelined=True), run(False, zero=True)
    return {
        "training": {"tiles": T, "seeds": seeds, "steps_per_seed": args.ws_collect_steps,
                     "collect_s": t_collect, "adam_steps": args.ws_train_steps, "batch_tiles": 32,
                     "train_s": t_train, "final_loss": float(np.mean(losses[-100:])),
                     "val_tile_residual_ratio": val_ratio},
        "plain": plain,
        "warm_started": learned,
        "cg_iters_saved_per_step": plain["cg_iters_per_step"] - learned["cg_iters_per_step"],
        "substep_time_ratio_warm_over_plain": learned["ms_per_step"] / plain["ms_per_step"],
        "solve_time_ratio_warm_over_plain": learned["solve_ms_per_step"] / plain["solve_ms_per_step"],
        # Both added techniques together: the pipelined CG at mixed precision, warm-started.
        "warm_started_pipelined_mixed": combined,
        "solve_speedup_pipelined_mixed_warm_over_plain": plain["solve_ms_per_step"] / combined["solve_ms_per_step"],
        # Plain FP32 CG from p = 0 (config 5's fallback guess), the same b and tolerance.
        "plain_from_zero": zero,
        "solve_speedup_warm_over_plain_from_zero": zero["solve_ms_per_step"] / learned["solve_ms_per_step"],
        "solve_speedup_pipelined_mixed_warm_over_plain_from_zero": zero["solve_ms_per_step"] / combined["solve_ms_per_step"],
        "note": "solve time includes the warm start (inputs, network, guess check); tol and b are the same",
    }


def stated_radius_report(args) -> dict:
    """The workload's stated geometry: the radius-0.3 sphere filled at 8 particles per cell on
    256^3 (15.2M particles) instead of the stated 2,097,152 particles, with the same capsule,
    CFL and CG. Timed like the main run (graph steps, particles resident)."""
    from natray_amd import Sim, scenes

    s = scenes.stated_radius_impact(iters=args.iters)
    sim = Sim.from_scene(s)
    wall, c, times, st = timed_run(sim, args.warmup, args.stated_steps, args.iters)
    sim.close()
    steps = max(c["steps"], 1)
    t_cg = times["solve"][0] * 1e-3
    return {"workload": s.meta["workload"], "n_particles": s.n, "steps": args.stated_steps,
            "particle_substeps_per_s": c["particles"] / wall if wall > 0 else 0.0,
            "ms_per_step": wall * 1e3 / max(args.stated_steps, 1),
            "cg_iters_per_step": c["cg_iters"] / steps, "blocks_per_step": c["blocks"] / steps,
            "cg_cell_iters_per_s": c["cg_cell_iters"] / t_cg if t_cg > 0 else 0.0,
            "solve_ms_per_step": times["solve"][0] / steps, "unconverged_steps": st.unconverged}


def batched_report(args) -> dict:
    """Config 4: 64 independent balloon impacts (16k..512k particles, 64^3..192^3, capsule
    200..400 m/s in a random direction, mixed-precision CG) stepped as one batch: the
    autotuner picks the launch shape, then timed batch steps; against the same scenes stepped
    one after another, each on its own (its own launches and graphs)."""
    from natray_amd import Sim, scenes
    from natray_amd.batch import Batch

    t0 = time.perf_counter()
    sl = scenes.batch_impacts()
    sims = [Sim.from_scene(s) for s in sl]
    t_setup = time.perf_counter() - t0
    b = Batch(sims)
    tuned = b.tune(args.batch_tune_steps, iters=args.iters)

    def totals():
        c = [s.counters() for s in sims]
        return {k: sum(x[k] for x in c) for k in c[0]}

    b.run(args.warmup, iters=args.iters)
    c0 = totals()
    t0 = time.perf_counter()
    b.run(args.batch_steps, iters=args.iters)
    wall = time.perf_counter() - t0
    c1 = totals()
    # The per-phase breakdown: kernel-by-kernel batch steps with events around each phase.
    b.set_graph(False)
    b.set_profiling(True)
    b.run(args.batch_steps, iters=args.iters)
    times = b.phase_times()
    b.set_profiling(False)
    b.set_graph(True)
    c1p = totals()
    # The same scenes one after another, each stepped alone for the same number of steps.
    t0 = time.perf_counter()
    for s in sims:
        s.run(args.batch_steps, iters=args.iters)
    wall_alone = time.perf_counter() - t0
    c2 = totals()
    st = [s.state() for s in sims]
    b.close()
    for s in sims:
        s.close()
    parts = c1["particles"] - c0["particles"]
    return {
        "workload": "64 balloon impacts, particles log-uniform in [16k, 512k], grids 64^3..192^3, capsule "
                    "200..400 m/s in a random direction, mixed-precision CG (FP64 dots, refresh every 20), tol 1e-6",
        "scenes": len(sl), "particles": int(sum(x.n for x in sl)), "setup_s": t_setup,
        "autotuner": tuned,
        "steps": args.batch_steps,
        "ms_per_batch_step": wall * 1e3 / max(args.batch_steps, 1),
        "particle_substeps_per_s": parts / wall if wall > 0 else 0.0,
        "cg_iters_per_scene_step": (c1["cg_iters"] - c0["cg_iters"]) / max(c1["steps"] - c0["steps"], 1),
        # Blocks per batch step (all scenes) and particles per block: the batch's small scenes
        # hold more surface blocks per particle than one large scene.
        "blocks_per_batch_step": (c1["blocks"] - c0["blocks"]) / max(args.batch_steps, 1),
        "particles_per_block": parts / max(c1["blocks"] - c0["blocks"], 1),
        "phases_ms_per_batch_step_kernel_by_kernel": {k: v[0] / max(v[1], 1) for k, v in times.items()},
        "scenes_one_after_another": {
            "ms_per_step_of_all_scenes": wall_alone * 1e3 / max(args.batch_steps, 1),
            "particle_substeps_per_s": (c2["particles"] - c1p["particles"]) / wall_alone if wall_alone > 0 else 0.0,
        },
        "speedup_batch_over_one_after_another": wall_alone / wall if wall > 0 else None,
        "unconverged_steps": int(sum(x.unconverged for x in st)),
    }


def _cells_by_slot(g: dict) -> dict:
    return {int(slot): i for i, slot in enumerate(g["block_list"])}


def _divergence(g: dict, U: np.ndarray) -> np.ndarray:
    """Discrete divergence sum over dims of u(up face) - u(own face), per cell (in units of
    velocity), from the face velocities U (3 x 64n) of the blocks in `g`; a face of a missing
    neighbour counts as 0."""
    n = g["n_blocks"]
    lane = np.arange(64)
    loc = np.stack([lane >> 4, (lane >> 2) & 3, lane & 3])
    div = np.zeros(64 * n)
    nb = g["block_nb"].astype(np.int64)
    for dim in range(3):
        up_local = loc.copy()
        up_local[dim] += 1
        inside = up_local[dim] < 4
        up_local[dim] %= 4
        up_lane = up_local[0] * 16 + up_local[1] * 4 + up_local[2]
        link = 13 + (9, 3, 1)[dim]  # the +dim neighbour in the 27-block table
        blk = np.repeat(np.arange(n), 64)
        lanes = np.tile(np.arange(64), n)
        other = nb[blk, link]
        up_blk = np.where(inside[lanes], blk, other)
        ok = up_blk != 0xFFFFFFFF
        up_cell = np.where(ok, up_blk * 64 + up_lane[lanes], 0)
        u_up = np.where(ok, U[dim][up_cell], 0.0)
        div += u_up - U[dim]
    return div


def _errors(g: dict, P: np.ndarray, g_ref: dict, cf_ref: np.ndarray, P_ref: np.ndarray, h: float) -> dict:
    """Errors of a GPU grid and particle set against the FP64 oracle's, on the cells of blocks
    both allocated (matched by domain slot) and the particles alive in both."""
    from natray_amd import abi

    slots_ref = _cells_by_slot(g_ref)
    common = [(i, slots_ref[int(slot)]) for i, slot in enumerate(g["block_list"]) if int(slot) in slots_ref]
    gi = np.concatenate([np.arange(64) + 64 * i for i, _ in common])
    ri = np.concatenate([np.arange(64) + 64 * j for _, j in common])
    water = (g["cf"][abi.F_TYPE][gi] == abi.WATER) & (cf_ref[abi.F_TYPE][ri] == abi.WATER)
    p_g, p_r = g["cf"][abi.F_P][gi][water].astype(np.float64), cf_ref[abi.F_P][ri][water]
    u_g = g["cf"][abi.F_U:abi.F_U + 3][:, gi].astype(np.float64)
    u_r = cf_ref[abi.F_U:abi.F_U + 3][:, ri]
    div_g = _divergence(g, g["cf"][abi.F_U:abi.F_U + 3].astype(np.float64))[gi][water]
    div_r = _divergence(g_ref, cf_ref[abi.F_U:abi.F_U + 3])[ri][water]
    alive = (P[:, 7] == 1.0) & (P_ref[:, 7] == 1.0)
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30)) if b.size else 0.0
    return {
        "max_pressure_error_rel": rel(p_g, p_r),
        "max_grid_velocity_error_rel": rel(u_g, u_r),
        "max_divergence_error_rel": rel(div_g, div_r),
        "max_position_error_cells": float(np.abs(P[alive, 0:3].astype(np.float64) - P_ref[alive, 0:3]).max() / h)
        if alive.any() else 0.0,
        "alive_mismatch": int(((P[:, 7] == 1.0) != (P_ref[:, 7] == 1.0)).sum()),
        "blocks_not_shared": int(g["n_blocks"] - len(common)),
    }


def accuracy_vs_fp64(scene, checkpoints, iters: int, one_from: int = 20) -> dict:
    """Config 3's report: the plain FP32, the mixed-precision and the pipelined CG run from the scene's
    initial state on the GPU, the FP64 oracle on the host; errors of the GPU runs against it
    after each checkpoint's number of steps, at the same tolerance. Relative errors are the
    largest difference over the largest reference magnitude."""
    import oracle
    from natray_amd import Sim, abi

    oracle.set_threads(min(16, os.cpu_count() or 1))
    h = float(scene.params.origin[3])
    ref = oracle.Oracle.from_scene(scene, "f64")
    refs, done = {}, 0
    for k in checkpoints:
        ref.run(k - done, iters)
        done = k
        sc_ref = ref.solver_scalars()
        refs[k] = (ref.grid(), ref.cf_f64(), ref.particles_f64(),
                   sc_ref["r_norm"] / max(sc_ref["b_norm"], 1.0), int(sc_ref["solve_iters"]))
    ref.close()
    out = {"tol": float(scene.params.solve[0]),
           "criteria": "grid velocity <= 1e-5 relative and positions <= 1e-4 dx after one substep"}
    sc = scene.scene
    keep = (sc.precision, sc.refresh_every)
    try:
        for label, prec, refresh in (("plain_fp32", abi.CG_FP32, 0), ("mixed_f32_f64", abi.CG_MIXED, 20),
                                     ("pipelined_f32_f64", abi.CG_PIPELINED, 10)):
            sc.precision, sc.refresh_every = prec, refresh
            sim = Sim.from_scene(scene)
            done = 0
            for k in checkpoints:
                c0 = sim.counters()
                st = sim.run(k - done, iters=iters)
                c1 = sim.counters()
                done = k
                g_ref, cf_ref, P_ref, res_ref, it_ref = refs[k]
                e = _errors(sim.grid(), sim.particles(), g_ref, cf_ref, P_ref, h)
                e["cg_iters"] = int(c1["cg_iters"] - c0["cg_iters"])
                e["cg_relative_residual_last_solve"] = st.r_norm / max(st.b_norm, 1.0)
                e["fp64_cg_relative_residual_last_solve"] = res_ref
                e["fp64_cg_iters_last_step"] = it_ref
                out.setdefault(label, {})[f"after_{k}_steps"] = e
            # One substep from identical state: the GPU's particles and State after `one_from`
            # steps are loaded into a fresh FP64 oracle, then both take one step.
            st = sim.run(one_from - done, iters=iters)
            ref = oracle.Oracle(scene.params, sim.particles(), scene.scene, sim.state(), "f64")
            ref.run(1, iters)
            sc_ref = ref.solver_scalars()
            g_ref, cf_ref, P_ref = ref.grid(), ref.cf_f64(), ref.particles_f64()
            ref.close()
            c0 = sim.counters()
            st = sim.run(1, iters=iters)
            c1 = sim.counters()
            e = _errors(sim.grid(), sim.particles(), g_ref, cf_ref, P_ref, h)
            e["cg_iters"] = int(c1["cg_iters"] - c0["cg_iters"])
            e["fp64_cg_iters"] = int(sc_ref["solve_iters"])
            e["cg_relative_residual"] = st.r_norm / max(st.b_norm, 1.0)
            e["within_criteria"] = bool(e["max_grid_velocity_error_rel"] <= 1e-5 and e["max_position_error_cells"] <= 1e-4)
            out[label][f"one_substep_from_identical_state_at_step_{one_from}"] = e
            sim.close()
    finally:
        sc.precision, sc.refresh_every = keep
    return out


def host_cpu() -> dict:
    """The host's CPU model and counts, from /proc/cpuinfo and os.cpu_count()."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def cpu_threads() -> int:
    """The host cores this job is given: OMP_NUM_THREADS where the environment sets it (the
    GPU box allots a share of its host to each GPU and sets it), else every logical CPU."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else (os.cpu_count() or 1)


def cpu_baseline(scene, iters: int, budget_s: float, runs: int = 5, one_thread_steps: int = 2) -> dict:
    """The FP32 C restatement of the reference kernels (oracle/; the reference itself has no
    CPU path and its WGSL-on-lavapipe route cannot run here) on the same workload from the same
    initial state: one untimed step, then `runs` timed runs of consecutive steps on all the
    job's host cores, each until budget_s / runs of wall time; the value is the median run.
    Then `one_thread_steps` steps on one thread."""
    import oracle

    threads = cpu_threads()
    oracle.set_threads(threads)
    o = oracle.Oracle.from_scene(scene, "f32")
    o.run(1, iters)
    rates, steps_all, t_all = [], 0, 0.0
    for _ in range(runs):
        steps, t0 = 0, time.perf_counter()
        while steps < 1 or time.perf_counter() - t0 < budget_s / runs:
            o.run(1, iters)
            steps += 1
        dt = time.perf_counter() - t0
        rates.append(scene.n * steps / dt)
        steps_all += steps
        t_all += dt
    oracle.set_threads(1)
    t0 = time.perf_counter()
    o.run(one_thread_steps, iters)
    dt1 = time.perf_counter() - t0
    oracle.set_threads(threads)
    o.close()
    rates.sort()
    return {"value": float(np.median(rates)), "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "label": "CPU restatement, not the lavapipe reference",
            "runs": [float(r) for r in rates], "one_thread_value": scene.n * one_thread_steps / dt1,
            "host": host_cpu(),
            "sample": f"median of {runs} runs ({steps_all} steps in all, {t_all:.1f} s) of the same {scene.n}-particle "
                      f"workload after one untimed step, FP32 C restatement of the reference kernels (oracle/), "
                      f"OpenMP on {threads} threads; then {one_thread_steps} steps on 1 thread ({dt1:.1f} s)"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["impact", "sphere"], default="impact")
    ap.add_argument("--precision", choices=["fp32", "mixed", "pipelined"], default="fp32")
    ap.add_argument("--iters", type=int, default=1000, help="CG iteration cap per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-compare", action="store_true", help="skip the mixed-precision and pipelined comparison runs")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the error report against the FP64 oracle")
    ap.add_argument("--accuracy-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="wall time of the CPU baseline's timed runs")
    ap.add_argument("--no-warm-start", action="store_true", help="skip the learned warm-start report")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes report")
    ap.add_argument("--no-stated", action="store_true", help="skip the stated-radius (15.2M particles) report")
    ap.add_argument("--stated-steps", type=int, default=40, help="timed steps of the stated-radius workload")
    ap.add_argument("--batch-steps", type=int, default=30, help="timed steps of the 64-scene batch")
    ap.add_argument("--batch-tune-steps", type=int, default=4, help="autotuner rounds (one timed step per candidate each)")
    ap.add_argument("--ws-tiles", type=int, default=50000, help="training tiles to log for the warm start")
    ap.add_argument("--ws-collect-steps", type=int, default=100, help="steps logged per training seed")
    ap.add_argument("--ws-train-steps", type=int, default=6000)
    ap.add_argument("--ws-steps", type=int, default=100, help="timed steps with and without the warm start")
    ap.add_argument("--no-bullets", action="store_true", help="skip the bullet report")
    ap.add_argument("--bullet-steps", type=int, default=40, help="steps of the bullet report")
    ap.add_argument("--bullets-only", action="store_true", help="print the bullet report alone")
    ap.add_argument("--write-scene", metavar="PATH",
                    help="write the workload's scene file for natray_amd/natray_drive and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: the library has no multi-GPU path (asked for --gpus {args.gpus}); "
              "the scene would shard by whole scenes or by Morton-ordered block ranges with halo exchange, "
              "which is not built.", file=sys.stderr)
        return 2

    from natray_amd import Sim, scenes

    if args.bullets_only:
        if args.write_scene:
            s, b, patches = bullet_workload(args)
            scenes.write_scene_file(s, args.write_scene, bullets=([b], patches[None]))
            print(f"wrote {args.write_scene}: {s.meta['workload']}")
            return 0
        print(json.dumps({"bullets": bullet_report(args)}), flush=True)
        return 0
    scene, cfg = make_scene(args)
    if args.write_scene:
        scenes.write_scene_file(scene, args.write_scene)
        print(f"wrote {args.write_scene}: {cfg['workload']}")
        return 0
    note(f"main run: {cfg['workload']}, {args.warmup} warmup + {args.steps} timed steps")
    sim = Sim.from_scene(scene)
    wall, c, _, st = timed_run(sim, args.warmup, args.steps, args.iters, events=False)
    sim.close()
    note(f"main run: {wall * 1e3 / max(args.steps, 1):.3f} ms per step")
    # The solve's launches timed with events: the same steps from the same state, with the
    # solve launched between two graphs so that events can bracket it.
    sim = Sim.from_scene(scene)
    wall_ev, _, times, _ = timed_run(sim, args.warmup, args.steps, args.iters)
    sim.close()
    # The per-phase breakdown: the same steps from the same state, every kernel launched and
    # timed on its own (the timed run above replays graphs and times the solve only).
    sim = Sim.from_scene(scene)
    wall_eager, _, times_eager, _ = timed_run(sim, args.warmup, args.steps, args.iters, graph=False)
    sim.close()
    steps_done = c["steps"]
    particles = c["particles"]
    value = particles / wall if wall > 0 else 0.0

    pb = phase_bytes(c)
    ptr = measured_phase_traffic() if args.workload == "impact" and args.precision == "fp32" else None
    single = ("p2g", "grid_prep", "solve", "project", "g2p")  # phases that are one kernel
    dominant = max(single, key=lambda k: times_eager[k][0])
    # The dominant kernel's launches are timed in the graph run when it is the solve (launched
    # between the graphs), else in the kernel-by-kernel run.
    ms, launches = times[dominant] if times[dominant][1] > 0 else times_eager[dominant]
    avg_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = pb[dominant] / max(launches, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    kernel_names = {"p2g": "k_p2g", "grid_prep": "k_grid_prep", "solve": "k_solve", "project": "k_project",
                    "g2p": "k_g2p"}
    t_cg = times["solve"][0] * 1e-3
    t_all_kernels = sum(v[0] for v in times_eager.values()) * 1e-3
    out = {
        "metric": "particle-substeps/s + CG cell-iters/s, mixed/warm vs plain FP32 at same div tol",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f32+f64",
        "data": "synthetic (seeded particle fill; no checkpoint)",
        "config": cfg,
        "cg_cell_iters_per_s": c["cg_cell_iters"] / t_cg if t_cg > 0 else 0.0,
        "particle_substeps_per_s_outside_cg": particles / max(wall - t_cg, 1e-12),
        "cg_iters_per_step": c["cg_iters"] / max(steps_done, 1),
        "blocks_per_step": c["blocks"] / max(steps_done, 1),
        "particles_per_block": particles / max(c["blocks"], 1),
        "water_cells_per_step": c["water_cells"] / max(steps_done, 1),
        "final_state": {"t_ns": st.t_ns, "step": st.step, "dt_ns": st.dt_ns, "unconverged": st.unconverged,
                        "max_speed": st.max_speed},
        "ms_per_step_kernel_by_kernel": wall_eager * 1e3 / max(args.steps, 1),
        # The same steps with events around the solve (two graphs around it, not one).
        "ms_per_step_with_solve_events": wall_ev * 1e3 / max(args.steps, 1),
        "phases_ms_per_step": {k: v[0] / max(steps_done, 1) for k, v in times_eager.items()},
        "phases_algorithmic_GBps": {k: (pb[k] / (times_eager[k][0] * 1e-3) / 1e9 if times_eager[k][0] > 0 else 0.0)
                                    for k in pb},
        "kernel_time_fraction_of_wall_kernel_by_kernel": t_all_kernels / wall_eager if wall_eager > 0 else 0.0,
        # Counter-side figures per phase (the profile's HBM bytes per launch, this run's phase
        # times), and the counter bytes over the algorithmic ones: above 1 means re-reads.
        "phases_counter_GBps": ({k: v / (times_eager[k][0] * 1e-3 / max(steps_done, 1)) / 1e9
                                 for k, v in ptr.items() if times_eager[k][0] > 0} if ptr else None),
        "phases_counter_over_algorithmic": ({k: v * max(steps_done, 1) / pb[k] for k, v in ptr.items() if pb[k] > 0}
                                            if ptr else None),
        "roofline": {
            "kernel": kernel_names[dominant],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # Counter-measured HBM bytes per launch of this kernel on this workload (FP32 form);
            # None where no profile covers it.
            "traffic": (measured_traffic("k_solve<false,One>" if dominant == "solve" else kernel_names[dominant] + "<One>")
                        if args.workload == "impact" and args.precision == "fp32" else None),
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if args.workload == "impact" and args.precision == "fp32" and not args.no_compare:
        note("CG forms: plain from p = 0, mixed, pipelined")
        out["precision_compare"] = compare_precision(
            scene, args, {"cg_iters_per_step": out["cg_iters_per_step"],
                          "solve_ms_per_step": times["solve"][0] / max(steps_done, 1),
                          "ms_per_step": wall_ev * 1e3 / max(args.steps, 1), "unconverged_steps": st.unconverged})
    if args.workload == "impact" and not args.no_stated:
        note("stated radius: 15.2M particles")
        out["stated_radius"] = stated_radius_report(args)
    if args.workload == "impact" and not args.no_batch:
        note("batch of 64 scenes")
        out["batched_scenes"] = batched_report(args)
    if args.workload == "impact" and not args.no_warm_start:
        note("learned warm start: logging, training, timed runs")
        out["learned_warm_start"] = warm_start_report(scene, args)
    if args.workload == "impact" and not args.no_bullets:
        note("bullets")
        out["bullets"] = bullet_report(args)
    if args.workload == "impact" and not args.no_accuracy:
        note("accuracy against the FP64 oracle")
        out["accuracy_vs_fp64_oracle"] = accuracy_vs_fp64(scene, (1, args.accuracy_steps), args.iters)
    if not args.no_cpu:
        note("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(scene, args.iters, args.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
