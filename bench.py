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

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

# Algorithmic bytes (unique bytes of the reference data layout) per alive particle and per
# allocated cell for each phase, and per water cell per CG iteration.
PER_PARTICLE = {"bin": 44, "p2g": 68, "g2p": 104}
PER_CELL = {"bin": 76, "p2g": 36, "grid_prep": 64, "solve": 36, "project": 32, "extend": 76, "g2p": 32}
PER_CG_CELL_ITER = 60


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
        prec = abi.CG_MIXED if args.precision == "mixed" else abi.CG_FP32
        s = scenes.balloon_impact(precision=prec, iters=args.iters)
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


def timed_run(sim, warmup: int, steps: int, iters: int):
    """Warmup steps, then `steps` timed steps with per-phase events; wall time, counter deltas,
    phase times and the final State."""
    if warmup:
        sim.run(warmup, iters=iters)
    sim.set_profiling(True)
    c0 = sim.counters()
    t0 = time.perf_counter()
    st = sim.run(steps, iters=iters)  # returns after the stream is idle
    t1 = time.perf_counter()
    c1 = sim.counters()
    times = sim.phase_times()
    sim.set_profiling(False)
    return t1 - t0, {k: c1[k] - c0[k] for k in c0}, times, st


def compare_precision(scene, args, plain: dict) -> dict:
    """The same workload from the same initial state with the other CG precision (FP32 SpMV
    and axpy, FP64 dots, norms and convergence test, FP64 true residual every 20 iterations),
    at the same tolerance in the reference's convergence form: iterations and solve time per
    step against the plain FP32 run."""
    from natray_amd import Sim, abi

    sc = scene.scene
    keep = (sc.precision, sc.refresh_every)
    sc.precision, sc.refresh_every = abi.CG_MIXED, 20
    try:
        sim = Sim.from_scene(scene)
        wall, c, times, st = timed_run(sim, args.warmup, args.steps, args.iters)
        sim.close()
    finally:
        sc.precision, sc.refresh_every = keep
    steps = max(c["steps"], 1)
    mixed = {"cg_iters_per_step": c["cg_iters"] / steps, "solve_ms_per_step": times["solve"][0] / steps,
             "ms_per_step": wall * 1e3 / max(args.steps, 1), "unconverged_steps": st.unconverged}
    return {
        "tol": scene.params.solve[0],
        "plain_fp32": plain,
        "mixed_f32_f64": mixed,
        "solve_speedup_mixed_over_plain": plain["solve_ms_per_step"] / mixed["solve_ms_per_step"]
        if mixed["solve_ms_per_step"] > 0 else None,
        "warm_start": "particle-carried pressure (the reference's p0) in both runs",
    }


def cpu_baseline(scene, iters: int, budget_s: float, max_steps: int) -> dict:
    """The FP32 oracle (OpenMP over the host's cores) on the same workload from the same
    initial state: one untimed step, then steps until `budget_s` of wall time or max_steps."""
    import oracle

    threads = min(16, os.cpu_count() or 1)
    oracle.set_threads(threads)
    o = oracle.Oracle.from_scene(scene, "f32")
    o.run(1, iters)
    steps, t0 = 0, time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.run(1, iters)
        steps += 1
    dt = time.perf_counter() - t0
    o.close()
    return {"value": scene.n * steps / dt, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps (steps 2..{steps + 1}) of the same {scene.n}-particle workload, "
                      f"FP32 C restatement of the reference kernels (oracle/), OpenMP, {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["impact", "sphere"], default="impact")
    ap.add_argument("--precision", choices=["fp32", "mixed"], default="fp32")
    ap.add_argument("--iters", type=int, default=1000, help="CG iteration cap per step")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-compare", action="store_true", help="skip the mixed-precision comparison run")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--write-scene", metavar="PATH",
                    help="write the workload's scene file for natray_amd/natray_drive and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "the scene would shard by whole scenes or by Morton-ordered block ranges with halo exchange, "
              "which is not built.", file=sys.stderr)
        return 2

    from natray_amd import Sim, scenes

    scene, cfg = make_scene(args)
    if args.write_scene:
        scenes.write_scene_file(scene, args.write_scene)
        print(f"wrote {args.write_scene}: {cfg['workload']}")
        return 0
    sim = Sim.from_scene(scene)
    wall, c, times, st = timed_run(sim, args.warmup, args.steps, args.iters)
    sim.close()
    steps_done = c["steps"]
    particles = c["particles"]
    value = particles / wall if wall > 0 else 0.0

    pb = phase_bytes(c)
    single = ("p2g", "grid_prep", "solve", "project", "g2p")  # phases that are one kernel
    dominant = max(single, key=lambda k: times[k][0])
    ms, launches = times[dominant]
    avg_s = ms * 1e-3 / max(launches, 1)
    bytes_per_launch = pb[dominant] / max(launches, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    kernel_names = {"p2g": "k_p2g", "grid_prep": "k_grid_prep", "solve": "k_solve", "project": "k_project",
                    "g2p": "k_g2p"}
    t_cg = times["solve"][0] * 1e-3
    t_all_kernels = sum(v[0] for v in times.values()) * 1e-3
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
        "water_cells_per_step": c["water_cells"] / max(steps_done, 1),
        "final_state": {"t_ns": st.t_ns, "step": st.step, "dt_ns": st.dt_ns, "unconverged": st.unconverged,
                        "max_speed": st.max_speed},
        "phases_ms_per_step": {k: v[0] / max(steps_done, 1) for k, v in times.items()},
        "phases_algorithmic_GBps": {k: (pb[k] / (times[k][0] * 1e-3) / 1e9 if times[k][0] > 0 else 0.0) for k in pb},
        "kernel_time_fraction_of_wall": t_all_kernels / wall if wall > 0 else 0.0,
        "roofline": {
            "kernel": kernel_names[dominant],
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            # Counter-measured HBM bytes per launch of this kernel on this workload (FP32 form);
            # None where no profile covers it.
            "traffic": (measured_traffic("k_solve<false>" if dominant == "solve" else kernel_names[dominant])
                        if args.workload == "impact" and args.precision == "fp32" else None),
            "avg_launch_us": avg_s * 1e6,
            "algorithmic_bytes_per_launch": bytes_per_launch,
        },
    }
    if args.workload == "impact" and args.precision == "fp32" and not args.no_compare:
        out["precision_compare"] = compare_precision(
            scene, args, {"cg_iters_per_step": out["cg_iters_per_step"],
                          "solve_ms_per_step": out["phases_ms_per_step"]["solve"],
                          "ms_per_step": out["ms_per_step"], "unconverged_steps": st.unconverged})
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.iters, args.cpu_seconds, 40)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
