"""GPU parity: every phase of the substep through the C ABI against the FP32 oracle.

Both sides start from the same seeded scene and run phase by phase. Integer data (block
lists, pointers, starts, the fixed-point transfer sums acc) must be bit-exact, and so must the
float fields: the kernels evaluate every float expression in the oracle's order without
contraction, and the solver reduces in the same fixed order on both sides. The one set that is
compared as a multiset is `sorted` within each block (ranks come from atomics on the GPU).
"""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from natray_amd.params import SimParams

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_reference():
    s = scenes.reference_scene(SimParams(resolution=16), capacity=4096)
    s.params.limits[2] = 64
    s.params.limits[3] = 0xFFFFFFFF
    return s


SCENES = {
    "sphere_jacobi_32k": lambda: scenes.water_sphere(),
    "impact_sgs_32k": lambda: scenes.balloon_impact(side=64, n=32768, seed=3, iters=400, capacity=4096),
    "impact_mixed_32k": lambda: scenes.balloon_impact(side=64, n=32768, seed=4, iters=400, capacity=4096,
                                                       precision=abi.CG_MIXED, refresh_every=5),
    "reference_balloon_res16": small_reference,
    # Replacement every 5 iterations: the first solves replace their residual once.
    "impact_pipelined_32k": lambda: scenes.balloon_impact(side=64, n=32768, seed=5, iters=400, capacity=4096,
                                                           precision=abi.CG_PIPELINED, refresh_every=5),
}


def few_solve_groups(name, groups=2):
    """The same scene with the pressure solve on `groups` workgroups, so each holds more blocks
    than its waves keep in registers and the solve runs on the fields in memory."""
    def make():
        s = SCENES[name]()
        s.scene.solve_groups = groups
        return s
    return make


def dense_sphere(per_cell, extra_cell=0):
    """The water sphere's 4,096 cells on 64^3 with `per_cell` particles each (copies of the
    seeded fill at other seeds), plus `extra_cell` more in one cell. Full blocks of more than
    606 particles take P2G's 32-bit atomics, of more than 2,048 its unordered splat, and a cell
    of more than 64 particles its splat in stored order."""
    def make():
        s = scenes.water_sphere()
        parts = [s.particles] + [scenes.sphere_particles(64, 32768, seed=100 + k) for k in range(1, per_cell // 8)]
        if extra_cell:
            P = np.zeros((extra_cell, abi.PARTICLE_FLOATS), np.float32)
            u = scenes.uniform01(77, 3 * extra_cell).reshape(-1, 3)
            P[:, :3] = ((np.array([32.0, 32.0, 32.0]) + u) / 64.0).astype(np.float32)
            P[:, 3] = 1.0
            P[:, 7] = 1.0
            parts.append(P)
        s.particles = np.concatenate(parts)
        s.params.dims[3] = s.n
        return s
    return make


SCENES["sphere_16ppc"] = dense_sphere(16)
SCENES["sphere_40ppc"] = dense_sphere(40)
SCENES["sphere_crowded_cell"] = dense_sphere(8, extra_cell=72)
SCENES["sphere_jacobi_32k_mem"] = few_solve_groups("sphere_jacobi_32k")
SCENES["impact_mixed_32k_mem"] = few_solve_groups("impact_mixed_32k")
SCENES["impact_pipelined_32k_mem"] = few_solve_groups("impact_pipelined_32k")


def assert_same_float(name, a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (name, a.shape, b.shape)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    if not same.all():
        bad = np.argwhere(~same)
        i = tuple(bad[0])
        d = np.abs(a.astype(np.float64) - b.astype(np.float64))
        raise AssertionError(f"{name}: {len(bad)} of {a.size} differ; first {i}: gpu {a[i]!r} oracle {b[i]!r}; max abs diff {d.max():.3e}")


def compare_grid(tag, g, o, cf_fields=range(abi.CELL_FIELDS), acc=True):
    assert g["n_blocks"] == o["n_blocks"], (tag, g["n_blocks"], o["n_blocks"])
    np.testing.assert_array_equal(g["block_list"], o["block_list"], err_msg=tag)
    np.testing.assert_array_equal(g["block_start"], o["block_start"], err_msg=tag)
    if acc:
        np.testing.assert_array_equal(g["acc"], o["acc"], err_msg=f"{tag} acc")
    for f in cf_fields:
        assert_same_float(f"{tag} cf[{f}]", g["cf"][f], o["cf"][f])


def compare_sorted(tag, sim, orc, grid):
    """`sorted` as a multiset within each block: both sides sorted by (block, particle)."""
    bp_g, so_g = sim.bins()
    bp_o, so_o = orc.bins()
    np.testing.assert_array_equal(bp_g, bp_o, err_msg=f"{tag} block_ptr")
    bs = grid["block_start"].astype(np.int64)
    n = grid["n_blocks"]
    blk = np.repeat(np.arange(n), np.diff(bs[:n + 1]))
    lo, hi = int(bs[0]), int(bs[n])
    a = so_g[lo:hi][np.lexsort((so_g[lo:hi], blk))]
    b = so_o[lo:hi][np.lexsort((so_o[lo:hi], blk))]
    if not np.array_equal(a, b):
        k = int(np.argmax(a != b))
        raise AssertionError(f"{tag}: particles of block {int(blk[k])} differ")


def compare_state(tag, a, b, keys=("t_ns", "dt_ns", "step", "n_blocks", "n_blocks_prev", "overflow", "solve_iters",
                                  "iters_max", "unconverged", "free_dt_ns")):
    for k in keys:
        assert getattr(a, k) == getattr(b, k), (tag, k, getattr(a, k), getattr(b, k))
    for k in ("max_speed", "rz", "alpha", "beta", "b_norm", "r_norm", "dt_flow"):
        assert_same_float(f"{tag} state.{k}", np.float32(getattr(a, k)), np.float32(getattr(b, k)))


def phase_parity(name, s, steps=4):
    """Steps the scene phase by phase on the GPU and on the oracle and compares every phase's
    State and grid, each binning's `sorted`, and the particles after G2P."""
    say(f"[parity] scene {name}: N={s.n} dims={tuple(s.params.dims[:3])}")
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    for step in range(steps):
        for ph in range(abi.N_PHASES):
            say(f"[parity] {name} step {step} phase {abi.PHASE_NAMES[ph]}")
            sim.phase(ph)
            orc.phase(ph)
            tag = f"{name} step {step} after {abi.PHASE_NAMES[ph]}"
            compare_state(tag, sim.state(), orc.state())
            if ph >= abi.PHASE_BIN:
                g, o = sim.grid(), orc.grid()
                if ph == abi.PHASE_BIN:
                    compare_grid(tag, g, o, cf_fields=(), acc=False)
                    compare_sorted(tag, sim, orc, g)
                else:
                    compare_grid(tag, g, o)
                    if ph >= abi.PHASE_P2G:
                        np.testing.assert_array_equal(g["block_nb"], o["block_nb"], err_msg=tag)
            if ph == abi.PHASE_G2P:
                assert_same_float(f"{tag} particles", sim.particles(), orc.particles())
    st = sim.state()
    sim.close()
    orc.close()
    return st


@pytest.mark.parametrize("name", list(SCENES))
def test_phases_bit_exact(name):
    phase_parity(name, SCENES[name]())


@pytest.mark.parametrize("name", ["sphere_jacobi_32k", "impact_sgs_32k", "impact_pipelined_32k"])
def test_run_matches_oracle(name):
    """Ten whole steps through natray_sim_run against the oracle's run."""
    s = SCENES[name]()
    say(f"[parity] run 10 steps of {name}")
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    a = sim.run(10, iters=int(s.params.limits[2]), stop_ns=0xFFFFFFFF)
    b = orc.run(10, iters=int(s.params.limits[2]), stop_ns=0xFFFFFFFF)
    compare_state(name, a, b)
    assert_same_float(f"{name} particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()


def test_repeat_runs_identical():
    """Two handles stepping the same scene 30 steps end bit-identical, and equal the oracle:
    the GPU path has no order-dependent float arithmetic and no races."""
    s = SCENES["impact_sgs_32k"]()
    say("[parity] determinism: 2 x 30 steps of impact_sgs_32k")
    a = Sim.from_scene(s)
    b = Sim.from_scene(s)
    a.run(30, iters=400)
    for _ in range(3):
        b.run(10, iters=400)
    assert_same_float("repeat particles", a.particles(), b.particles())
    orc = oracle.Oracle.from_scene(s, "f32")
    orc.run(30, iters=400)
    assert_same_float("oracle particles", a.particles(), orc.particles())
    compare_state("repeat", a.state(), orc.state())


@pytest.mark.parametrize("name", ["impact_sgs_32k", "impact_mixed_32k"])
def test_graph_replays_match_kernel_by_kernel(name):
    """natray_sim_run replays captured HIP graphs: one of 8 steps for each 8 of a run, one of a
    step for the rest. Every replay gives the bits of the same steps launched kernel by kernel,
    with other work and readbacks between the replays."""
    s = SCENES[name]()
    say(f"[parity] graph replays vs eager launches: {name}")
    g = Sim.from_scene(s)
    e = Sim.from_scene(s)
    e.set_graph(False)
    for k, n in enumerate((3, 8, 11, 17)):
        a = g.run(n, iters=400)
        b = e.run(n, iters=400)
        compare_state(f"{name} run {k}", a, b)
        assert_same_float(f"{name} run {k} particles", g.particles(), e.particles())
        gg, ge = g.grid(), e.grid()
        compare_grid(f"{name} run {k}", gg, ge)
    g.close()
    e.close()
