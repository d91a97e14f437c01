"""GPU parity at the benchmarked sizes, bit for bit against the FP32 oracle.

The small parity scenes (tests/test_gpu_parity.py) hold at most ~180 blocks, so the solve's
chunk fold runs with one block per chunk and each workgroup holds at most one block. Here the
scenes are the workloads themselves:

- the 2,097,152-particle balloon impact on 256^3 (~5,630 blocks): chunks of K = 22 blocks, 1-2
  blocks per wave in the register form, the memory form's streaming passes with 64 solve
  workgroups (88 blocks each), and its per-block passes with 8 (~700 each);
- a 262,144-particle impact on 128^3 (~700 blocks, K = 3);
- the same 2M impact with the mixed-precision and the pipelined CG, in both forms;
- the 64-scene batch of balloon impacts (16k-512k particles on 64^3-192^3), two steps against
  the scenes stepped alone, and its largest scene against the oracle;
- the warm start on the 2M impact with the hand-set Richardson networks.

Each runs phase by phase for three steps (the first step starts from rest and needs no CG
iterations; the next two solve with ~26-28 iterations).
"""
import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_batch import alone, batched, check_same
from tests.test_gpu_parity import assert_same_float, compare_state, phase_parity, say
from tests.test_warm_start import warm_start_parity

pytestmark = pytest.mark.gpu

STEPS = 3


def solve_groups(s, groups):
    s.scene.solve_groups = groups
    return s


AT_SIZE = {
    "impact_2M": lambda: scenes.balloon_impact(),
    "impact_2M_mem": lambda: solve_groups(scenes.balloon_impact(), 64),
    "impact_262k_128": lambda: scenes.balloon_impact(side=128, n=262144, seed=6),
    "impact_2M_mixed": lambda: scenes.balloon_impact(precision=abi.CG_MIXED),
    "impact_2M_mixed_mem": lambda: solve_groups(scenes.balloon_impact(precision=abi.CG_MIXED), 64),
    "impact_2M_pipelined": lambda: scenes.balloon_impact(precision=abi.CG_PIPELINED),
    "impact_2M_pipelined_mem": lambda: solve_groups(scenes.balloon_impact(precision=abi.CG_PIPELINED), 64),
    # 8 workgroups of ~700 blocks: the memory form's per-block passes (beyond the streaming
    # passes' 512 blocks per workgroup), with its fenced reductions, in FP32 and mixed.
    "impact_2M_mem8": lambda: solve_groups(scenes.balloon_impact(), 8),
    "impact_2M_mixed_mem8": lambda: solve_groups(scenes.balloon_impact(precision=abi.CG_MIXED), 8),
    # The stated geometry of the water sphere: radius 0.25 on 64^3 at 8 particles per cell.
    "sphere_stated_r025_138k": lambda: scenes.stated_radius_sphere(),
}

# Blocks the scenes hold: more than 256 (several blocks per chunk) in every case.
MIN_BLOCKS = {"impact_262k_128": 600, "sphere_stated_r025_138k": 500}


@pytest.mark.parametrize("name", list(AT_SIZE))
def test_phases_bit_exact_at_size(name):
    st = phase_parity(name, AT_SIZE[name](), steps=STEPS)
    assert st.n_blocks >= MIN_BLOCKS.get(name, 5000), (name, st.n_blocks)
    assert st.solve_iters > 10, (name, st.solve_iters)


def test_batch_of_the_benchmark_set():
    """The 64 impacts of the batch workload (10.76M particles): two batch steps against the
    scenes stepped alone, then the largest scene's two steps alone against the oracle."""
    sl = scenes.batch_impacts()
    say(f"[at size] batch of {len(sl)} scenes, {sum(s.n for s in sl)} particles")
    want = alone(sl, 2)
    got, _ = batched(sl, 2)
    check_same("batch_impacts", got, want)
    k = int(np.argmax([s.n for s in sl]))
    s = sl[k]
    say(f"[at size] scene {k}: N={s.n} side {int(s.params.dims[0])} against the oracle")
    orc = oracle.Oracle.from_scene(s, "f32")
    b = orc.run(2, iters=int(s.params.limits[2]))
    Pw, sw = want[k]
    compare_state(f"scene {k}", sw, b)
    assert_same_float(f"scene {k} particles", Pw, orc.particles())
    assert sw.n_blocks > 256 and sw.solve_iters > 10, (sw.n_blocks, sw.solve_iters)
    orc.close()


@pytest.mark.parametrize("name,c", [("impact_2M", 0.1), ("impact_2M", -0.1), ("impact_2M_pipelined_mixed", 0.1)])
def test_warm_start_at_size(name, c):
    if name == "impact_2M_pipelined_mixed":
        s = scenes.balloon_impact(precision=abi.CG_PIPELINED)
    else:
        s = AT_SIZE[name]()
    taken = warm_start_parity(name, s, c, steps=STEPS, max_tiles=1024)
    # The first step is at rest (r0 = 0, nothing to take); a c > 0 guess is taken after that.
    assert taken == (STEPS - 1 if c > 0 else 0), (name, c, taken)


def test_stated_radius_impact_runs_like_the_oracle():
    """The balloon impact's stated geometry: radius 0.3 on 256^3 at 8 particles per cell,
    15,179,264 particles in ~35,000 blocks (K = 138 blocks per chunk, the solve's memory form
    at 256 workgroups). Two whole steps through natray_sim_run against the oracle's run."""
    s = scenes.stated_radius_impact()
    say(f"[at size] stated radius impact N={s.n}")
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    a = sim.run(2, iters=int(s.params.limits[2]))
    b = orc.run(2, iters=int(s.params.limits[2]))
    compare_state("stated radius", a, b)
    assert_same_float("stated radius particles", sim.particles(), orc.particles())
    assert a.n_blocks > 30000 and a.solve_iters > 10, (a.n_blocks, a.solve_iters)
    sim.close()
    orc.close()
