"""Batches of scenes on the GPU: every scene of a batch gives the same bits as the same scene
stepped alone (which tests/test_gpu_parity.py holds bit-exact against the oracle), with the
steps replayed as graphs or launched kernel by kernel, with the solve in register form (few
scenes) and in memory form (many scenes sharing the solve's workgroups), for both CG
precisions and the pipelined CG, with the solve launched plain or cooperatively, over step
counts that use the 8-step graph and the one-step graph; and the autotuner's choice leaves the
results unchanged."""
import sys

import numpy as np
import pytest

from natray_amd import Sim, abi, scenes
from natray_amd.batch import Batch
from tests.test_gpu_parity import assert_same_float, compare_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def small_set(precision):
    specs = [(32, 4096, (1.0, 0.0, 0.0)), (48, 8192, (0.0, 1.0, 1.0)), (40, 2048, (-1.0, 0.3, 0.0)),
             (64, 16384, (0.2, -1.0, 0.5))]
    return [scenes.balloon_impact(side=side, n=n, seed=10 + k, speed=250.0 + 40 * k, direction=d, precision=precision,
                                  iters=400, capacity=(side // 4) ** 3)
            for k, (side, n, d) in enumerate(specs)]


def many_set(count=64, precision=abi.CG_MIXED):
    """Scenes with more blocks than their share of the solve's workgroups holds in registers."""
    return [scenes.balloon_impact(side=48, n=32768, seed=100 + k, speed=300.0, precision=precision, iters=400,
                                  direction=(1.0, 0.1 * k, 0.0), capacity=(48 // 4) ** 3)
            for k in range(count)]


def alone(scene_list, steps):
    out = []
    for s in scene_list:
        sim = Sim.from_scene(s)
        sim.run(steps, iters=int(s.params.limits[2]))
        out.append((sim.particles(), sim.state()))
        sim.close()
    return out


def batched(scene_list, steps, graph=True, tune=False, launch=abi.SOLVE_PLAIN):
    sims = [Sim.from_scene(s) for s in scene_list]
    b = Batch(sims)
    b.set_graph(graph)
    b.set_solve_launch(launch)
    chosen = None
    if tune:
        chosen = b.tune(1, iters=400)
        steps -= 2 * 9  # the tuner's steps (two rounds of one step per candidate) are real steps
    b.run(steps, iters=400)
    out = [(s.particles(), s.state()) for s in sims]
    b.close()
    for s in sims:
        s.close()
    return out, chosen


STATE_KEYS = ("t_ns", "dt_ns", "step", "n_blocks", "n_blocks_prev", "overflow", "solve_iters", "iters_max",
              "unconverged", "free_dt_ns")


def check_same(tag, got, want, keys=STATE_KEYS):
    for k, ((Pg, sg), (Pw, sw)) in enumerate(zip(got, want)):
        compare_state(f"{tag} scene {k}", sg, sw, keys=keys)
        assert_same_float(f"{tag} scene {k} particles", Pg, Pw)


@pytest.mark.parametrize("precision", [abi.CG_FP32, abi.CG_MIXED, abi.CG_PIPELINED])
@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_scenes_alone(precision, graph):
    say(f"[batch] 4 scenes, precision {precision}, graph {graph}")
    sl = small_set(precision)
    want = alone(sl, 11)
    got, _ = batched(sl, 11, graph=graph)
    check_same(f"precision {precision} graph {graph}", got, want)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_cooperative_solve_launch(graph):
    say(f"[batch] 4 scenes, cooperative solve launch, graph {graph}")
    sl = small_set(abi.CG_MIXED)
    want = alone(sl, 9)
    got, _ = batched(sl, 9, graph=graph, launch=abi.SOLVE_COOPERATIVE)
    check_same(f"cooperative graph {graph}", got, want)


def test_batch_solve_launch_modes():
    """A batch takes the plain and the cooperative launch; the test-only mode of a single
    scene's solve is refused with NATRAY_E_ARG, and the batch still steps afterwards."""
    say("[batch] solve launch modes")
    sims = [Sim.from_scene(s) for s in small_set(abi.CG_FP32)[:2]]
    b = Batch(sims)
    with pytest.raises(abi.NatrayError) as e:
        b.set_solve_launch(abi.SOLVE_PLAIN_TEST_CHECK)
    assert e.value.code == abi.E_ARG
    b.set_solve_launch(abi.SOLVE_COOPERATIVE)
    b.set_solve_launch(abi.SOLVE_PLAIN)
    b.run(2, iters=400)
    assert all(s.state().solve_iters > 0 for s in sims)
    b.close()
    for s in sims:
        s.close()


@pytest.mark.parametrize("precision", [abi.CG_MIXED, abi.CG_PIPELINED])
def test_batch_of_many_scenes_memory_form(precision):
    say(f"[batch] 64 scenes of 32k particles, precision {precision}")
    sl = many_set(precision=precision)
    want = alone(sl, 3)
    got, _ = batched(sl, 3)
    check_same(f"64 scenes precision {precision}", got, want)


def test_autotuned_batch_matches_scenes_alone():
    say("[batch] autotuner")
    sl = small_set(abi.CG_MIXED)
    want = alone(sl, 20)
    got, chosen = batched(sl, 20, tune=True)
    say(f"[batch] tuner chose {chosen['best']}")
    assert chosen["best"]["blocks_per_cu"] in (4, 8, 16) and chosen["best"]["waves_per_cu"] in (4, 8, 16)
    # iters_max is the largest iteration count of a run call's steps; the tuner makes many calls.
    check_same("tuned", got, want, keys=tuple(k for k in STATE_KEYS if k != "iters_max"))
