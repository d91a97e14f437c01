"""Bullets on the GPU (natray_sim_set_bullets) against the oracle, step by step: two bullets
fired into a water sphere, one arriving later and at an angle, through their flight in the
air, their entry and the cavity they open. Particles, State, the bullets, their cavity
sections and their pose log bit for bit, with and without the step graphs."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_gpu_parity import assert_same_float, compare_state

pytestmark = pytest.mark.gpu

KEYS = ("t_ns", "dt_ns", "step", "n_blocks", "n_blocks_prev", "overflow", "solve_iters", "iters_max", "unconverged",
        "free_dt_ns")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def two_shots():
    b0, p0 = scenes.bullet_shot(arrive_ns=30000)
    b1, p1 = scenes.bullet_shot(at=(0.5, 0.8, 0.3), direction=(0.0, -1.0, 0.6), speed=250.0, arrive_ns=120000,
                                spin_rev_s=2000.0, yaw=0.05)
    return [b0, b1], np.stack([p0, p1])


@pytest.mark.parametrize("graph", [True, False])
def test_bullets_bit_exact(graph):
    s = scenes.bullet_scene()
    bullets, patches = two_shots()
    say(f"[bullet] water sphere N={s.n}, 2 bullets, graph {graph}")
    sim = Sim.from_scene(s)
    sim.set_graph(graph)
    sim.set_bullets(bullets, patches)
    orc = oracle.Oracle.from_scene(s, "f32")
    orc.set_bullets(bullets, patches)
    for step in range(36):
        say(f"[bullet] step {step}")
        a = sim.run(1, iters=200)
        b = orc.run(1, 200)
        compare_state(f"step {step}", a, b, keys=KEYS)
        gb, gs = sim.bullets()
        ob, osec = orc.bullets()
        for k in range(2):
            assert_same_float(f"step {step} bullet {k}", gb[k].as_array(), ob[k].as_array())
        assert_same_float(f"step {step} sections", gs, osec.astype(np.float32))
        np.testing.assert_array_equal(sim.pose_log().view(np.uint32), orc.pose_log().view(np.uint32),
                                      err_msg=f"step {step} pose log")
        assert_same_float(f"step {step} particles", sim.particles(), orc.particles())
    # both joined, the first is in the water with its cavity open
    assert gb[0].info[2] == 1 and gb[1].info[2] == 1
    assert gb[0].info[3] >= 10 and gb[0].reach[3] > 0
    assert np.linalg.norm(list(gb[0].v[:3])) < 290.0
    sim.close()
    orc.close()


def test_bullets_removed_leave_no_source():
    s = scenes.bullet_scene()
    bullets, patches = two_shots()
    say(f"[bullet] removal, N={s.n}")
    sim = Sim.from_scene(s)
    sim.set_bullets(bullets, patches)
    orc = oracle.Oracle.from_scene(s, "f32")
    orc.set_bullets(bullets, patches)
    sim.run(20, iters=200)
    orc.run(20, 200)
    empty = np.zeros((0, abi.BULLET_PATCHES, 8), np.float32)
    sim.set_bullets([], empty)
    orc.set_bullets([], empty)
    for step in range(4):
        a = sim.run(1, iters=200)
        b = orc.run(1, 200)
        compare_state(f"after removal {step}", a, b, keys=KEYS)
        assert_same_float(f"after removal {step} particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()


def test_bad_bullets_are_refused_before_any_launch():
    s = scenes.bullet_scene(side=32, n=3200)
    say(f"[bullet] argument checks, N={s.n}")
    sim = Sim.from_scene(s)
    b, pt = scenes.bullet_shot()
    five = np.stack([pt] * 5)
    with pytest.raises(Exception):
        abi.check(abi.lib().natray_sim_set_bullets(sim._h, 5, abi.ptr(np.frombuffer(bytes(b) * 5, np.uint8).copy()),
                                                   abi.ptr(five)))
    bad = abi.Bullet.from_buffer_copy(bytes(b))
    bad.info[1] = abi.BULLET_PATCHES + 1
    with pytest.raises(Exception):
        sim.set_bullets([bad], pt[None])
    # the scene still steps as one without bullets, bit for bit
    orc = oracle.Oracle.from_scene(s, "f32")
    compare_state("no bullets", sim.run(3, iters=200), orc.run(3, 200), keys=KEYS)
    assert_same_float("no bullets particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()
