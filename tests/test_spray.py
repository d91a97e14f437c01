"""Spray and wet ground on the GPU (NATRAY_SCENE_SPRAY) against the oracle, step by step on a
scene where lone particles break up, drops fly and land, and ground particles wet the ground:
particles, State (with the spray count) and the wet map bit for bit; the spray records and
where they landed as a multiset of rows, also bit for bit (slots come from an atomic add, so
their order differs; a record's motion does not depend on its slot)."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_gpu_parity import assert_same_float, compare_state

pytestmark = pytest.mark.gpu

KEYS = ("t_ns", "dt_ns", "step", "n_blocks", "n_blocks_prev", "overflow", "solve_iters", "iters_max", "unconverged",
        "free_dt_ns", "spray_count")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rows(sp):
    r = np.concatenate([sp["records"].astype(np.float32), sp["landed"].astype(np.float32)], axis=1)
    return r[np.lexsort(r.T[::-1])]


@pytest.mark.parametrize("graph", [True, False])
def test_spray_and_wet_ground_bit_exact(graph):
    s = scenes.breakup_scene()
    say(f"[spray] breakup scene N={s.n}, graph {graph}")
    sim = Sim.from_scene(s)
    sim.set_graph(graph)
    orc = oracle.Oracle.from_scene(s, "f32")
    landed = 0
    for step in range(10):
        say(f"[spray] step {step}")
        a = sim.run(1, iters=200)
        b = orc.run(1, 200)
        compare_state(f"step {step}", a, b, keys=KEYS)
        assert_same_float(f"step {step} particles", sim.particles(), orc.particles())
        np.testing.assert_array_equal(sim.wet(), orc.wet(), err_msg=f"step {step} wet map")
        g, o = sim.spray(), orc.spray()
        assert g["count"] == o["count"] and len(g["records"]) == len(o["records"])
        assert_same_float(f"step {step} spray rows", rows(g), rows(o))
        landed = int((g["landed"][:, 3] > 0).sum())
    assert a.spray_count >= 300 and landed > 50
    sim.close()
    orc.close()
