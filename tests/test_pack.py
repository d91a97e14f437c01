"""The keyframe pack of the water on the GPU against the oracle, bit for bit, after steps of
the impact (live and dead particles, positions anywhere in the box)."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim
from tests.test_gpu_parity import SCENES

pytestmark = pytest.mark.gpu


def test_pack_water_matches_oracle():
    print("[pack] impact_sgs_32k", flush=True)
    s = SCENES["impact_sgs_32k"]()
    s.particles[::97, 7] = 0.0  # some dead particles
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    for step in range(3):
        np.testing.assert_array_equal(sim.pack_water(), orc.pack_water(), err_msg=f"step {step}")
        sim.run(2, iters=400)
        orc.run(2, iters=400)
    w = sim.pack_water()
    np.testing.assert_array_equal(w, orc.pack_water())
    assert ((w[:, 1] >> 31) == 0).sum() >= len(w) // 97
    sim.close()
    orc.close()
