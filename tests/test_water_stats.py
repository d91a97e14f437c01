"""natray_sim_water_stats (debug_stats' water line) on the GPU against the oracle's restatement
over the same particles: the counts, the box, the top speed and the fastest particle bit for
bit; the float64 sums to 1e-12 relative (both sum exact f32 values in float64, in different
orders)."""
import numpy as np
import pytest

import oracle
from natray_amd import Sim, scenes
from tests.test_gpu_parity import say

pytestmark = pytest.mark.gpu


def check(sim, U, label):
    P = sim.particles()
    want = oracle.water_stats(P, list(U.origin)[:3], float(U.origin[3]))
    got = sim.water_stats()
    say(f"[water] {label}: alive {got['alive']} full cells {got['full_cells']} vmax {got['vmax']:.3f}")
    for k in ("alive", "full_cells"):
        assert got[k] == want[k], (label, k, got[k], want[k])
    for k in ("lo", "hi", "fastest"):
        np.testing.assert_array_equal(np.float32(got[k]), np.float32(want[k]), err_msg=f"{label} {k}")
    assert np.float32(got["vmax"]) == np.float32(want["vmax"]), label
    for k in ("mean", "vmean"):
        np.testing.assert_allclose(got[k], want[k], rtol=1e-12, atol=1e-300, err_msg=f"{label} {k}")
    assert got["ke"] == pytest.approx(want["ke"], rel=1e-12)
    assert got["volume_l"] == pytest.approx(want["volume_l"], rel=1e-12)
    return got


def test_water_stats_match_the_oracle():
    s = scenes.balloon_impact(side=64, n=32768, seed=2, iters=400, capacity=4096)
    sim = Sim.from_scene(s)
    first = check(sim, s.params, "at rest")
    assert first["alive"] == s.n and first["vmax"] == 0.0
    sim.run(6, iters=400)
    moving = check(sim, s.params, "after 6 steps")
    assert moving["vmax"] > 0.0 and moving["full_cells"] > 0
    line = sim.debug_stats()
    assert f"water {moving['alive']}/{s.n}" in line and "volume" in line
    # Every particle dead: the reference's empty fold (box at +-FLT_MAX, nothing fastest).
    P = sim.particles()
    P[:, 7] = 0.0
    sim.write_particles(P)
    dead = check(sim, s.params, "all dead")
    assert dead["alive"] == 0 and dead["full_cells"] == 0
    sim.close()


def test_water_stats_at_size():
    """The 2M impact after 3 steps (2,097,152 particles: a 4M-slot cell table, 1024 workgroups)."""
    s = scenes.balloon_impact()
    sim = Sim.from_scene(s)
    sim.run(3, iters=int(s.params.limits[2]))
    got = check(sim, s.params, "2M after 3 steps")
    assert got["alive"] == s.n and got["full_cells"] > 100000
    sim.close()
