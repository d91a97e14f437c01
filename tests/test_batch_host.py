"""Host side of the batches of scenes: the workgroup shares and the batched scene set."""
import math

import numpy as np
import pytest

from natray_amd import abi, scenes
from natray_amd.batch import plan


def test_plan_is_proportional_and_fills_the_budget():
    w = np.array([1, 2, 3, 4], np.uint64)
    out = plan(w, 100)
    assert out.sum() == 100
    np.testing.assert_array_equal(out, [10, 20, 30, 40])
    out = plan(np.array([1, 1, 1], np.uint64), 10)  # largest remainder, ties to the lowest index
    np.testing.assert_array_equal(out, [4, 3, 3])


def test_plan_respects_min_and_max():
    w = np.array([1000, 1, 1, 1], np.uint64)
    out = plan(w, 64, min_each=2)
    assert out.sum() == 64 and out.min() >= 2
    out = plan(w, 64, min_each=1, max_each=np.array([10, 100, 100, 100], np.uint32))
    assert out[0] == 10 and out.sum() == 64  # the capped scene's share goes to the others
    out = plan(w, 8, min_each=1, max_each=np.array([2, 2, 2, 2], np.uint32))
    np.testing.assert_array_equal(out, [2, 2, 2, 2])  # the bounds stop it short of the budget


@pytest.mark.parametrize("n,budget", [(64, 256), (7, 2048), (1, 5)])
def test_plan_random_weights(n, budget):
    rng = np.random.default_rng(n)
    w = rng.integers(1, 1 << 20, n).astype(np.uint64)
    out = plan(w, budget, min_each=1)
    assert out.sum() == budget and out.min() >= 1
    ideal = budget * w / w.sum()
    assert np.abs(out - ideal).max() <= 1.0 + 1e-9 + (1 if budget < 2 * n else 0)


def test_batched_scene_set():
    """The 64 scenes: counts log-uniform in [16k, 512k] (multiples of 8), sides multiples of 4
    in [64, 192] that hold the ball and the capsule, speeds in [200, 400], unit directions,
    mixed-precision CG; the same set for the same seed."""
    a = scenes.batch_impacts()
    b = scenes.batch_impacts()
    assert len(a) == 64
    for s, t in zip(a, b):
        np.testing.assert_array_equal(s.particles, t.particles)
        n, side = s.n, int(s.params.dims[0])
        assert 16384 <= n <= 524288 and n % 8 == 0
        assert 64 <= side <= 196 and side % 4 == 0
        assert (3.0 * (n // 8) / (4.0 * math.pi)) ** (1 / 3) + 10 <= side / 2
        v = np.array(s.scene.capsule_v[:3], np.float64)
        assert 200.0 - 1e-3 <= np.linalg.norm(v) <= 400.0 + 1e-3
        assert s.scene.precision == abi.CG_MIXED and s.scene.refresh_every == 20
        assert int(s.params.bdims[3]) <= (side // 4) ** 3
    total = sum(s.n for s in a)
    assert 5e6 < total < 15e6
