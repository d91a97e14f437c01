"""The oracle's water diagnostics (debug_stats' water line, src/sim.rs:691-716) against values
worked out by hand on a handful of particles. Each figure below fails for a plausible
mistake: counting dead particles, keeping the last of two equal top speeds, a cell threshold
of > 4, or cells taken from the position without the origin."""
import numpy as np
import pytest

import oracle


def record(x, v, alive=1.0):
    p = np.zeros(20, np.float32)
    p[0:3], p[4:7], p[7] = x, v, alive
    return p


def hand_set():
    # h = 0.5 with the origin at (-0.25, 0, 0): cell (2, 1, 1) spans x in [0.75, 1.25),
    # y and z in [0.5, 1.0).
    P = [
        record((0.80, 0.60, 0.60), (3.0, 4.0, 0.0)),   # speed 5, the first of the two fastest
        record((1.00, 0.70, 0.90), (0.0, 1.0, 0.0)),
        record((1.20, 0.99, 0.51), (0.0, 3.0, 4.0)),   # speed 5 too
        record((0.90, 0.75, 0.80), (1.0, 0.0, 0.0)),   # 4 particles in cell (2, 1, 1)
        record((0.10, 0.20, 0.30), (0.0, 0.0, -2.0)),  # cell (0, 0, 0), alone
        record((9.0, 9.0, 9.0), (100.0, 0.0, 0.0), alive=0.0),  # dead: ignored everywhere
    ]
    return np.stack(P), (-0.25, 0.0, 0.0), 0.5


def test_water_stats_known_answers():
    P, origin, h = hand_set()
    w = oracle.water_stats(P, origin, h)
    assert w["alive"] == 5
    np.testing.assert_array_equal(w["lo"], np.float32([0.10, 0.20, 0.30]).tolist())
    np.testing.assert_array_equal(w["hi"], np.float32([1.20, 0.99, 0.90]).tolist())
    xs = np.float32([[0.80, 0.60, 0.60], [1.00, 0.70, 0.90], [1.20, 0.99, 0.51], [0.90, 0.75, 0.80], [0.10, 0.20, 0.30]])
    np.testing.assert_allclose(w["mean"], xs.astype(np.float64).mean(0), rtol=1e-15)
    assert w["vmean"] == pytest.approx([4.0 / 5, 8.0 / 5, 2.0 / 5], rel=1e-15)
    # 0.5 (25 + 1 + 25 + 1 + 4) = 28
    assert w["ke"] == 28.0
    assert w["vmax"] == 5.0 and w["fastest"] == np.float32([0.80, 0.60, 0.60]).tolist()
    assert w["full_cells"] == 1 and w["volume_l"] == pytest.approx(125.0)


def test_water_stats_without_alive_particles():
    P, origin, h = hand_set()
    P[:, 7] = 0.0
    w = oracle.water_stats(P, origin, h)
    big = float(np.finfo(np.float32).max)
    assert w["alive"] == 0 and w["full_cells"] == 0 and w["ke"] == 0.0 and w["vmax"] == 0.0
    assert w["lo"] == [big] * 3 and w["hi"] == [-big] * 3 and w["fastest"] == [0.0, 0.0, 0.0]


def test_water_stats_cells_by_brute_force():
    """Cell membership against a plain loop with Python floor on the same f32 quotients, on
    random particles around cell boundaries."""
    rng = np.random.default_rng(5)
    n = 400
    x = (rng.integers(0, 6, (n, 3)) * 0.125 + rng.choice([0.0, 1e-7, 0.0624], (n, 3))).astype(np.float32)
    P = np.stack([record(x[i], (0, 0, 0)) for i in range(n)])
    origin, h = (0.0, 0.0, 0.0), 0.125
    counts = {}
    for i in range(n):
        c = tuple(int(np.floor(float((x[i, a] - np.float32(0.0)) / np.float32(h)))) for a in range(3))
        counts[c] = counts.get(c, 0) + 1
    assert oracle.water_stats(P, origin, h)["full_cells"] == sum(1 for c in counts.values() if c >= 4)
