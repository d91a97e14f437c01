"""The oracle's particle-to-grid transfer (shaders/fluid.wgsl:229-303) against the float64
golden vectors of tests/golden/ and the closed forms of the quadratic B-spline.

The goldens use the textbook piecewise B-spline over every lattice node, not the per-axis
weights of the kernels, so a wrong weight, lattice offset, node offset, the h factor of the
affine term or the MPa scale of the pressure fails them. Tolerance: each fixed-point
contribution rounds by at most 1/2 unit, and its f32 evaluation is good to ~1e-6 relative.
"""
import os

import numpy as np
import pytest

import oracle
from natray_amd import abi
from tests.helpers import box_scene, cell_coords, records

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SCALES = {0: abi.FP_M, 1: abi.FP_W, 2: abi.FP_M, 3: abi.FP_W, 4: abi.FP_M, 5: abi.FP_W, 6: abi.FP_W, 7: abi.FP_P, 8: 1.0}


def p2g_oracle(P, side, precision="f32"):
    s = box_scene(P, side=side)
    o = oracle.Oracle.from_scene(s, precision)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G):
        o.phase(ph)
    g = o.grid()
    return s, o, g


@pytest.mark.parametrize("name", ["p2g_one.npz", "p2g_seven.npz", "p2g_fifty.npz"])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_p2g_matches_float64_golden(name, precision):
    z = np.load(os.path.join(GOLD, name))
    P = records(z["pos"], z["vel"], z["C"], z["pres"])
    s, o, g = p2g_oracle(P, int(z["side"]), precision)
    cc = cell_coords(g, s.params.bdims)
    index = {tuple(c): i for i, c in enumerate(cc.tolist())}
    acc = g["acc"]
    seen = np.zeros_like(acc, dtype=bool)
    for (f, x, y, zc), val, cnt, mag in zip(z["key"].tolist(), z["val"], z["cnt"], z["mag"]):
        i = index[(x, y, zc)]
        scale = SCALES[f]
        tol = 0.5 * cnt + 2e-6 * mag * scale + 1e-9
        got = float(acc[f, i])
        assert abs(got - val * scale) <= tol, (name, f, (x, y, zc), got, val * scale, tol)
        seen[f, i] = True
    # Everything the golden does not list is zero (up to rounding of weights below half a unit).
    assert np.all(acc[:9][~seen[:9]] == 0)
    assert np.all(acc[9:] == 0)


def test_partition_of_unity_and_first_moment():
    """Per particle and lattice, the weights sum to 1 (2^24 +- 27/2 in fixed point) and the
    first moment vanishes, so a particle's momentum comes out as v exactly (+- 27/2 units)."""
    rng = np.random.default_rng(5)
    side = 16
    for _ in range(20):
        pos = rng.uniform(5, 11, 3) / side
        v = rng.normal(0, 2, 3)
        C = rng.normal(0, 30, (3, 3))
        s, o, g = p2g_oracle(records(pos, v, C, 0.0), side)
        acc = g["acc"].astype(np.int64)
        for d in range(3):
            assert abs(acc[2 * d + 1].sum() - (1 << 24)) <= 14
            # momentum sum: v * 2^16; the affine term's first moment vanishes
            assert abs(acc[2 * d].sum() - v[d] * 65536.0) <= 14 + 1e-5 * 65536 * (abs(v[d]) + 1)
        assert abs(acc[6].sum() - (1 << 24)) <= 14
        assert acc[8].sum() == 1


def test_second_moment_gives_the_affine_term():
    """With v = 0, sum_i m_i (x_i - x_p) = C h^2 sum w (x_i - x_p)^2 / h ... = C / 4 h^2 per axis:
    the quadratic B-spline has variance 1/4 cell^2."""
    side = 16
    h = 1.0 / side
    pos = np.array([7.3, 8.6, 6.45]) / side
    C = np.array([[3.0, -2.0, 5.0], [1.0, 4.0, -6.0], [-2.5, 0.5, 2.0]])
    s, o, g = p2g_oracle(records(pos, [0, 0, 0], C, 0.0), side)
    cc = cell_coords(g, s.params.bdims).astype(np.float64)
    offs = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]
    x = np.array(pos, np.float32).astype(np.float64) / h
    for d in range(3):
        m = g["acc"][2 * d].astype(np.float64) / 65536.0
        node = cc + np.array(offs[d])
        first = (m[:, None] * (node - x) * h).sum(axis=0)  # sum_i m_i (x_i - x_p) in metres
        # = sum w C_d . (x_i - x_p) (x_i - x_p) = C_d h^2 / 4 per axis (the cross terms vanish)
        np.testing.assert_allclose(first, C[d] * h * h / 4.0, rtol=2e-3, atol=30 * h / 65536.0)


def test_permutation_and_whole_cell_translation():
    """Sums are integer, so any particle order gives identical bits; moving every particle by
    exactly one cell moves every sum by one cell."""
    rng = np.random.default_rng(9)
    side = 16
    n = 40
    # Positions on a 2^-20 lattice, so shifting by 1/16 is exact in f32.
    pos = np.round(rng.uniform(5, 10, (n, 3)) / side * 2**20) / 2**20
    P = records(pos, rng.normal(0, 1, (n, 3)), rng.normal(0, 10, (n, 3, 3)), rng.normal(0, 1e4, n))
    s1, o1, g1 = p2g_oracle(P, side)
    perm = rng.permutation(n)
    s2, o2, g2 = p2g_oracle(P[perm], side)
    np.testing.assert_array_equal(g1["acc"], g2["acc"])
    Q = P.copy()
    Q[:, 0] += np.float32(1.0 / side)  # exact: positions in [5/16, 11/16] and 1/16 are dyadic
    assert np.all(Q[:, 0] - P[:, 0] == np.float32(1.0 / side))
    s3, o3, g3 = p2g_oracle(Q, side)
    c1 = {tuple(c): i for i, c in enumerate(cell_coords(g1, s1.params.bdims).tolist())}
    c3 = {tuple(c): i for i, c in enumerate(cell_coords(g3, s3.params.bdims).tolist())}
    nz = np.argwhere(g1["acc"][:9] != 0)
    assert len(nz) > 100
    for f, i in nz:
        x, y, z = cell_coords(g1, s1.params.bdims)[i]
        j = c3[(x + 1, y, z)]
        assert g3["acc"][f, j] == g1["acc"][f, i]
