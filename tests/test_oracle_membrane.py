"""The oracle's membrane coupling (shaders/membrane.wgsl:151-256 and grid_prep's RUBBER and
COVERED cells, shaders/fluid.wgsl:339-355), pinned by closed forms and brute force:

- mem_splat's push and area sum to the vertex's values over the 27 cells (B-spline partition
  of unity, up to the fixed-point rounding of each cell's share), and their centroid is the
  vertex (the B-spline's first moment); the own cell gets the area and area times velocity;
- mark_rubber claims exactly the blocks of cells c-2..c+3 around rubber off the water;
- grid_prep types a GAS cell with rubber in it RUBBER, an open non-water cell under the rubber's
  pull COVERED with pressure push / area;
- mem_sample's face velocity is the weighted average of the face velocities around the vertex,
  restated here in float64 numpy; its fill is about 1 deep in the water and less outside it.
"""
import numpy as np
import pytest

import oracle
from natray_amd import abi, scenes
from tests.helpers import cell_coords

FP_A, FP_F, FP_W = 65536.0, 4194304.0, 16777216.0


def ball(side=32, n=8 * 900, seed=5):
    s = scenes.water_sphere(side=side, n=n, seed=seed, precond=abi.PRECOND_BLOCK_SGS, fixed_dt=2e-4)
    return s


def one_vertex(pos, push=(3e-3, 2e-5), vel=(0.1, -0.2, 0.3), mass=1e-4):
    x = np.array([[pos[0], pos[1], pos[2], 1.0]], np.float32)
    v = np.array([[vel[0], vel[1], vel[2], 0.0]], np.float32)
    return x, v, np.array([mass], np.float32), np.array([push], np.float32)


def to_fp(x, scale):
    return np.rint(np.clip(np.float32(x) * np.float32(scale), -2e9, 2e9))


def run_to(o, phase_last):
    for ph in range(phase_last + 1):
        o.phase(ph)


def test_splat_partition_and_first_moment():
    s = ball()
    h = float(s.params.origin[3])
    c = np.array([0.5, 0.5, 0.5]) + np.array([0.13, -0.21, 0.07]) * h  # inside the water ball
    x, v, m, p = one_vertex(c)
    o = oracle.Oracle.from_scene(s, "f32")
    o.set_membrane(x, v, m, p)
    o.run(1, 200)  # mem_sample marks the vertex as on the water
    assert o.membrane_water()[0, 3] > 0.5
    run_to(o, abi.PHASE_P2G)
    g = o.grid()
    mf = g["acc"][abi.A_MF].astype(np.float64)
    ma = g["acc"][abi.A_MA].astype(np.float64)
    # At most the 27 cells around the vertex (a share that rounds to 0 in fixed point leaves none).
    assert 0 < np.count_nonzero(mf) <= 27 and 0 < np.count_nonzero(ma) <= 27
    assert abs(mf.sum() - p[0, 0] * FP_F) <= 14.0
    assert abs(ma.sum() - p[0, 1] * 1e6 * FP_A) <= 14.0
    cc = cell_coords(g, s.params.bdims).astype(np.float64)
    base = np.floor((c - np.array(s.params.origin[:3], np.float64)) / h - 1.0)
    span = cc[mf != 0] - base
    assert span.min() >= 0 and span.max() <= 2
    centroid = (cc * mf[:, None]).sum(axis=0) / mf.sum()
    want = (c - np.array(s.params.origin[:3], np.float64)) / h - 0.5
    np.testing.assert_allclose(centroid, want, atol=1e-3)  # fixed-point rounding of 27 shares
    # The own cell: area in mm^2 and area times velocity, in FP_A.
    own = np.floor((c - np.array(s.params.origin[:3])) / h).astype(int)
    k = np.nonzero((cc == own).all(axis=1))[0]
    assert len(k) == 1
    a = np.float32(p[0, 1]) * np.float32(1e6)
    assert g["acc"][abi.A_MW][k[0]] == to_fp(a, FP_A)
    for d in range(3):
        assert g["acc"][abi.A_MV + d][k[0]] == to_fp(np.float32(a) * v[0, d], FP_A)
    assert np.count_nonzero(g["acc"][abi.A_MW]) == 1


def test_mark_rubber_claims_blocks_off_the_water():
    s = ball()
    h = float(s.params.origin[3])
    pos = np.array([0.5, 0.5, 0.5]) + np.array([0.3, 0.0, 0.0])  # well outside the ball
    x, v, m, p = one_vertex(pos)
    o = oracle.Oracle.from_scene(s, "f32")
    ref = oracle.Oracle.from_scene(s, "f32")
    o.set_membrane(x, v, m, p)
    run_to(o, abi.PHASE_BIN)
    run_to(ref, abi.PHASE_BIN)
    base = set(ref.grid()["block_list"].tolist())
    got = set(o.grid()["block_list"].tolist())
    c = np.floor(pos / h).astype(int)
    bmax = np.array(s.params.bdims[:3]) - 1
    lo = np.clip((c - 2) >> 2, 0, bmax)
    hi = np.clip((c + 3) >> 2, 0, bmax)
    bd = s.params.bdims
    want = {(bx * bd[1] + by) * bd[2] + bz for bx in range(lo[0], hi[0] + 1) for by in range(lo[1], hi[1] + 1)
            for bz in range(lo[2], hi[2] + 1)}
    assert got == base | want
    assert want - base  # the vertex claims blocks the water does not
    # A vertex without area claims nothing.
    x2, v2, m2, p2 = one_vertex(pos, push=(3e-3, 0.0))
    o2 = oracle.Oracle.from_scene(s, "f32")
    o2.set_membrane(x2, v2, m2, p2)
    run_to(o2, abi.PHASE_BIN)
    assert set(o2.grid()["block_list"].tolist()) == set(ref.grid()["block_list"].tolist())


def test_rubber_and_covered_cells():
    s = ball()
    x, v, m, p = scenes.membrane_shell(s, n_verts=600, gap_cells=0.6)
    o = oracle.Oracle.from_scene(s, "f32")
    o.set_membrane(x, v, m, p)
    o.run(2, 200)
    run_to(o, abi.PHASE_GRID)
    g = o.grid()
    ty = g["cf"][abi.F_TYPE]
    acc = g["acc"]
    rubber = ty == abi.RUBBER
    covered = ty == abi.COVERED
    assert rubber.any() and covered.any()
    assert (acc[abi.A_MW][rubber] > 0).all()
    assert (acc[abi.A_MA][covered] > 0).all()
    # A covered cell holds the rubber's pressure: push / area (MPa back to Pa).
    mf = acc[abi.A_MF][covered].astype(np.float32) / np.float32(FP_F)
    ma = acc[abi.A_MA][covered].astype(np.float32) / np.float32(FP_A)
    want = (mf / ma * np.float32(1e6)).astype(np.float32)
    np.testing.assert_array_equal(g["cf"][abi.F_P][covered], want)
    # No water cell and no solid cell is retyped.
    assert not (rubber & (acc[abi.A_WC] > 0) & (ty == abi.WATER)).any()


def test_sample_matches_numpy_restatement():
    s = ball()
    x, v, m, p = scenes.membrane_shell(s, n_verts=300, gap_cells=-1.5, extra_off_water=2)
    o = oracle.Oracle.from_scene(s, "f64")
    o.set_membrane(x, v, m, p)
    o.run(2, 200)
    water = o.membrane_water()
    g = o.grid()
    cf = o.cf_f64()
    acc = g["acc"].astype(np.float64)
    h = float(s.params.origin[3])
    cc = cell_coords(g, s.params.bdims)
    index = {tuple(c): i for i, c in enumerate(cc.tolist())}

    def weights(fx):
        return np.array([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])

    def sample(xv, fu_row, fw_row):
        base = np.floor(xv - 0.5).astype(int)
        w = [weights(xv[k] - base[k]) for k in range(3)]
        sx = sy = 0.0
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    i = index.get((base[0] + a, base[1] + b, base[2] + c))
                    if i is None:
                        continue
                    wt = w[0][a] * w[1][b] * w[2][c] * fw_row[i] / FP_W
                    sx += wt * fu_row[i]
                    sy += wt
        return sx, sy

    for i in range(len(m)):
        xv = (x[i, :3].astype(np.float64) - np.array(s.params.origin[:3])) / h
        for d in range(3):
            xs = xv - np.where(np.arange(3) == d, 0.0, 0.5)
            sx, sy = sample(xs, cf[abi.F_U + d], acc[abi.a_wt(d)])
            want = sx / sy if sy > 1e-6 else 0.0
            assert abs(water[i, d] - want) <= 1e-5 * max(1.0, abs(want)), (i, d, water[i, d], want)
        _, sy = sample(xv - 0.5, cf[abi.F_TYPE], acc[abi.A_WC])
        assert abs(water[i, 3] - sy / 8.0) <= 1e-6
    # Deep in the water the cells are full; the two vertices on the ring outside see less.
    assert np.median(water[:300, 3]) > 0.8
    assert (water[300:, 3] < 0.5).all()
