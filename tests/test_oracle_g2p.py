"""The oracle's grid update and grid-to-particle transfer against closed forms.

  - grid_prep (shaders/fluid.wgsl:319-356): a lone particle gives u = v + dt g at every face
    it reaches with weight > 1e-6, and cells are typed by the scene's solids;
  - APIC (shaders/fluid.wgsl:780-861): P2G then G2P of an affine velocity field returns the
    same velocity and the same C (affine reproduction), advection is x + dt v;
  - the death rules: an escaped or isolated particle dies with vel.w = 0.
"""
import numpy as np
import pytest

import oracle
from natray_amd import abi
from tests.helpers import box_scene, cell_coords, records


def test_lone_particle_face_velocity_closed_form():
    side = 16
    v = np.array([0.3, -1.2, 2.5], np.float32)
    s = box_scene(records([0.41, 0.52, 0.47], v), side=side, fixed_dt=1e-4)
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
        o.phase(ph)
    g = o.grid()
    dt = np.float32(o.state().dt_ns) * np.float32(1e-9)
    gvec = np.array([0.0, -9.81, 0.0], np.float32)
    for d in range(3):
        w = g["acc"][2 * d + 1] / 2.0**24
        u = g["cf"][d]
        sel = w > 1e-6
        assert sel.sum() >= 8
        np.testing.assert_allclose(u[sel], v[d] + dt * gvec[d], rtol=0, atol=14 / 65536.0 / w[sel].min() * 2 + 1e-6)
        assert np.all(u[~sel] == 0)


def test_cell_types():
    """Water where particles are, walls solid, the rest gas (shaders/fluid.wgsl:321-356)."""
    side = 16
    P = records([[0.5, 0.5, 0.5], [0.53, 0.5, 0.5], [0.04, 0.5, 0.5]])
    s = box_scene(P, side=side)
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
        o.phase(ph)
    g = o.grid()
    cc = cell_coords(g, s.params.bdims)
    ty = g["cf"][abi.F_TYPE]
    for i, c in enumerate(cc.tolist()):
        wall = min(c) == 0 or max(c) == side - 1
        has = g["acc"][abi.A_N][i] > 0
        expect = abi.SOLID if wall else (abi.WATER if has else abi.GAS)
        assert ty[i] == expect, (c, ty[i], expect)
    # The particle in the wall cell x = 0 is in a SOLID cell (solids win over water).
    i0 = [i for i, c in enumerate(cc.tolist()) if c == [0, 8, 8]][0]
    assert g["acc"][abi.A_N][i0] == 1 and ty[i0] == abi.SOLID


def affine_block(side=32, seed=3):
    """A dense lattice of particles (8 per cell) with v(x) = v0 + C (x - x0)."""
    rng = np.random.default_rng(seed)
    h = 1.0 / side
    lo, hi = 10, 22
    cells = np.stack(np.meshgrid(*[np.arange(lo, hi)] * 3, indexing="ij"), -1).reshape(-1, 3)
    pos = (np.repeat(cells, 8, axis=0) + rng.uniform(0.05, 0.95, (len(cells) * 8, 3))) * h
    x0 = np.array([0.5, 0.5, 0.5])
    v0 = np.array([0.4, -0.2, 0.1])
    C = np.array([[1.5, -0.8, 0.3], [0.6, -1.1, 0.9], [-0.4, 0.7, 0.2]])
    vel = v0 + (pos - x0) @ C.T
    P = records(pos, vel, np.broadcast_to(C, (len(pos), 3, 3)))
    return P, v0, C, x0, lo, hi, h


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_apic_affine_reproduction(precision):
    P, v0, C, x0, lo, hi, h = affine_block()
    s = box_scene(P, side=32, fixed_dt=1e-7, gravity=0.0)
    o = oracle.Oracle.from_scene(s, precision)
    # Transfer to the grid and straight back: no solve, no projection.
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID, abi.PHASE_G2P):
        o.phase(ph)
    Q = o.particles_f64()
    dt = float(np.float32(o.state().dt_ns) * np.float32(1e-9))
    # Particles whose whole stencil is inside the filled region.
    c = np.floor(P[:, :3] / h).astype(int)
    inner = np.all((c >= lo + 2) & (c <= hi - 3), axis=1)
    assert inner.sum() > 1000
    vel_ref = P[inner, 4:7].astype(np.float64)
    # Both precisions transfer through the same fixed-point sums (2^-16 m/s units per node).
    tol = 2e-4 if precision == "f32" else 5e-5
    np.testing.assert_allclose(Q[inner, 4:7], vel_ref, atol=tol)
    for d, sl in enumerate((slice(8, 11), slice(12, 15), slice(16, 19))):
        np.testing.assert_allclose(Q[inner, sl], np.broadcast_to(C[d], (inner.sum(), 3)), atol=tol * 200)
    # Advection: x + dt v.
    np.testing.assert_allclose(Q[inner, :3], P[inner, :3].astype(np.float64) + dt * Q[inner, 4:7], atol=1e-7)
    assert np.all(Q[:, 7] == 1.0)


def test_constant_field_has_no_affine_term():
    P, v0, C, x0, lo, hi, h = affine_block()
    P[:, 4:7] = np.array([0.25, 0.5, -0.75], np.float32)
    P[:, 8:19] = 0
    s = box_scene(P, side=32, fixed_dt=1e-7, gravity=0.0)
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID, abi.PHASE_G2P):
        o.phase(ph)
    Q = o.particles()
    c = np.floor(P[:, :3] / h).astype(int)
    inner = np.all((c >= lo + 2) & (c <= hi - 3), axis=1)
    np.testing.assert_allclose(Q[inner][:, [8, 9, 10, 12, 13, 14, 16, 17, 18]], 0.0, atol=2e-3)
    np.testing.assert_allclose(Q[inner][:, 4:7], np.broadcast_to([0.25, 0.5, -0.75], (inner.sum(), 3)), atol=5e-5)


def test_isolated_and_escaped_particles_die():
    """fill < 0.2 away from cover kills a particle, and so does leaving the domain's margin
    (shaders/fluid.wgsl:821-846); a dead particle keeps its position and reports its velocity."""
    side = 16
    P = records([[0.5, 0.5, 0.5], [0.25, 0.3, 0.5]], [[0.0, 0.0, 0.0], [-400.0, 0.0, 0.0]])
    s = box_scene(P, side=side, fixed_dt=1e-4, gravity=0.0)
    o = oracle.Oracle.from_scene(s)
    for ph in range(abi.N_PHASES):
        o.phase(ph)
    Q = o.particles()
    assert np.all(Q[:, 7] == 0.0)
    np.testing.assert_array_equal(Q[:, :3], P[:, :3])
    assert o.state().max_speed == 0.0


def test_ground_clamp_and_carried_pressure():
    """A particle that would sink below 0.3 spacing is put back there with no downward speed
    (shaders/fluid.wgsl:847-853); c0.w carries the weighted water pressure."""
    side = 16
    h = 1.0 / side
    cells = np.stack(np.meshgrid(np.arange(5, 11), np.arange(0, 4), np.arange(5, 11), indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(0)
    pos = (np.repeat(cells, 8, axis=0) + rng.uniform(0.05, 0.95, (len(cells) * 8, 3))) * h
    pos[0] = [8.5 * h, 0.01, 8.5 * h]
    vel = np.zeros_like(pos)
    vel[pos[:, 1] < h, 1] = -30.0  # the bottom layer falls fast
    pres = np.full(len(pos), 1500.0)
    P = records(pos, vel, None, pres)
    s = box_scene(P, side=side, fixed_dt=1e-4, gravity=0.0, flags=0)  # no walls: the ground rule alone
    s.params.limits[2] = 0  # no solve: the carried pressure stays what the particles brought
    o = oracle.Oracle.from_scene(s)
    for ph in range(abi.N_PHASES):
        o.phase(ph)
    Q = o.particles()
    spacing = np.float32(0.5 * h)
    assert Q[0, 7] == 1.0
    assert Q[0, 1] == np.float32(0.3) * spacing and Q[0, 5] == 0.0
    # c0.w = sum w p / sum w over water cells: every particle carries 1500 Pa. The pressure
    # sums are fixed point on MPa with 2^16 units (15.3 Pa each), and node contributions
    # below half a unit round away, so the carried value sits within ~2 units below 1500.
    alive = Q[:, 7] == 1.0
    assert np.all(np.abs(Q[alive, 11] - 1500.0) < 2 * 1e6 / 65536.0)
