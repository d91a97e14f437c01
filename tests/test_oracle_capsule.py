"""The oracle's scene extensions for the balloon impact against closed forms: the moving capsule
solid (cells inside it SOLID, `bounded` with the capsule's velocity on its faces), its `carve`
in G2P, and the fixed spherical COVERED mask.

These stand in for the reference's patch bullet (shaders/bullet.wgsl:395-424) and take the
place of the ground in the face rule `bounded` (shaders/fluid.wgsl:360-365):

  - a cell is SOLID when its centre lies closer than r to the capsule's axis segment, which
    moves as a + v t, b + v t with t the step's clock (checked against float64 geometry);
  - `bounded` on a face whose -x (-y, -z) cell is inside the capsule gives max(u, v_capsule),
    on a face whose own cell is inside gives min(u, v_capsule), between two solid cells 0;
  - `carve`: a particle that ends a step inside the capsule is put on its surface, at the
    point nearest to where it ended (on the ray from the axis through it), and its velocity
    along that ray is raised to the capsule's if it was lower; tangential velocity is kept;
    a particle on the axis itself goes out along cross(b - a, y);
  - mask: a cell that is neither WATER nor SOLID, with its centre inside the mask's sphere,
    is COVERED at the mask's pressure; other gas cells keep p = 0.

Each check fails for a plausible mistake: 0 in place of the capsule's velocity, hi/lo
swapped, the capsule taken at the start of the step instead of its clock, the push's sign,
the surface point measured from the particle instead of the axis, corner instead of centre.
"""
import numpy as np

import oracle
from natray_amd import abi
from tests.helpers import box_scene, cell_coords, records

SIDE = 32
H = 1.0 / SIDE
A0 = np.array([0.35, 0.5, 0.5])
B0 = np.array([0.55, 0.5, 0.5])
R = 0.1  # 3.2 cells
VCAP = np.array([300.0, 0.0, 40.0])
DT = 1e-4


def f32(x):
    return np.float32(x)


def now_of(o):
    return float(f32(o.state().t_ns) * f32(1e-9))


def lattice(lo, hi):
    """8 particles per cell (a 2 x 2 x 2 lattice) over the cells lo..hi-1 of each axis."""
    g = [np.arange(2 * lo[k], 2 * hi[k]) + 0.5 for k in range(3)]
    X, Y, Z = np.meshgrid(*g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) * (H / 2)


def capsule_scene(extra=None, capsule=True, mask=False, mask_p=0.0, gravity=9.81, sphere=(0.5, 0.75, 0.5, 0.12)):
    P = lattice((8, 9, 9), (24, 23, 23))
    if extra is not None:
        P = np.concatenate([P, np.asarray(extra, np.float64).reshape(-1, 3)])
    s = box_scene(records(P), side=SIDE, fixed_dt=DT, flags=0, gravity=gravity)
    sc = s.scene
    sc.flags = (abi.SCENE_CAPSULE if capsule else 0) | (abi.SCENE_MASK if mask else 0)
    sc.capsule_a[:] = [A0[0], A0[1], A0[2], R]
    sc.capsule_b[:] = [B0[0], B0[1], B0[2], 0.0]
    sc.capsule_v[:] = [VCAP[0], VCAP[1], VCAP[2], 0.0]
    sc.mask[:] = list(sphere)
    sc.mask_p = mask_p
    return s


def axis_distance(p, t):
    """Distance of points p (n, 3) to the capsule's axis segment at time t, float64."""
    a, b = A0 + VCAP * t, B0 + VCAP * t
    ab = b - a
    f = np.clip(((p - a) @ ab) / (ab @ ab), 0.0, 1.0)
    on = a + f[:, None] * ab
    return np.linalg.norm(p - on, axis=1), on


def test_capsule_cells_are_solid():
    s = capsule_scene()
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
        o.phase(ph)
    g = o.grid()
    t = now_of(o)
    assert t > 0
    cc = cell_coords(g, s.params.bdims)
    d, _ = axis_distance((cc + 0.5) * H, t)
    ty = g["cf"][abi.F_TYPE]
    clear = np.abs(d - R) > 1e-4
    inside = (d < R) & clear
    assert inside.sum() > 100
    assert np.all(ty[inside] == abi.SOLID)
    assert not np.any(ty[(d > R) & clear] == abi.SOLID)
    # At t = 0 the capsule would be 0.03 (about a cell) further back: some cells differ.
    d0, _ = axis_distance((cc + 0.5) * H, 0.0)
    assert np.any((d0 < R) != (d < R))
    o.close()


def test_bounded_takes_the_capsule_velocity():
    s = capsule_scene()
    o = oracle.Oracle.from_scene(s)
    for ph in range(abi.PHASE_PROJECT + 1):
        o.phase(ph)
    g = o.grid()
    t = now_of(o)
    cc = cell_coords(g, s.params.bdims)
    index = {tuple(int(v) for v in c): i for i, c in enumerate(cc)}
    d, _ = axis_distance((cc + 0.5) * H, t)
    solid = d < R
    cf = g["cf"].copy()
    # No water, so the extension changes nothing and only `bounded` acts.
    cf[abi.F_TYPE][:] = np.where(solid, abi.SOLID, abi.GAS)
    cf[abi.F_U:abi.F_U + 3] = 0.0
    # The axis at t runs along y = 16, z = 16.13 (in cells); the cases sit at the edges of the
    # cells inside it, found from the float64 geometry.
    ins = lambda c: bool(solid[index[c]])
    x_in = [x for x in range(8, 24) if ins((x, 15, 16))]
    x_lo, x_hi = min(x_in), max(x_in)
    xm = (x_lo + x_hi) // 2
    assert x_hi - x_lo >= 6
    y_in = [y for y in range(8, 24) if ins((xm, y, 16))]
    z_in = [z for z in range(8, 24) if ins((xm, 16, z))]
    # (cell, axis, where the capsule is: "lo" = only the -axis cell, "hi" = only the cell itself,
    # "both", u before, u after), the results worked out from the rules in the module's text.
    cases = [
        ((x_hi + 1, 15, 16), 0, "lo", 1.0, 300.0),    # max(1, 300)
        ((x_hi + 1, 16, 16), 0, "lo", 500.0, 500.0),  # max(500, 300)
        ((x_lo, 15, 16), 0, "hi", 500.0, 300.0),      # min(500, 300)
        ((x_lo, 16, 16), 0, "hi", -7.0, -7.0),        # min(-7, 300)
        ((x_lo + 2, 15, 16), 0, "both", 123.0, 0.0),
        ((xm, max(y_in) + 1, 16), 1, "lo", -5.0, 0.0),  # max(-5, 0): the capsule's v_y is 0
        ((xm, max(y_in) + 1, 15), 1, "lo", 2.5, 2.5),
        ((xm, min(y_in), 16), 1, "hi", 3.0, 0.0),       # min(3, 0)
        ((xm, 16, max(z_in) + 1), 2, "lo", 10.0, 40.0),  # max(10, 40)
        ((xm, 15, max(z_in) + 1), 2, "lo", 55.0, 55.0),
        ((xm, 16, min(z_in)), 2, "hi", 70.0, 40.0),      # min(70, 40)
    ]
    for c, ax, where, before, _ in cases:
        lo = list(c)
        lo[ax] -= 1
        assert (ins(tuple(lo)), ins(c)) == {"lo": (True, False), "hi": (False, True), "both": (True, True)}[where], (c, ax, where)
        cf[abi.F_U + ax][index[c]] = before
    o.write_grid(None, cf)
    o.phase(abi.PHASE_EXTEND)
    out = o.grid()["cf"]
    for c, ax, where, before, after in cases:
        got = float(out[abi.F_U + ax][index[c]])
        assert got == after, (c, ax, where, before, after, got)
    o.close()


def test_carve_closed_form():
    """G2P of the same grid with and without the capsule: particles that end outside it are
    identical; those that end inside are moved to the surface point and pushed as the closed
    form says; a particle exactly on the axis goes out along cross(b - a, y) = +z."""
    t_step = float(f32(100000) * f32(1e-9))
    # A particle on the axis at the clock of G2P: gathered velocity 0 (every face around it lies
    # between solid cells and `bounded` zeroes it), so it ends where it started.
    axis_pt = [0.45, 0.5, float(f32(0.5) + f32(40.0) * f32(t_step))]
    sa = capsule_scene(extra=[axis_pt])
    sb = capsule_scene(extra=[axis_pt], capsule=False)
    oa, ob = oracle.Oracle.from_scene(sa), oracle.Oracle.from_scene(sb)
    for ph in range(abi.PHASE_G2P):
        oa.phase(ph)
        ob.phase(ph)
    ga = oa.grid()
    gb = ob.grid()
    np.testing.assert_array_equal(ga["block_list"], gb["block_list"])
    ob.write_grid(ga["acc"], ga["cf"])
    assert now_of(oa) == t_step
    oa.phase(abi.PHASE_G2P)
    ob.phase(abi.PHASE_G2P)
    Pa, Pb = oa.particles().astype(np.float64), ob.particles().astype(np.float64)
    xb, vb = Pb[:, :3], Pb[:, 4:7]
    d, on = axis_distance(xb, t_step)
    out = d > R + 1e-5
    np.testing.assert_array_equal(Pa[out], Pb[out])
    inn = (d < R - 1e-5) & (d > 1e-5)
    assert inn.sum() > 200
    dirn = (xb[inn] - on[inn]) / d[inn, None]
    want_x = on[inn] + R * dirn
    vn = np.sum(vb[inn] * dirn, axis=1)
    wall = dirn @ VCAP
    want_v = vb[inn] + dirn * np.maximum(wall - vn, 0.0)[:, None]
    np.testing.assert_allclose(Pa[inn, :3], want_x, rtol=0, atol=2e-6)
    np.testing.assert_allclose(Pa[inn, 4:7], want_v, rtol=1e-5, atol=1e-3)
    # Pushed particles now move outwards at least as fast as the wall.
    got_vn = np.sum(Pa[inn, 4:7] * dirn, axis=1)
    assert np.all(got_vn >= wall - 1e-3 * np.maximum(1.0, np.abs(wall)))
    assert np.any(wall - vn > 1.0)  # some were pushed
    # Alive, and everything but position and velocity equal.
    assert np.all(Pa[inn, 7] == 1.0)
    np.testing.assert_array_equal(Pa[inn, 8:], Pb[inn, 8:])
    # The particle on the axis.
    k = Pa.shape[0] - 1
    np.testing.assert_array_equal(Pb[k, :3], np.asarray(axis_pt, np.float32).astype(np.float64))
    np.testing.assert_array_equal(Pb[k, 4:7], 0.0)
    np.testing.assert_allclose(Pa[k, :3], [0.45, 0.5, axis_pt[2] + R], rtol=0, atol=2e-7)
    np.testing.assert_allclose(Pa[k, 4:7], [0.0, 0.0, 40.0], rtol=0, atol=1e-5)
    oa.close()
    ob.close()


def test_mask_cells_are_covered():
    mask_p = 1234.5
    s = capsule_scene(capsule=False, mask=True, mask_p=mask_p)
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
        o.phase(ph)
    g = o.grid()
    cc = cell_coords(g, s.params.bdims)
    centre, radius = np.array([0.5, 0.75, 0.5]), 0.12  # the mask straddles the water's top
    r = np.linalg.norm((cc + 0.5) * H - centre, axis=1)
    ty, p = g["cf"][abi.F_TYPE], g["cf"][abi.F_P]
    clear = np.abs(r - radius) > 1e-5
    water = g["acc"][abi.A_N] > 0
    inside = (r < radius) & clear & ~water
    outside = (r > radius) & clear & ~water
    assert inside.sum() > 50 and outside.sum() > 50
    assert np.all(ty[inside] == abi.COVERED) and np.all(p[inside] == np.float32(mask_p))
    assert np.all(ty[outside] == abi.GAS) and np.all(p[outside] == 0.0)
    assert np.all(ty[water] == abi.WATER)
    # Measured from the corner instead of the centre, some cells would change side.
    rc = np.linalg.norm(cc * H - centre, axis=1)
    assert np.any((rc < radius) != (r < radius))
    o.close()
