"""The oracle's bullet model (bullet.wgsl:85-424, fluid.wgsl:36-50, 481-482, 818), pinned by
what the mathematics and the physics fix:

- sin and asin against the C library in double, to half an ulp of float on samples over the
  ranges the bullet uses, and at the special values;
- the surface: closed (the area-weighted normals cancel, as the reference's own
  surfaces_are_closed test checks, src/bullet.rs:147-164) and of the solid's area;
- free flight away from the water (float64): the centre of mass follows the ballistic path,
  the angular momentum is kept exactly, and the axis cones about it at |L| / Iy, the closed
  form of a torque-free symmetric top; the roll advances at L.axis / Ix; no cavity, no push;
- the clock: a step is cut to end at the arrival, and the bullet joins in the next step;
  one that leaves the domain is gone;
- the pose log holds each step's pose under step % 256, a zero time where the bullet waited;
- in the water: the cavity source the sections splat sums to their volume rate
  (pi d(R^2)/dt |b - a| each, the B-spline weights summing to 1), and bullet_apply adds
  exactly the fixed-point pushes to the face velocities a twin without bullets computes.
"""
import math

import numpy as np
import pytest

import oracle
from natray_amd import abi, scenes
from tests.test_oracle_spray import ulp_error


@pytest.mark.parametrize("which", ["sin", "asin"])
def test_sin_and_asin_are_correctly_rounded_on_samples(which):
    rng = np.random.default_rng(8)
    n = 200000
    if which == "sin":
        a = rng.uniform(-20.0, 20.0, n).astype(np.float32)
        want = np.sin(a.astype(np.float64))
    else:
        a = rng.uniform(-1.0, 1.0, n).astype(np.float32)
        want = np.arcsin(a.astype(np.float64))
    got = oracle.elementary(which, a)
    assert ulp_error(got, want).max() <= 0.5 + 1e-9


def test_sin_and_asin_special_values():
    assert oracle.elementary("sin", np.array([0.0]))[0] == 0.0
    assert oracle.elementary("asin", np.array([0.0]))[0] == 0.0
    assert oracle.elementary("asin", np.array([1.0]))[0] == np.float32(math.pi / 2)
    assert oracle.elementary("asin", np.array([-1.0]))[0] == np.float32(-math.pi / 2)
    assert oracle.elementary("asin", np.array([0.5]))[0] == np.float32(math.pi / 6)
    assert np.isnan(oracle.elementary("asin", np.array([1.5]))[0])


def test_surface_is_closed_and_has_the_solids_area():
    length, radius, nose = 0.02, 0.0045, 0.008
    pt, count, hull = scenes.bullet_surface(length, radius, nose, 0.45 * length)
    assert 0 < count <= abi.BULLET_PATCHES
    p = pt[:count].astype(np.float64)
    area = p[:, 3].sum()
    net = (p[:, 4:7] * p[:, 3:4]).sum(axis=0)
    assert np.linalg.norm(net) < 1e-3 * area
    want = 2 * math.pi * radius * (length - nose) + math.pi * radius * math.hypot(radius, nose) + math.pi * radius**2
    assert abs(area - want) < 1e-3 * want
    assert np.allclose(np.linalg.norm(p[:, 4:7], axis=1), 1.0, atol=1e-6)
    assert np.all(pt[count:] == 0)
    u_cg = 0.45 * length  # the farthest points: the base rim, the shoulder, the tip
    far = max(math.hypot(u_cg, radius), math.hypot(length - nose - u_cg, radius), length - u_cg)
    assert hull == pytest.approx(far, rel=1e-6)


def rodrigues(v, k, theta):
    k = k / np.linalg.norm(k)
    return v * math.cos(theta) + np.cross(k, v) * math.sin(theta) + k * np.dot(k, v) * (1 - math.cos(theta))


def qrot(q, v):
    qv, w = np.asarray(q[:3], np.float64), float(q[3])
    t = 2.0 * np.cross(qv, v)
    return v + w * t + np.cross(qv, t)


def test_free_flight_is_ballistic_and_the_axis_cones_about_L():
    s = scenes.bullet_scene()
    b, pt = scenes.bullet_shot(at=(0.1, 0.85, 0.1), direction=(0.0, 0.3, 1.0), speed=40.0, spin_rev_s=500.0, yaw=0.3)
    o = oracle.Oracle.from_scene(s, "f64")
    o.set_bullets([b], pt[None])
    x0 = np.array(b.x[:3], np.float64)
    v0 = np.array(b.v[:3], np.float64)
    L0 = np.array(b.l[:3], np.float64)
    ix, iy = float(b.shape[0]), float(b.shape[1])
    axis0 = qrot(np.array(b.q, np.float64), np.array([0.0, 0.0, 1.0]))
    g = np.array(s.params.gravity[:3], np.float64)
    t_in = 0.0
    for k in range(12):
        st = o.run(1, 200)
        (B,), sec = o.bullets()
        if k == 0:
            assert B.info[2] == 1  # arrive_ns 0: joins in the first step
        t_in += st.dt_ns * 1e-9
        x = np.array(B.x[:3])
        v = np.array(B.v[:3])
        np.testing.assert_allclose(v, v0 + g * t_in, rtol=0, atol=1e-4)
        # x_{k+1} = x_k + v_{k+1} dt: the path of the symplectic Euler step.
        assert np.linalg.norm(x - (x0 + v0 * t_in + g * t_in**2 / 2)) < 2e-3 * np.linalg.norm(v0) * t_in + 1e-9
        assert list(B.l[:3]) == pytest.approx(list(L0.astype(np.float32)), abs=0)
        axis = qrot(np.array(B.q, np.float64), np.array([0.0, 0.0, 1.0]))
        want = rodrigues(axis0, L0, np.linalg.norm(L0) / iy * t_in)
        np.testing.assert_allclose(axis, want, atol=1e-6)
        assert B.v[3] == pytest.approx(np.dot(L0, axis0) / ix * t_in, rel=1e-5)
        # far from the water the pace is a block per step: |v| CFL / 4
        assert B.l[3] == pytest.approx(np.linalg.norm(v) * float(s.params.fluid[1]) / 4, rel=1e-6)
        assert B.reach[3] == 0 and not (sec[0, :, 7] > 0).any()
    grid = o.grid()
    assert not grid["acc"][abi.A_BU:abi.A_SRC + 1].any()


def test_a_step_ends_at_the_arrival_and_the_bullet_joins_the_next():
    s = scenes.bullet_scene()
    b, pt = scenes.bullet_shot(at=(0.1, 0.9, 0.1), direction=(0.0, 0.0, 1.0), speed=400.0, arrive_ns=7000)
    o = oracle.Oracle.from_scene(s, "f32")
    o.set_bullets([b], pt[None])
    st = o.run(1, 200)
    (B,), _ = o.bullets()
    assert (st.t_ns, st.dt_ns, B.info[2]) == (7000, 7000, 0)
    assert list(B.x[:3]) == list(b.x[:3])
    st = o.run(1, 200)
    (B,), _ = o.bullets()
    assert B.info[2] == 1 and B.x[2] > b.x[2]
    # arrived but not yet in: its speed sets dt, CFL h / |v|
    cfl, h = float(s.params.fluid[1]), 1.0 / 64
    assert st.dt_ns == pytest.approx(cfl * h / 400.0 * 1e9, abs=1)
    first = st.dt_ns
    # in: its pace, CFL h / (|v| CFL / 4) = 4 h / |v| away from the water, held to 1.25x growth
    st = o.run(1, 200)
    assert st.dt_ns == pytest.approx(min(4.0 * h / 400.0 * 1e9, 1.25 * first), abs=1)


def test_a_bullet_that_leaves_the_domain_is_gone():
    s = scenes.bullet_scene()
    b, pt = scenes.bullet_shot(at=(0.1, 0.9, 0.995), direction=(0.0, 0.0, 1.0), speed=400.0)
    o = oracle.Oracle.from_scene(s, "f32")
    o.set_bullets([b], pt[None])
    o.run(2, 200)
    (B,), _ = o.bullets()
    assert B.info[2] == 2
    x = list(B.x)
    o.run(1, 200)
    (B,), _ = o.bullets()
    assert B.info[2] == 2 and list(B.x) == x


def into_water(kind, steps):
    s = scenes.bullet_scene()
    b, pt = scenes.bullet_shot(arrive_ns=30000)
    o = oracle.Oracle.from_scene(s, kind)
    o.set_bullets([b], pt[None])
    o.run(steps, 200)
    return s, b, pt, o


def test_cavity_source_sums_to_the_sections_volume_rate():
    s, b, pt, o = into_water("f64", 14)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
        o.phase(ph)
    st = o.state()
    (B,), sec = o.bullets()
    t = st.t_ns * 1e-9
    hh = float(s.params.origin[3])
    block_ptr, _ = o.bins()
    bd = [int(v) for v in s.params.bdims[:3]]

    def present(c):  # a cell of a block in use this step (the others take no source)
        if min(c) < 0 or max(c) >= 64:
            return False
        slot = ((c[0] >> 2) * bd[1] + (c[1] >> 2)) * bd[2] + (c[2] >> 2)
        return block_ptr[slot] < st.n_blocks

    def weights(fx):  # the quadratic B-spline weights (fluid.wgsl:221-223)
        return [0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2]

    want, thin, whole = 0.0, 0, 0
    for k in range(abi.BULLET_SECTIONS):
        a, e, c = sec[0, k, 0:4], sec[0, k, 4:8], sec[0, k, 8:12]
        r2 = c[0] + e[3] * max(t - a[3], 0.0)
        if e[3] > 0 and r2 < (1.5 * hh) ** 2:
            q = math.pi * e[3] * np.linalg.norm(e[:3] - a[:3])
            x = (0.5 * (a[:3] + e[:3]) - np.array(s.params.origin[:3])) / hh - 0.5
            base = np.floor(x - 0.5).astype(int)
            w = [weights(x[i] - base[i]) for i in range(3)]
            cover = sum(w[0][i] * w[1][j] * w[2][m] for i in range(3) for j in range(3) for m in range(3)
                        if present((base[0] + i, base[1] + j, base[2] + m)))
            want += q * cover
            thin += 1
            whole += abs(cover - 1.0) < 1e-12
    assert thin >= 3 and whole >= 3
    got = o.grid()["acc"][abi.A_SRC].astype(np.float64).sum() / 16777216.0
    assert abs(got - want) <= 27 * thin / 16777216.0 + 1e-9 * want


def test_bullet_apply_adds_the_pushes_to_the_face_velocities():
    sa, b, pt, a = into_water("f32", 12)
    _, _, _, c = into_water("f32", 12)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G):
        a.phase(ph)
        c.phase(ph)
    # the twin loses its bullet just before the grid phase: grid_prep alone
    c.set_bullets([], np.zeros((0, abi.BULLET_PATCHES, 8), np.float32))
    a.phase(abi.PHASE_GRID)
    c.phase(abi.PHASE_GRID)
    ga, gc = a.grid(), c.grid()
    assert ga["n_blocks"] == gc["n_blocks"]
    pushed = 0
    for d in range(3):
        du = ga["acc"][abi.A_BU + d]
        want = gc["cf"][abi.F_U + d].copy()
        nz = du != 0
        want[nz] = (want[nz] + du[nz].astype(np.float32) / np.float32(65536.0)).astype(np.float32)
        np.testing.assert_array_equal(ga["cf"][abi.F_U + d], want)
        pushed += int(nz.sum())
    assert pushed > 50
    assert not gc["acc"][abi.A_BU:abi.A_SRC + 1].any()


def test_the_water_slows_the_bullet_and_the_cavity_opens():
    s, b, pt, o = into_water("f32", 30)
    (B,), sec = o.bullets()
    speed = np.linalg.norm(list(B.v[:3]))
    assert 150.0 < speed < 290.0
    assert B.info[3] >= 10 and B.reach[3] > 0
    made = sec[0][sec[0, :, 7] > 0]
    assert (made[:, 8] > 0).all() and (np.diff(np.sort(made[:, 3])) > 0).all()


def test_the_pose_log_records_each_step_in_the_grid():
    s = scenes.bullet_scene()
    b, pt = scenes.bullet_shot(arrive_ns=30000)
    o = oracle.Oracle.from_scene(s, "f32")
    o.set_bullets([b], pt[None])
    for k in range(6):
        st = o.run(1, 200)
        (B,), _ = o.bullets()
        e = o.pose_log()[0, st.step % abi.BULLET_LOG]
        if B.info[2] == 1:
            assert list(e[:3]) == list(B.x[:3]) and e[3:4].view(np.uint32)[0] == st.t_ns
            assert list(e[4:8]) == list(B.q) and list(e[8:11]) == list(B.v[:3]) and e[11] == B.v[3]
        else:  # waiting for its arrival: the entry's position and time are zero
            assert k == 0 and not e[:4].any()
    log = o.pose_log()[0]
    assert (log[1:st.step + 1, 3].view(np.uint32) > 0).sum() == st.step - 1
