"""The oracle's spray and wet ground (fluid.wgsl:676-765, sim_defs.wgsl:233-240), with the
elementary functions they use, pinned by what the mathematics and the physics fix:

- exp, log, cos and pow against the C library in double, to half an ulp of float on samples
  over the ranges the spray uses;
- break_up: four drops whose volume (pi/6 d^3 each) is the blob's, hs^3, for slow blobs; for
  fast ones (the fine branch) the drops and the parcel (its weight times a drop of its
  diameter) together hold hs^3; the parcel at the blob's position and velocity; the drops
  within half a particle spacing of it and no faster than the spread allows;
- spray_update (float64): the drag is isotropic (a drop's horizontal direction is kept), a
  drop's radius follows d(r^2)/dt = -EVAPORATION exactly, positions advance by dt v, and a
  landed single drop reports its own radius;
- the wet map holds the time of the step in which ground particles died, once per cell.
"""
import numpy as np
import pytest

import oracle
from natray_amd import abi, scenes

EVAPORATION = 1.3e-10


def ulp_error(got, want):
    want32 = want.astype(np.float32)
    spacing = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


@pytest.mark.parametrize("which", ["exp", "log", "cos", "pow"])
def test_elementary_functions_are_correctly_rounded_on_samples(which):
    rng = np.random.default_rng(4)
    n = 200000
    if which == "exp":
        a = rng.uniform(-80.0, 80.0, n).astype(np.float32)
        want = np.exp(a.astype(np.float64))
        got = oracle.elementary("exp", a)
    elif which == "log":
        a = np.exp(rng.uniform(-80.0, 80.0, n)).astype(np.float32)
        want = np.log(a.astype(np.float64))
        got = oracle.elementary("log", a)
    elif which == "cos":
        a = rng.uniform(-20.0, 20.0, n).astype(np.float32)
        want = np.cos(a.astype(np.float64))
        got = oracle.elementary("cos", a)
    else:
        a = np.exp(rng.uniform(-12.0, 12.0, n)).astype(np.float32)
        b = rng.uniform(-2.0, 2.0, n).astype(np.float32)
        want = np.power(a.astype(np.float64), b.astype(np.float64))
        got = oracle.elementary("pow", a, b)
    assert ulp_error(got, want).max() <= 0.5 + 1e-9


def test_elementary_special_values():
    assert oracle.elementary("exp", np.array([0.0]))[0] == 1.0
    assert oracle.elementary("log", np.array([1.0]))[0] == 0.0
    assert oracle.elementary("cos", np.array([0.0]))[0] == 1.0
    assert oracle.elementary("pow", np.array([8.0]), np.array([1.0 / 3.0]))[0] == np.float32(2.0)
    assert oracle.elementary("exp", np.array([-200.0]))[0] == 0.0
    assert np.isinf(oracle.elementary("exp", np.array([200.0]))[0])


def breakup(kind="f32", steps=1, **kw):
    s = scenes.breakup_scene(**kw)
    o = oracle.Oracle.from_scene(s, kind)
    o.run(steps, 200)
    return s, o


def groups(rec):
    """The spray in groups of five slots (four drops and a parcel), as break_up lays them out."""
    assert len(rec) % 5 == 0
    return rec.reshape(-1, 5, 12)


def test_break_up_conserves_the_blob_volume():
    s, o = breakup("f64", steps=1)
    sp = o.spray()
    hs = float(s.params.fluid[3])
    G = groups(sp["records"])
    assert len(G) >= 60 and sp["count"] == 5 * len(G)
    slow = fine = 0
    for g in G:
        drops, parcel = g[:4], g[4]
        vol_drops = (np.pi / 6.0 * (2.0 * drops[:, 3]) ** 3).sum()
        if parcel[3] > 0:  # the fine branch: drops plus the parcel's drops hold the blob
            fine += 1
            vol = vol_drops + parcel[7] * np.pi / 6.0 * (2.0 * parcel[3]) ** 3
            assert abs(vol - hs ** 3) <= 2e-4 * hs ** 3, (vol, hs ** 3)
        elif parcel[7] == 2.0 and abs(vol_drops - hs ** 3) <= 2e-4 * hs ** 3:
            slow += 1  # all the volume in the four visible drops (0.5236 ~ pi/6)
        # Drops start within half a particle spacing of the blob, made at the same time.
        assert np.all(np.abs(drops[:, 0:3] - parcel[0:3]) <= 0.5 * hs + 1e-12)
        assert np.all(drops[:, 11] == parcel[11]) and parcel[11] > 0
        assert np.all(drops[:, 7] == 1.0)
    assert slow >= 30 and fine >= 10, (slow, fine)


def test_spray_update_drag_evaporation_and_landing():
    s, o = breakup("f64", steps=1)
    before = o.spray()["records"].copy()
    o.run(1, 200)
    after = o.spray()
    rec, landed = after["records"], after["landed"]
    dt = o.state().dt_ns * 1e-9
    moved = 0
    for i in range(len(before)):
        b, a = before[i], rec[i]
        if b[3] <= 0:
            np.testing.assert_array_equal(a, b)  # gone drops and non-drawn parcels stay put
            continue
        moved += 1
        # isotropic drag: the horizontal direction is kept
        np.testing.assert_allclose(a[4] * b[6], a[6] * b[4], rtol=1e-12, atol=1e-15)
        if a[3] > 0:
            np.testing.assert_allclose(a[3] ** 2, b[3] ** 2 - EVAPORATION * dt, rtol=1e-9)
            np.testing.assert_allclose(a[0:3], b[0:3] + dt * a[4:7], rtol=1e-12, atol=1e-15)
    assert moved > 100
    o.run(8, 200)
    sp = o.spray()
    down = np.nonzero(sp["landed"][:, 3] > 0)[0]
    assert len(down) > 50
    single = [i for i in down if sp["records"][i, 7] == 1.0]
    for i in single:
        assert sp["records"][i, 3] == 0.0  # a landed drop is gone
        assert 0 < sp["landed"][i, 2] < 0.01


def test_wet_map_marks_where_ground_particles_died():
    s, o = breakup("f32", steps=1)
    wet = o.wet()
    touched = wet != 0xFFFFFFFF
    t = np.float32(o.state().t_ns) * np.float32(1e-9)
    assert 0 < touched.sum() <= 32
    assert np.all(wet[touched].view(np.float32) == t)
    # ground cells under the ground group: x, z in [0.25, 0.75] of the box map to these cells
    res, hw = int(s.params.limits[0]), float(s.params.wet[0])
    zs, xs = np.nonzero(touched)
    for c in (xs, zs):
        u = (c + 0.5) / res * 2.0 * hw - hw
        assert np.all((u > 0.2) & (u < 0.8))


def test_spray_off_changes_nothing_else():
    """The flag adds spray and the wet map; the particles and the grid are unchanged."""
    s1 = scenes.breakup_scene()
    s2 = scenes.breakup_scene()
    s2.scene.flags &= ~abi.SCENE_SPRAY
    a, b = oracle.Oracle.from_scene(s1, "f32"), oracle.Oracle.from_scene(s2, "f32")
    a.run(4, 200)
    b.run(4, 200)
    np.testing.assert_array_equal(a.particles(), b.particles())
    assert a.state().spray_count > 0 and b.state().spray_count == 0
