"""Host side, no GPU: the C-ABI library loads and exports every entry point the header
declares with byte-identical layouts; the reference's host rules (batch sizing and CG budget
adaptation, src/sim.rs:498-522,636-645); argument checking that fails before any GPU call;
the SimParams / presets / set_kv mirror of src/params.rs and src/presets.rs; and the scene
builders."""
import ctypes as C
import math

import numpy as np
import pytest

from natray_amd import abi, params, scenes


def test_library_exports_every_declared_symbol():
    L = abi.lib()
    names = abi.exported_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_layouts_are_byte_identical():
    assert abi.abi_sizes() == (192, 96, 80, C.sizeof(abi.Scene))
    assert C.sizeof(abi.Params) == 192 and C.sizeof(abi.State) == 96
    # Word 16 of State is the spray count the keyframe packer reads (sim_defs.wgsl:23).
    assert abi.State.spray_count.offset == 16 * 4
    # iters_max / unconverged are words 13-14, which Sim::run resets (src/sim.rs:593).
    assert abi.State.iters_max.offset == 13 * 4 and abi.State.unconverged.offset == 14 * 4
    assert abi.Params.limits.offset == 6 * 16 and abi.Params.knot.offset == 11 * 16


def test_create_rejects_bad_arguments_before_touching_the_gpu():
    s = scenes.water_sphere(side=16, n=8 * 20)
    P = s.particles
    h = C.c_void_p()
    U = s.params
    U.dims[3] = 7  # != particle count
    rc = abi.lib().natray_sim_create(C.byref(U), abi.ptr(P), P.shape[0], None, None, C.byref(h))
    assert rc == abi.E_ARG and b"particle count" in abi.lib().natray_last_error()
    U.dims[3] = P.shape[0]
    U.bdims[0] = 3  # != dims / 4
    rc = abi.lib().natray_sim_create(C.byref(U), abi.ptr(P), P.shape[0], None, None, C.byref(h))
    assert rc == abi.E_ARG
    assert abi.lib().natray_sim_create(None, None, 0, None, None, C.byref(h)) == abi.E_ARG


@pytest.mark.parametrize(
    "stop,t,free,step_ms,left,expect",
    [
        (1_000_000, 0, 0, 0.0, 1000, 1),           # first batch: one step (no timing yet)
        (1_000_000, 0, 0, 0.5, 1000, 1),           # DT_MAX steps: 1 ms needs one
        (1_000_000, 0, 10_000, 0.5, 1000, 8),      # 100 steps wanted, 4 ms / 0.5 ms fit
        (1_000_000, 0, 10_000, 0.01, 1000, 100),   # cheap steps: as many as needed
        (10_000_000, 0, 10_000, 0.001, 1000, 256),  # capped at MAX_BATCH
        (10_000_000, 0, 10_000, 0.001, 30, 30),    # and at the budget left
        (1_000_500, 1_000_000, 300, 0.001, 9, 2),  # rounds up to reach the stop
        (1_000_000, 1_000_000, 300, 0.001, 9, 0),  # nothing left to do
    ],
)
def test_batch_sizing(stop, t, free, step_ms, left, expect):
    assert abi.lib().natray_plan_batch(stop, t, free, step_ms, left) == expect


@pytest.mark.parametrize(
    "iters,iters_max,unconverged,expect",
    [(64, 0, 0, 64), (64, 40, 0, 52), (64, 3, 0, 8), (64, 5000, 0, 4096), (64, 64, 1, 128), (3000, 10, 2, 4096)],
)
def test_cg_budget_adaptation(iters, iters_max, unconverged, expect):
    assert abi.lib().natray_next_iters(iters, iters_max, unconverged) == expect


# ---- SimParams, presets, set_kv (src/params.rs, src/presets.rs) ------------------------------


def test_presets_fit_the_shot_limit():
    """src/presets.rs:123-128"""
    ps = params.all_presets()
    assert len(ps) == 8
    for p in ps:
        assert 1 <= len(p.sim.shots) <= params.MAX_SHOTS, p.name
    assert params.named("weightless").sim.gravity == 0.0
    with pytest.raises(params.ParamError, match="no preset"):
        params.named("nope")


def test_bullet_spin_and_defaults():
    assert params.BB.spin_rpm() == 0.0
    assert math.isclose(params.LR22.spin_rpm(), 330.0 / (16.0 * 0.0254) * 60.0)
    d = params.SimParams()
    assert d.resolution == 64 and d.balloon_diameter_cm == 16.0 and d.shots[0].bullet == params.LR22


def test_set_kv():
    sim, look = params.SimParams(), params.LookParams()
    params.set_kv(sim, look, "shot3.bullet", "3")
    assert len(sim.shots) == 3 and sim.shots[2].bullet == params.M855 and sim.shots[2].speed == 940.0
    # Added shots come a quarter turn further round each (src/params.rs:134-140).
    assert sim.shots[1].azimuth_deg == 90.0 and sim.shots[2].azimuth_deg == -180.0
    params.set_kv(sim, look, "shots", "1")
    assert len(sim.shots) == 1
    params.set_kv(sim, look, "gravity", "1.5")
    params.set_kv(sim, look, "resolution", "32")
    params.set_kv(sim, look, "dof", "false")
    params.set_kv(sim, look, "color", "0.1,0.2,0.3")
    assert sim.gravity == 1.5 and sim.resolution == 32 and look.dof is False and look.balloon_color == (0.1, 0.2, 0.3)
    for key, v, msg in [("shot0.speed", "1", "count from 1"), ("shot5.speed", "1", "at most"), ("shots", "0", "shots 1"),
                        ("bullet", "7", "bullet 0"), ("gravity", "x", "bad value"), ("nope", "1", "unknown"),
                        ("color", "1,2", "r,g,b"), ("dof", "yes", "bad value")]:
        with pytest.raises(params.ParamError, match=msg):
            params.set_kv(sim, look, key, v)


# ---- scenes ----------------------------------------------------------------------------------------


def test_splitmix64_known_answer():
    # First output of splitmix64 from state 0 (Steele, Lea, Flood 2014; Vigna's reference code).
    assert int(scenes.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF


def test_sphere_fill_is_the_nearest_cells_eight_per_cell():
    side, n = 16, 8 * 100
    P = scenes.sphere_particles(side, n, seed=3)
    assert P.shape == (n, 20) and np.all(P[:, 7] == 1.0) and np.all(P[:, 3] == 1.0)
    cells = np.floor(P[:, :3].astype(np.float64) * side).astype(int)
    uniq, counts = np.unique(cells, axis=0, return_counts=True)
    assert len(uniq) == n // 8 and np.all(counts == 8)
    # Brute force: the 100 cells of 16^3 nearest the centre, ties by linear index.
    allc = np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    d2 = ((allc + 0.5 - side / 2) ** 2).sum(1)
    lin = (allc[:, 0] * side + allc[:, 1]) * side + allc[:, 2]
    ref = allc[np.lexsort((lin, d2))[: n // 8]]
    assert {tuple(c) for c in ref.tolist()} == {tuple(c) for c in uniq.tolist()}
    # Same seed, same particles; another seed, other offsets.
    np.testing.assert_array_equal(P, scenes.sphere_particles(side, n, seed=3))
    assert not np.array_equal(P, scenes.sphere_particles(side, n, seed=4))


def test_config_scene_sizes():
    s = scenes.water_sphere()
    assert s.n == 32768 and tuple(s.params.dims[:3]) == (64, 64, 64) and s.params.origin[3] == np.float32(1 / 64)
    assert abs(s.meta["r_eff"] - 0.1551) < 1e-3
    b = scenes.balloon_impact(side=64, n=32768)
    assert b.scene.flags & abi.SCENE_CAPSULE and b.scene.capsule_v[0] == 300.0 and b.scene.capsule_a[3] == np.float32(0.01)
    # The capsule's nose starts two cells short of the water.
    nose = b.scene.capsule_b[0] + b.scene.capsule_a[3]
    assert abs((0.5 - b.meta["r_eff"]) - nose - 2 / 64) < 1e-6


def test_reference_balloon_scene():
    """Sim::new's domain layout (src/sim.rs:213-225) and the lattice water fill
    (src/balloon.rs:419-445): dx = 2 spacing, sides multiples of 4, particles about 8 per
    cell inside the balloon, jitter within 0.15 spacing of the offset lattice."""
    p = params.SimParams(resolution=16)
    s = scenes.reference_scene(p)
    sp = s.meta["spacing"]
    assert math.isclose(s.meta["dx"], 2 * sp)
    U = s.params
    assert all(int(U.dims[k]) % 4 == 0 and int(U.bdims[k]) * 4 == int(U.dims[k]) for k in range(3))
    assert U.origin[1] == np.float32(-3.0 * 2 * sp)
    n = s.n
    r = 0.5 * p.balloon_diameter_cm * 0.01
    assert 0.5 < n / (4.0 / 3.0 * math.pi * r**3 / sp**3) < 1.05
    P = s.particles
    bottom = p.balloon_height - r
    iy = np.round((P[:, 1] - bottom) / sp - 0.5).astype(int)
    off = 0.5 * (iy % 2)
    jx = P[:, 0] / sp - off - np.round(P[:, 0] / sp - off)
    assert np.all(np.abs(jx) <= 0.15 + 1e-4)
    assert np.all(P[:, 7] == 1.0)
