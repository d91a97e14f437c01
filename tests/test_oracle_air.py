"""The oracle's air boundary (shaders/air.wgsl:85-148, 271-277), pinned by what the gas must
be without the air's own substeps: ambient air carried unchanged from step to step as blocks
are renumbered, sealed cavities at the vapour pressure, GAS cells' pressures the gas's gauge
pressure, water and solid cells holding no gas; and that a scene without the flag keeps GAS
cells at gauge 0."""
import numpy as np

import oracle
from natray_amd import abi, scenes
from tests.helpers import cell_coords

GAMMA, P_ATM, P_VAPOUR, RHO_ATM, RHO_VAPOUR = 1.4, 101325.0, 2339.0, 1.204, 0.0278


def impact_with_air(side=32, n=8 * 1200, seed=3):
    s = scenes.balloon_impact(side=side, n=n, seed=seed, iters=400, capacity=(side // 4) ** 3, speed=250.0, mask=False)
    s.scene.flags |= abi.SCENE_AIR
    return s


def by_cell(g, bdims, rows):
    cc = cell_coords(g, bdims)
    return {tuple(c): tuple(r) for c, r in zip(cc.tolist(), np.asarray(rows).T.tolist())}


def gauge(rho, mx, my, mz, e):
    return (np.float32(GAMMA) - np.float32(1.0)) * (e - np.float32(0.5) * ((mx * mx + my * my) + mz * mz) / rho) - np.float32(P_ATM)


def test_air_boundary_closed_forms():
    s = impact_with_air()
    o = oracle.Oracle.from_scene(s, "f32")
    prev = None
    for step in range(12):
        o.run(1, 400)
        g = o.grid()
        air = o.air()
        ty = g["cf"][abi.F_TYPE]
        p = g["cf"][abi.F_P]
        wet = (ty == abi.WATER) | (ty == abi.SOLID)
        assert (air[:, wet] == 0).all()
        gas = ty == abi.GAS
        assert (air[0, gas] > 0).all()
        np.testing.assert_array_equal(p[gas], gauge(*air[:, gas]).astype(np.float32))
        ambient = np.isclose(air[0], np.float32(RHO_ATM)) & gas
        assert (np.abs(p[ambient]) < 0.05).all()
        cur = by_cell(g, s.params.bdims, np.vstack([air, ty[None, :]]))
        if prev is not None:
            # Gas in a cell that stays gas is carried unchanged (no air substeps move it).
            for c, row in cur.items():
                old = prev.get(c)
                if old is not None and old[5] == abi.GAS and row[5] == abi.GAS and old[0] > 0:
                    assert row[:5] == old[:5], (step, c, row, old)
        prev = cur


def dry_region(g, bdims, centre, half=1):
    """Cells of the (2 half + 1)^3 cube around `centre` (cell coordinates), as indices."""
    cc = cell_coords(g, bdims)
    d = np.abs(cc - np.asarray(centre)).max(axis=1)
    return np.nonzero(d <= half)[0]


def test_sealed_cells_fill_with_vapour():
    """Water cells that all stop being water at once, away from any gas, hold vapour: the
    27 cells of a cube in the middle of the ball lose their water (their particle counts zeroed
    after P2G), so none of them has a neighbour with gas to take."""
    s = impact_with_air()
    o = oracle.Oracle.from_scene(s, "f32")
    o.run(2, 400)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G):
        o.phase(ph)
    g = o.grid()
    side = int(s.params.dims[0])
    idx = dry_region(g, s.params.bdims, (side // 2, side // 2, side // 2))
    assert len(idx) == 27
    acc = g["acc"].copy()
    acc[abi.A_N, idx] = 0  # no particle of its own: the cell is not water (grid_prep)
    acc[abi.A_WC, idx] = 0
    o.write_grid(acc, None)
    o.phase(abi.PHASE_GRID)
    g2 = o.grid()
    air = o.air()
    assert (g2["cf"][abi.F_TYPE][idx] == abi.GAS).all()
    np.testing.assert_allclose(air[0, idx], RHO_VAPOUR, rtol=1e-6)
    np.testing.assert_allclose(g2["cf"][abi.F_P][idx], P_VAPOUR - P_ATM, rtol=1e-5)


def test_without_the_flag_gas_is_at_gauge_zero():
    s = impact_with_air()
    s.scene.flags &= ~abi.SCENE_AIR
    o = oracle.Oracle.from_scene(s, "f32")
    o.run(6, 400)
    g = o.grid()
    gas = g["cf"][abi.F_TYPE] == abi.GAS
    assert gas.any()
    np.testing.assert_array_equal(g["cf"][abi.F_P][gas], 0.0)
