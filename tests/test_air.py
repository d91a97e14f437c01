"""The air boundary on the GPU against the oracle, bit for bit: gas carried between steps as
blocks are renumbered, cells that stop being water filled from their neighbours or with
vapour, and the GAS and RUBBER pressures the solve takes as given (with the membrane shell, so
that RUBBER cells exist)."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_gpu_parity import assert_same_float, compare_grid, compare_state
from tests.test_oracle_air import dry_region, impact_with_air

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_air_boundary_bit_exact():
    s = impact_with_air(side=64, n=32768)
    say("[air] impact 32k with the air boundary and a membrane shell")
    x, v, m, p = scenes.membrane_shell(s, n_verts=1500)
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    sim.set_membrane(x, v, m, p)
    orc.set_membrane(x, v, m, p)
    rubber = 0
    for step in range(5):
        for ph in range(abi.N_PHASES):
            say(f"[air] step {step} phase {abi.PHASE_NAMES[ph]}")
            sim.phase(ph)
            orc.phase(ph)
            tag = f"air step {step} after {abi.PHASE_NAMES[ph]}"
            compare_state(tag, sim.state(), orc.state())
            if ph >= abi.PHASE_P2G:
                compare_grid(tag, sim.grid(), orc.grid())
                assert_same_float(f"{tag} air", sim.air(), orc.air())
            if ph == abi.PHASE_GRID:
                rubber += int((sim.grid()["cf"][abi.F_TYPE] == abi.RUBBER).sum())
            if ph == abi.PHASE_G2P:
                assert_same_float(f"{tag} particles", sim.particles(), orc.particles())
    assert rubber > 0
    sim.close()
    orc.close()


def test_vapour_fill_bit_exact():
    s = impact_with_air()
    say("[air] sealed cells fill with vapour")
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    sim.run(2, iters=400)
    orc.run(2, 400)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G):
        sim.phase(ph)
        orc.phase(ph)
    g = orc.grid()
    side = int(s.params.dims[0])
    idx = dry_region(g, s.params.bdims, (side // 2, side // 2, side // 2))
    acc = g["acc"].copy()
    acc[abi.A_N, idx] = 0
    acc[abi.A_WC, idx] = 0
    sim.write_grid(acc, None)
    orc.write_grid(acc, None)
    for ph in (abi.PHASE_GRID, abi.PHASE_SOLVE, abi.PHASE_PROJECT, abi.PHASE_EXTEND, abi.PHASE_G2P):
        sim.phase(ph)
        orc.phase(ph)
        compare_grid(f"vapour after {abi.PHASE_NAMES[ph]}", sim.grid(), orc.grid())
        assert_same_float(f"vapour air after {abi.PHASE_NAMES[ph]}", sim.air(), orc.air())
    assert_same_float("vapour particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()
