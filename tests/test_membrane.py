"""The membrane coupling on the GPU against the oracle, phase by phase and bit for bit: blocks
claimed around rubber off the water, the rubber's area, motion, push and area splatted to the
grid, the RUBBER and COVERED cells and their pressures that the solve takes as given, and the
water sampled at every vertex. A shell of 2000 vertices just outside the water ball, and 8 off
the water, moved between steps."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_gpu_parity import SCENES, assert_same_float, compare_grid, compare_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["impact_sgs_32k", "impact_mixed_32k_mem"])
def test_membrane_coupling_bit_exact(name):
    s = SCENES[name]()
    say(f"[membrane] {name}")
    x, v, m, p = scenes.membrane_shell(s)
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    rubber = covered = 0
    for step in range(4):
        xs = x.copy()
        xs[:, :3] += (step * 1e-4) * v[:, :3]  # the membrane's own step moved it
        sim.set_membrane(xs, v, m, p)
        orc.set_membrane(xs, v, m, p)
        for ph in range(abi.N_PHASES):
            say(f"[membrane] {name} step {step} phase {abi.PHASE_NAMES[ph]}")
            sim.phase(ph)
            orc.phase(ph)
            tag = f"{name} step {step} after {abi.PHASE_NAMES[ph]}"
            compare_state(tag, sim.state(), orc.state())
            if ph >= abi.PHASE_BIN:
                g, o = sim.grid(), orc.grid()
                compare_grid(tag, g, o, cf_fields=() if ph == abi.PHASE_BIN else range(abi.CELL_FIELDS),
                             acc=ph != abi.PHASE_BIN)
                if ph == abi.PHASE_GRID:
                    rubber += int((g["cf"][abi.F_TYPE] == abi.RUBBER).sum())
                    covered += int((g["cf"][abi.F_TYPE] == abi.COVERED).sum())
            if ph == abi.PHASE_EXTEND:
                assert_same_float(f"{tag} membrane water", sim.membrane_water(), orc.membrane_water())
            if ph == abi.PHASE_G2P:
                assert_same_float(f"{tag} particles", sim.particles(), orc.particles())
    say(f"[membrane] {name}: {rubber} rubber and {covered} covered cell-steps")
    assert rubber > 0 and covered > 0
    w = sim.membrane_water()
    assert (w[:2000, 3] > 0.05).mean() > 0.5  # most of the shell is on the water
    sim.close()
    orc.close()


def test_membrane_off_changes_nothing():
    """n = 0 vertices gives the same bits as a scene that never had a membrane."""
    s = SCENES["impact_sgs_32k"]()
    say("[membrane] n = 0")
    a = Sim.from_scene(s)
    b = Sim.from_scene(s)
    e = np.zeros((0, 4), np.float32)
    b.set_membrane(e, e, np.zeros(0, np.float32), np.zeros((0, 2), np.float32))
    a.run(4, iters=400)
    b.run(4, iters=400)
    compare_state("n = 0", a.state(), b.state())
    assert_same_float("n = 0 particles", a.particles(), b.particles())
    a.close()
    b.close()
