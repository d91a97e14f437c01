"""The oracle's velocity extension (extend_1, extend_2; shaders/fluid.wgsl:600-674) on hand-built
face patterns whose results are worked out by hand.

A 16^3 box with a particle at the centre of each 4^3 block has all 64 blocks in use. After the
projection, the test writes its own cell types and x-face velocities into the grid and runs the
extension phase. The patterns:

  - water cells W1 = (8,8,8), W2 = (8,9,9), W3 = (8,10,8). A cell's x face is "set" when the cell
    or its -x neighbour is water, so the set faces are those of W1, W2, W3 and of the cells
    one step +x of them. Every x velocity is u = -(100 x + 10 y + z), an integer, so the means
    below are exact up to the one f32 division;
  - pass 1 (F_EXT, F_EXT_SET): a set face keeps its u; an unset face with 1, 2 or 3 set
    neighbour faces takes their mean; one with none keeps its u and stays unset;
  - pass 2 writes F_U: a face set in pass 1 takes F_EXT; an unset face takes the mean of its
    set neighbours' F_EXT (not their F_U); one with none keeps its u; then `bounded` clamps it
    against SOLID cells: max(u, 0) when the -x neighbour is SOLID, min(u, 0) when the cell
    itself is, 0 when both are (no moving solid here).

Each check below fails for a plausible mistake: a wrong neighbour set, F_U read in pass 2, the
hi/lo roles in `bounded` swapped, or a mean that counts unset neighbours.
"""
import numpy as np

import oracle
from natray_amd import abi
from tests.helpers import box_scene, cell_coords, records

SIDE = 16
F_EXT = 4  # pass 1's output fields (F_R..F_Z reused, fluid.wgsl:604)
F_EXT_SET = 7
WATER = [(8, 8, 8), (8, 9, 9), (8, 10, 8)]
SOLID_HI = (11, 8, 8)   # the cell itself is SOLID: its x face is clamped to <= 0
SOLID_LO = (12, 10, 8)  # the -x neighbour of (13, 10, 8) is SOLID: clamped to >= 0
SOLID_BOTH = [(3, 3, 3), (4, 3, 3)]  # the face of (4, 3, 3) lies between two SOLID cells


def u0(c):
    x, y, z = c
    return -float(100 * x + 10 * y + z)


def prepared():
    """The oracle after the projection, with every block of the 16^3 box in use."""
    centres = [((4 * bx + 2) / SIDE, (4 * by + 2) / SIDE, (4 * bz + 2) / SIDE)
               for bx in range(4) for by in range(4) for bz in range(4)]
    s = box_scene(records(centres), side=SIDE, fixed_dt=1e-4, flags=0)
    o = oracle.Oracle.from_scene(s)
    for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID, abi.PHASE_SOLVE, abi.PHASE_PROJECT):
        o.phase(ph)
    g = o.grid()
    assert g["n_blocks"] == 64
    cc = cell_coords(g, s.params.bdims)
    index = {tuple(int(v) for v in c): i for i, c in enumerate(cc)}
    return s, o, g, cc, index


def run_pattern(u_override=None):
    s, o, g, cc, index = prepared()
    cf = g["cf"].copy()
    cf[abi.F_TYPE][:] = abi.GAS
    for c in WATER:
        cf[abi.F_TYPE][index[c]] = abi.WATER
    for c in [SOLID_HI, SOLID_LO] + SOLID_BOTH:
        cf[abi.F_TYPE][index[c]] = abi.SOLID
    for d in range(3):
        cf[abi.F_U + d][:] = [u0(c) for c in cc.tolist()]
    for c, v in (u_override or {}).items():
        cf[abi.F_U][index[c]] = v
    o.write_grid(None, cf)
    o.phase(abi.PHASE_EXTEND)
    out = o.grid()["cf"]
    o.close()
    at = lambda f, c: float(out[f][index[c]])
    return at


def test_pass_one_means():
    at = run_pattern()
    f32 = np.float32
    # Set faces keep their own u (W1, W2, W3 and their +x neighbours).
    for c in WATER + [(9, 8, 8), (9, 9, 9), (9, 10, 8)]:
        assert at(F_EXT, c) == u0(c), c
        assert int(at(F_EXT_SET, c)) & 1 == 1, c
    # One set neighbour: (7,8,8) sees only W1's face on +x.
    assert at(F_EXT, (7, 8, 8)) == -888.0
    # Three: (8,9,8) has W1 on -y, W3 on +y and W2 on +z.
    assert at(F_EXT, (8, 9, 8)) == float(f32(-888 - 899 - 908) / f32(3))
    # Three: (9,9,8) between the set faces (9,8,8) (-y), (9,10,8) (+y) and (9,9,9) (+z).
    # One: (8,9,10) has W2 on -z only. Two: (8,10,9) has W2 on -y and W3 on -z.
    assert at(F_EXT, (9, 9, 8)) == float(f32(-988 - 999 - 1008) / f32(3))
    assert at(F_EXT, (8, 9, 10)) == -899.0
    assert at(F_EXT, (8, 10, 9)) == float(f32(-899 - 908) / f32(2))
    # No set neighbour: unchanged and unset.
    assert at(F_EXT, (6, 8, 8)) == u0((6, 8, 8))
    assert int(at(F_EXT_SET, (6, 8, 8))) & 1 == 0
    assert int(at(F_EXT_SET, (7, 8, 8))) & 1 == 1


def test_pass_two_reads_pass_one():
    at = run_pattern()
    f32 = np.float32
    # A set face of pass 1 takes F_EXT.
    assert at(abi.F_U, (7, 8, 8)) == -888.0
    # (6,8,8) was unset; its +x neighbour (7,8,8) is set with F_EXT = -888 (F_U = -788).
    assert at(abi.F_U, (6, 8, 8)) == -888.0
    # (7,9,8) has no set face around it (none of its neighbours is water or has water on -x),
    # so it is unset after pass 1. In pass 2 four of its neighbours are set: (8,9,8) (+x, F_EXT
    # -2695/3), (7,8,8) (-y, -888), (7,10,8) (+y, -908, from W3) and (7,9,9) (+z, -899, from
    # W2), summed in the order -x, +x, -y, +y, -z, +z.
    assert int(at(F_EXT_SET, (7, 9, 8))) & 1 == 0
    e889 = f32(-2695) / f32(3)
    want = (((f32(0) + e889) + f32(-888)) + f32(-908)) + f32(-899)
    assert at(abi.F_U, (7, 9, 8)) == float(want / f32(4))
    # Far away: no set neighbour in either pass, u unchanged.
    assert at(abi.F_U, (2, 13, 5)) == u0((2, 13, 5))


def test_bounded_last():
    # The cell (11,8,8) is SOLID itself: its face takes min(u, 0). Its -x neighbour (10,8,8) is
    # set in pass 1 (F_EXT -988 from (9,8,8)), so pass 2 gives -988, kept by min(., 0).
    # With a positive pass-2 value the clamp shows: override (9,8,8)'s u to +500 and (10,8,8)
    # then extends +500 into (11,8,8), which is clamped to 0.
    at = run_pattern({(9, 8, 8): 500.0})
    assert at(F_EXT, (10, 8, 8)) == 500.0
    assert at(abi.F_U, (11, 8, 8)) == 0.0
    # (13,10,8): -x neighbour SOLID, u = -1408 with no set neighbour: max(-1408, 0) = 0.
    assert at(abi.F_U, (13, 10, 8)) == 0.0
    # (12,10,8) is SOLID and its -x neighbour is not: min(-1308, 0) keeps -1308.
    assert at(abi.F_U, (12, 10, 8)) == -1308.0
    # Between two SOLID cells: 0. (3,3,3) is SOLID with gas on -x: min(-333, 0) = -333. (5,3,3)
    # has SOLID on -x: max(-533, 0) = 0.
    assert at(abi.F_U, (4, 3, 3)) == 0.0
    assert at(abi.F_U, (3, 3, 3)) == -333.0
    assert at(abi.F_U, (5, 3, 3)) == 0.0
    # A face untouched by solids and extension keeps its u.
    assert at(abi.F_U, (6, 3, 3)) == u0((6, 3, 3))


def test_other_axes_follow_their_own_faces():
    """The y faces: set where the cell or its -y neighbour is water. (8,11,8) has W3 as its -y
    neighbour, so its own y face is set and keeps u; (8,12,8) takes it in pass 1."""
    s, o, g, cc, index = prepared()
    cf = g["cf"].copy()
    cf[abi.F_TYPE][:] = abi.GAS
    cf[abi.F_TYPE][index[(8, 10, 8)]] = abi.WATER
    for d in range(3):
        cf[abi.F_U + d][:] = [u0(c) * (d + 1) for c in cc.tolist()]
    o.write_grid(None, cf)
    o.phase(abi.PHASE_EXTEND)
    out = o.grid()["cf"]
    o.close()
    at = lambda f, c: float(out[f][index[c]])
    # Its own y face is set (not the mean of its neighbours, which would be W3's -1816).
    assert at(F_EXT + 1, (8, 11, 8)) == 2 * u0((8, 11, 8))
    assert int(at(F_EXT_SET, (8, 11, 8))) & 2 == 2
    assert at(F_EXT + 1, (8, 12, 8)) == 2 * u0((8, 11, 8))
    # Its x face is not set: x faces need water in (8,11,8) or (7,11,8).
    assert at(F_EXT, (8, 12, 8)) == u0((8, 12, 8))
    # z: (8,10,9) has W3 as its -z neighbour, so its z face is set.
    assert at(F_EXT + 2, (8, 10, 9)) == 3 * u0((8, 10, 9))
    assert int(at(F_EXT_SET, (8, 10, 9))) & 4 == 4
