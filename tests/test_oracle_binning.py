"""The oracle's binning (clear/mark/claim/scan/sort, shaders/fluid.wgsl:91-196) against brute
force: the claimed slots are exactly the union of the particles' stencil blocks (cells c-1 to
c+2 of the own cell c), counts are the own-block histogram, starts are its exclusive prefix
sum, every alive particle appears once in its own block's range, and an undersized capacity
raises the overflow flag."""
import numpy as np
import pytest

import oracle
from natray_amd import abi
from tests.helpers import box_scene, records


def brute(pos, alive, side):
    cells = np.floor(pos.astype(np.float32) / np.float32(1.0 / side)).astype(int)
    bd = side // 4
    claimed = set()
    own = {}
    for i, c in enumerate(cells):
        if not alive[i]:
            continue
        lo = np.clip((c - 1) >> 2, 0, bd - 1)
        hi = np.clip((c + 2) >> 2, 0, bd - 1)
        for x in range(lo[0], hi[0] + 1):
            for y in range(lo[1], hi[1] + 1):
                for z in range(lo[2], hi[2] + 1):
                    claimed.add((x * bd + y) * bd + z)
        b = c >> 2
        own.setdefault(int((b[0] * bd + b[1]) * bd + b[2]), []).append(i)
    return sorted(claimed), own


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_binning_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    side = 32
    n = 800
    # Clustered particles, some dead, some at the domain edge.
    pos = np.concatenate([rng.uniform(0.3, 0.5, (n // 2, 3)), rng.uniform(0.0, 1.0, (n // 2, 3))])
    pos = np.clip(pos, 0.0, 0.999)
    alive = rng.uniform(size=n) > 0.1
    P = records(pos)
    P[:, 7] = alive.astype(np.float32)
    s = box_scene(P, side=side, capacity=(side // 4) ** 3)
    o = oracle.Oracle.from_scene(s)
    o.phase(abi.PHASE_BEGIN)
    o.phase(abi.PHASE_BIN)
    g = o.grid()
    claimed, own = brute(P[:, :3], alive, side)
    assert g["n_blocks"] == len(claimed)
    np.testing.assert_array_equal(g["block_list"], np.array(claimed, np.uint32))
    counts = np.array([len(own.get(sl, [])) for sl in claimed])
    np.testing.assert_array_equal(g["block_start"], np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32))
    bp, so = o.bins()
    idx = {sl: i for i, sl in enumerate(claimed)}
    for sl in range(len(bp)):
        assert bp[sl] == idx.get(sl, abi.NONE)
    for sl, members in own.items():
        b = idx[sl]
        got = np.sort(so[g["block_start"][b]:g["block_start"][b + 1]])
        np.testing.assert_array_equal(got, np.sort(np.array(members, np.uint32)))
    assert o.state().overflow == 0


def test_overflow_sets_the_flag_and_keeps_capacity():
    rng = np.random.default_rng(3)
    side = 32
    P = records(rng.uniform(0.1, 0.9, (500, 3)))
    s = box_scene(P, side=side, capacity=8)
    o = oracle.Oracle.from_scene(s)
    o.phase(abi.PHASE_BEGIN)
    rc = o.phase(abi.PHASE_BIN)
    st = o.state()
    assert rc == abi.E_OVERFLOW and st.overflow == 1
    assert st.n_blocks > 8 and o.n_blocks() == 8
    g = o.grid()
    claimed, _ = brute(P[:, :3], np.ones(500, bool), side)
    # The blocks kept are the first `capacity` claimed slots in slot order.
    np.testing.assert_array_equal(g["block_list"], np.array(claimed[:8], np.uint32))


def test_dead_step_bins_nothing():
    """A step that starts at or past the stop time has dt = 0 and every phase does nothing
    (shaders/fluid.wgsl:62-88)."""
    P = records(np.full((10, 3), 0.5))
    s = box_scene(P, side=16)
    s.params.limits[3] = 0
    o = oracle.Oracle.from_scene(s)
    for ph in range(abi.N_PHASES):
        o.phase(ph)
    st = o.state()
    assert st.dt_ns == 0 and st.t_ns == 0 and st.step == 0 and st.n_blocks == 0
    np.testing.assert_array_equal(o.particles(), P)
