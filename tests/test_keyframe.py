"""A whole keyframe on the GPU (natray_sim_pack_keyframe) against the oracle, on the spray
scene with a membrane and triangle flags: header, water, membrane and alive words bit for bit;
the spray's part bit for bit as a multiset of live entries (slots come from an atomic add, so
their order is not fixed), its live bits and prefix words consistent with the GPU's own spray
records, and each slot's weight and time made as a multiset of rows."""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.test_oracle_keyframe import keyframe_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def sorted_rows(a):
    return a[np.lexsort(a.T[::-1])] if len(a) else a


def test_keyframe_matches_the_oracle():
    s, orc, shell, alive = keyframe_case(steps=0)
    say(f"[keyframe] breakup scene N={s.n}, {shell[2].size} vertices, {alive.size} triangles")
    sim = Sim.from_scene(s)
    sim.set_membrane(*shell)
    for step in range(6):
        say(f"[keyframe] step {step}")
        sim.run(1, iters=200)
        orc.run(1, 200)
        g, gw = sim.pack_keyframe(alive)
        o, ow = orc.pack_keyframe(alive)
        assert g.size == o.size and g[0] == o[0] and g[1] == o[1] and not g[2:4].any()
        N, V, T, n = s.n, shell[2].size, alive.size, int(g[0])
        spw = 4 + 2 * (N + V) + (T + 31) // 32
        np.testing.assert_array_equal(g[:spw], o[:spw], err_msg=f"step {step}: water, membrane or alive words")
        nw = (n + 31) // 32
        np.testing.assert_array_equal(sorted_rows(g[spw + 2 * nw:].reshape(-1, 3)),
                                      sorted_rows(o[spw + 2 * nw:].reshape(-1, 3)), err_msg=f"step {step}: spray entries")
        rec = sim.spray()["records"][:n]
        live = rec[:, 3] > 0
        bits = g[spw:spw + nw]
        got = (bits[np.arange(n) // 32] >> (np.arange(n) % 32)) & 1
        np.testing.assert_array_equal(got.astype(bool), live)
        counts = np.array([bin(int(b)).count("1") for b in bits], np.int64)
        np.testing.assert_array_equal(g[spw + nw:spw + 2 * nw], np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.uint32))
        np.testing.assert_array_equal(sorted_rows(gw), sorted_rows(ow))
    assert n > 100
    with pytest.raises(Exception):
        abi.check(abi.lib().natray_sim_pack_keyframe(sim._h, alive.size, abi.ptr(alive), abi.ptr(np.zeros(8, np.uint32)), 8,
                                                     abi.ptr(np.zeros(1, np.uint32)), None))
    sim.close()
    orc.close()
