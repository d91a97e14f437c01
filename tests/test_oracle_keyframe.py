"""A whole keyframe in the oracle (shaders/pack.wgsl:33-124, layout shaders/keys.wgsl:1-54),
pinned by what the layout and the data fix:

- the size is the reference's key_bytes word count (src/cache.rs:408-416, before rounding);
- the header holds the spray slot and live counts;
- the water's words are pack_water's, the membrane's unpack (keys.wgsl:26-31) to the vertices
  within half a quantisation step (plus the f32 rounding of the scaling), and the alive bits are the flags;
- the spray: a live bit for every slot with a radius, the exclusive prefix of the live counts,
  and 3 words per live slot that unpack to its position and hold its radius's bits, in slot
  order; spray_w is each slot's weight and time made.
"""
import numpy as np

import oracle
from natray_amd import scenes
from tests.test_oracle_pack import STEP, unpack

# Half a step, plus the f32 rounding of (p - lo) / 8 * (2^21 - 1) before it is truncated.
BOUND = 0.5 * STEP + 8.0 * 2.0 ** -24 * 2097151 * STEP


def keyframe_case(steps=6, n_tris=1000, seed=3):
    s = scenes.breakup_scene()
    o = oracle.Oracle.from_scene(s, "f32")
    shell = scenes.membrane_shell(scenes.water_sphere(side=32, n=3200, seed=9), n_verts=300)
    o.set_membrane(*shell)
    o.run(steps, 200)
    alive = (np.random.default_rng(seed).uniform(size=n_tris) < 0.8).astype(np.uint32)
    return s, o, shell, alive


def test_keyframe_layout_and_contents():
    s, o, shell, alive = keyframe_case()
    words, sw = o.pack_keyframe(alive)
    N, V, T = s.n, shell[2].size, alive.size
    sp = o.spray()
    n = min(sp["count"], int(s.params.limits[1]))
    rec = sp["records"][:n]
    live = rec[:, 3] > 0
    nw = (n + 31) // 32
    assert n > 100 and live.sum() > 50 and (~live).sum() > 0
    assert words.size == 4 + 2 * (N + V) + (T + 31) // 32 + 2 * nw + 3 * int(live.sum())
    assert list(words[:4]) == [n, int(live.sum()), 0, 0]
    np.testing.assert_array_equal(words[4:4 + 2 * N].reshape(-1, 2), o.pack_water())
    mem = words[4 + 2 * N:4 + 2 * (N + V)].reshape(-1, 2)
    assert np.abs(unpack(mem) - shell[0][:, :3].astype(np.float64)).max() <= BOUND
    assert not (mem[:, 1] >> 31).any()
    aw = 4 + 2 * (N + V)
    bits = words[aw:aw + (T + 31) // 32]
    got_alive = (bits[np.arange(T) // 32] >> (np.arange(T) % 32)) & 1
    np.testing.assert_array_equal(got_alive, alive)
    spw = aw + (T + 31) // 32
    live_bits = words[spw:spw + nw]
    got_live = (live_bits[np.arange(n) // 32] >> (np.arange(n) % 32)) & 1
    np.testing.assert_array_equal(got_live.astype(bool), live)
    counts = np.array([bin(int(b)).count("1") for b in live_bits])
    np.testing.assert_array_equal(words[spw + nw:spw + 2 * nw], np.concatenate([[0], np.cumsum(counts)[:-1]]))
    entries = words[spw + 2 * nw:].reshape(-1, 3)
    np.testing.assert_array_equal(entries[:, 2].view(np.float32), rec[live, 3].astype(np.float32))
    assert np.abs(unpack(entries[:, :2]) - rec[live, :3].astype(np.float32)).max() <= BOUND
    np.testing.assert_array_equal(sw, rec[:, [7, 11]].astype(np.float32))


def test_keyframe_without_spray_or_membrane_is_the_water():
    s = scenes.water_sphere(side=16, n=512)
    o = oracle.Oracle.from_scene(s, "f32")
    o.run(2, 200)
    words, sw = o.pack_keyframe()
    assert list(words[:4]) == [0, 0, 0, 0] and words.size == 4 + 2 * s.n and sw.size == 0
    np.testing.assert_array_equal(words[4:].reshape(-1, 2), o.pack_water())
