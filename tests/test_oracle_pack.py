"""The keyframe pack of the water (shaders/keys.wgsl:21-37, shaders/pack.wgsl:33-39) in the
oracle, pinned by known answers, the unpacking the reference's viewer applies (keys.wgsl:26-31)
and the quantisation bound, the clamp at the box and the live flag."""
import numpy as np

import oracle
from natray_amd import abi, scenes

Q_LO = np.array([-4.0, -0.5, -4.0])
STEP = 8.0 / 2097151.0


def unpack(w):
    x = w[:, 0] & 0x1FFFFF
    y = (w[:, 0] >> 21) | ((w[:, 1] & 0x3FF) << 11)
    z = (w[:, 1] >> 10) & 0x1FFFFF
    return Q_LO + np.stack([x, y, z], axis=1).astype(np.float64) * STEP


def packed(P):
    s = scenes.water_sphere(side=16, n=P.shape[0] if P.shape[0] % 8 == 0 else 8)
    s.particles = np.ascontiguousarray(P, np.float32)
    s.params.dims[3] = P.shape[0]
    o = oracle.Oracle.from_scene(s, "f32")
    return o.pack_water()


def records(pos, alive):
    P = np.zeros((len(pos), abi.PARTICLE_FLOATS), np.float32)
    P[:, :3] = pos
    P[:, 3] = 1.0
    P[:, 7] = alive
    return P


def test_known_answers():
    pos = np.array([[-4.0, -0.5, -4.0], [4.0, 7.5, 4.0], [0.0, 0.0, 0.0], [10.0, -9.0, 0.5]] + [[0.0, 0.0, 0.0]] * 4)
    w = packed(records(pos, [1, 0, 1, 1, 0, 0, 0, 0]))
    # The box's low corner packs to zeros, its high corner to all ones.
    assert w[0, 0] == 0 and w[0, 1] == 1 << 31
    assert w[1, 0] == 0xFFFFFFFF and w[1, 1] == 0x7FFFFFFF
    # The origin: x = 0.5 * 2097151 + 0.5 truncated, y = 0.0625 * 2097151 + 0.5, z as x.
    qx, qy = int(np.float32(0.5) * np.float32(2097151) + np.float32(0.5)), int(np.float32(0.0625) * np.float32(2097151) + np.float32(0.5))
    assert w[2, 0] == (qx | (qy << 21)) & 0xFFFFFFFF
    assert w[2, 1] == ((qy >> 11) | (qx << 10) | (1 << 31)) & 0xFFFFFFFF
    # Outside the box: clamped to its faces.
    u = unpack(w[3:4])[0]
    np.testing.assert_allclose(u, [4.0, -0.5, (0.5 + 4.0) / 8.0 * 8.0 - 4.0], atol=STEP)


def test_round_trip_and_flag():
    rng = np.random.default_rng(3)
    pos = Q_LO + rng.random((4096, 3)) * 8.0
    alive = (rng.random(4096) < 0.7).astype(np.float32)
    w = packed(records(pos, alive))
    back = unpack(w)
    # float32 inputs, rounded to the nearest of 2^21 - 1 steps: within half a step and the
    # float32 rounding of the scaled position.
    np.testing.assert_array_less(np.abs(back - pos.astype(np.float32)), 0.5 * STEP + 8.0 * 2.0 ** -24 * 2097151 * STEP)
    np.testing.assert_array_equal((w[:, 1] >> 31) == 1, alive > 0.5)
