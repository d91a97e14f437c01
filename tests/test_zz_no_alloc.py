"""natray_sim_pack_keyframe_static, the keyframe pack that allocates nothing (its spray index in
a fixed device array, the caller's device buffers for flags, words and weights), against
natray_sim_pack_keyframe on the same state: the same words and weights, bit for bit, on two
calls in a row, the second on a smaller scene (fewer particles, vertices, triangles and spray
slots) so that it runs over what the first left in the array."""
import ctypes as C
import sys

import numpy as np
import pytest

from natray_amd import Sim, abi, scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself uses (the test allocates; the
    entry point under test does not)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = C.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
            h.hipDeviceSynchronize.argtypes = []
            cls._hip = h
        return cls._hip

    def __init__(self, nbytes, fill=0xAB):
        self.nbytes = max(int(nbytes), 4)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert self.hip().hipMemset(self.ptr, fill, self.nbytes) == 0  # stale bytes the pack must overwrite

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    def free(self):
        self.hip().hipFree(self.ptr)


def static_pack(sim, alive):
    L = abi.lib()
    dalive = DeviceBuffer(alive.nbytes)
    dalive.upload(alive)
    n = C.c_uint32()
    abi.check(L.natray_sim_pack_keyframe_static(sim._h, alive.size, dalive.ptr, None, 0, C.byref(n), None))  # the size only
    total = int(n.value)
    dwords = DeviceBuffer(4 * total)
    n_spray = min(sim.spray()["count"], int(sim.params.limits[1]))
    dw = DeviceBuffer(8 * n_spray)
    # Too small a buffer is refused before anything is written.
    with pytest.raises(abi.NatrayError):
        abi.check(L.natray_sim_pack_keyframe_static(sim._h, alive.size, dalive.ptr, dwords.ptr, total - 1, C.byref(n), dw.ptr))
    abi.check(L.natray_sim_pack_keyframe_static(sim._h, alive.size, dalive.ptr, dwords.ptr, total, C.byref(n), dw.ptr))
    assert int(n.value) == total
    words = dwords.download(np.uint32, total)
    weights = dw.download(np.float32, 2 * n_spray).reshape(-1, 2)
    for b in (dalive, dwords, dw):
        b.free()
    return words, weights


def case(n_ball, n_slow, n_fast, n_ground, n_verts, n_tris, steps):
    s = scenes.breakup_scene(n_ball=n_ball, n_slow=n_slow, n_fast=n_fast, n_ground=n_ground)
    shell = scenes.membrane_shell(scenes.water_sphere(side=32, n=3200, seed=9), n_verts=n_verts)
    alive = (np.random.default_rng(n_tris).uniform(size=n_tris) < 0.8).astype(np.uint32)
    sim = Sim.from_scene(s)
    sim.set_membrane(*shell)
    sim.run(steps, iters=200)
    return sim, alive


def test_static_pack_matches_the_allocating_pack():
    got = []
    for k, (args, tag) in enumerate([((8 * 400, 48, 32, 32, 300, 1000, 6), "first"), ((8 * 200, 24, 16, 16, 120, 100, 4), "second, smaller")]):
        sim, alive = case(*args)
        say(f"[no-alloc] {tag}: N={sim.n}, {alive.size} triangles")
        want, want_w = sim.pack_keyframe(alive)
        words, weights = static_pack(sim, alive)
        n_spray = int(want[0])
        say(f"[no-alloc] {tag}: {words.size} words, {n_spray} spray slots, {int(want[1])} live")
        np.testing.assert_array_equal(words, want, err_msg=f"{tag}: keyframe words")
        np.testing.assert_array_equal(weights.view(np.uint32), np.asarray(want_w, np.float32).view(np.uint32), err_msg=f"{tag}: weights")
        # A call without the weights' buffer gives the same words.
        dalive = DeviceBuffer(alive.nbytes)
        dalive.upload(alive)
        dwords = DeviceBuffer(4 * want.size)
        n = C.c_uint32()
        abi.check(abi.lib().natray_sim_pack_keyframe_static(sim._h, alive.size, dalive.ptr, dwords.ptr, want.size, C.byref(n), None))
        np.testing.assert_array_equal(dwords.download(np.uint32, want.size), want, err_msg=f"{tag}: without weights")
        dalive.free()
        dwords.free()
        got.append((sim.n, alive.size, n_spray, words.size))
        sim.close()
    assert got[0][2] > 100 and got[1][2] > 0
    assert got[1][0] < got[0][0] and got[1][1] < got[0][1] and got[1][3] < got[0][3], got
