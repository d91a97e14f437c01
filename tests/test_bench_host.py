"""bench.py's contract pieces that need no GPU: it refuses more than one GPU with a message, it
writes the native driver's scene file for the profiles (with its bullets trailer), and it finds the counter-measured HBM
traffic of the solve in the committed profile summary."""
import os
import struct
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          cwd=ROOT, timeout=300)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2")
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout.strip() == ""


def test_write_scene_file(tmp_path):
    path = tmp_path / "sphere.scene"
    r = run_bench("--workload", "sphere", "--write-scene", str(path))
    assert r.returncode == 0, r.stderr
    raw = path.read_bytes()
    assert raw[:8] == b"NATRAYSC"
    version, n = struct.unpack("<II", raw[8:16])
    assert version == 1 and n == 32768
    assert len(raw) > 16 + 192 + 96 + n * 80


def test_scene_file_carries_bullets(tmp_path):
    import ctypes as C

    import numpy as np

    from natray_amd import abi, scenes

    s = scenes.water_sphere(side=16, n=512)
    b, pt = scenes.bullet_shot(arrive_ns=1234)
    path = tmp_path / "bullet.scene"
    scenes.write_scene_file(s, str(path), bullets=([b], pt[None]))
    raw = path.read_bytes()
    head = 16 + C.sizeof(abi.Params) + C.sizeof(abi.Scene) + C.sizeof(abi.State) + s.n * 80
    assert raw[head:head + 8] == b"BULLETS\0" and struct.unpack("<I", raw[head + 8:head + 12]) == (1,)
    rec = abi.Bullet.from_buffer_copy(raw[head + 12:head + 12 + C.sizeof(abi.Bullet)])
    assert C.sizeof(abi.Bullet) == 112 and list(rec.info) == [1234, list(b.info)[1], 0, 0]
    got = np.frombuffer(raw[head + 12 + 112:], np.float32).reshape(1, abi.BULLET_PATCHES, 8)
    np.testing.assert_array_equal(got, pt[None])


def test_solve_traffic_is_read_from_the_profile():
    import bench

    t = bench.measured_traffic("k_solve<false,One>")
    assert t is not None and 1e8 < t < 1e9
    assert bench.measured_traffic("no_such_kernel") is None


def test_cpu_baseline_reports_median_threads_and_host(monkeypatch):
    """The CPU baseline: the median of its timed runs on the job's cores (OMP_NUM_THREADS), a
    one-thread figure, and the host's CPU model, on a small scene."""
    import numpy as np

    import bench
    from natray_amd import scenes

    monkeypatch.setenv("OMP_NUM_THREADS", "2")
    s = scenes.water_sphere(side=32, n=4096)
    r = bench.cpu_baseline(s, 200, budget_s=0.5, runs=5, one_thread_steps=1)
    assert r["cores"] == 2 and r["kind"] == "port" and len(r["runs"]) == 5
    assert r["value"] == float(np.median(r["runs"])) and r["value"] > 0 and r["one_thread_value"] > 0
    assert "not the lavapipe" in r["label"] and r["host"]["logical_cpus"] == os.cpu_count()


def test_stated_radius_scenes():
    """The stated geometry: every cell whose centre lies inside the radius holds 8 particles."""
    import numpy as np

    from natray_amd import scenes

    s = scenes.stated_radius_sphere()
    side = 64
    cells = np.unique(np.floor(s.particles[:, :3].astype(np.float64) * side).astype(int), axis=0)
    d = np.linalg.norm((cells + 0.5) / side - 0.5, axis=1)
    assert s.n == 8 * len(cells) == 138048 and d.max() < 0.25
    # Brute force over the whole grid: exactly those cells lie inside.
    g = (np.indices((side, side, side)).reshape(3, -1).T + 0.5) / side
    assert int((np.linalg.norm(g - 0.5, axis=1) < 0.25).sum()) == len(cells)
    assert scenes.cells_within(256, 0.3) * 8 == 15179264


def test_phase_traffic_is_read_from_the_profile():
    """Every phase's kernels are in the counter summary, so the bench can put counter-side
    bytes beside the algorithmic ones."""
    import bench

    t = bench.measured_phase_traffic()
    assert t is not None and set(t) == set(bench.PHASE_KERNELS)
    assert t["g2p"] > t["p2g"] > 1e8 and all(v > 0 for v in t.values())
