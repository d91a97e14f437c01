"""begin_step's clock (shaders/fluid.wgsl:31-89) against a table of closed forms: the step is
min(DT_MAX, CFL h / vmax, 1.25 x the last free step, the rubber and air substep limits) in
whole nanoseconds (at least 20), it lands exactly on the stop time, and a step that starts
at or past the stop does nothing and keeps the clock."""
import numpy as np
import pytest

import oracle
from natray_amd import abi
from tests.helpers import box_scene, records

F32MAX = np.finfo(np.float32).max


def clock(vmax, free_ns=0, stop=0xFFFFFFFF, t_ns=0, cfl=0.5, fixed=0.0, mem_dt=F32MAX, air_dt=F32MAX, side=16):
    s = box_scene(records([[0.5, 0.5, 0.5]]), side=side, fixed_dt=fixed, cfl=cfl)
    s.params.limits[3] = stop
    st = abi.State()
    st.t_ns = t_ns
    st.max_speed = vmax
    st.free_dt_ns = free_ns
    st.mem_dt = mem_dt
    st.cap_dt = F32MAX
    st.air_dt = air_dt
    o = oracle.Oracle(s.params, s.particles, s.scene, st)
    o.phase(abi.PHASE_BEGIN)
    return o.state()


def f32(x):
    return np.float32(x)


@pytest.mark.parametrize("vmax", [0.0, 0.5, 3.0, 40.0, 300.0])
def test_cfl_step(vmax):
    h = f32(1.0 / 16)
    st = clock(vmax)
    dt = f32(1e-3) if vmax == 0.0 else min(f32(1e-3), f32(0.5) * h / f32(vmax))
    assert st.dt_ns == max(int(np.float32(dt) * np.float32(1e9)), 20)
    assert st.t_ns == st.dt_ns and st.step == 1 and st.free_dt_ns == st.dt_ns
    assert st.dt_flow == dt


def test_growth_is_limited_to_a_quarter():
    st = clock(0.0, free_ns=100_000)
    assert st.dt_ns == int(f32(1.25) * f32(100_000) * f32(1e-9) * f32(1e9))
    assert 124_990 <= st.dt_ns <= 125_000


def test_lands_on_the_stop_time_and_keeps_the_free_step():
    st = clock(0.0, t_ns=1_000_000, stop=1_000_300)
    assert st.dt_ns == 300 and st.t_ns == 1_000_300
    assert st.free_dt_ns == 1_000_000  # the step it would have taken


def test_dead_step_at_or_past_the_stop():
    for t in (5000, 6000):
        st = clock(10.0, t_ns=t, stop=5000, free_ns=77)
        assert st.dt_ns == 0 and st.t_ns == t and st.step == 0 and st.free_dt_ns == 77


def test_substep_limits_and_floor():
    # Two air substeps of 1e-8 s each limit the step to 2e-8 s, which floors to 20 ns.
    st = clock(0.0, air_dt=1e-8)
    assert st.dt_ns == 20
    st = clock(0.0, air_dt=2e-7)
    assert st.dt_ns == int(f32(2.0) * f32(2e-7) * f32(1e9))


def test_fixed_step():
    st = clock(1000.0, fixed=1e-4)
    assert st.dt_ns == 100_000


def test_step_resets_per_step_state():
    """A live step zeroes the max speed and block count and carries the block count over."""
    from natray_amd import scenes

    s = scenes.water_sphere(side=32, n=8 * 200, fixed_dt=1e-4)
    o = oracle.Oracle.from_scene(s)
    o.run(2, iters=50)
    before = o.state()
    assert before.n_blocks > 0
    o.phase(abi.PHASE_BEGIN)
    st = o.state()
    assert st.n_blocks == 0 and st.n_blocks_prev == before.n_blocks and st.max_speed == 0.0
    assert st.mem_dt == F32MAX and st.cap_dt == F32MAX and st.air_dt == F32MAX
