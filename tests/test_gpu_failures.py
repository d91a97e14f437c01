"""The failure and limit paths of the substep through the C ABI, forced from the host, against
the oracle: block overflow is reported as NATRAY_E_OVERFLOW
with the first `capacity` claimed blocks kept (fluid.wgsl:117-122, src/sim.rs:635); a CG cap
below what the tolerance needs gives unconverged steps, bit for bit (fluid.wgsl:572-577); a
stop time truncates the last step onto it and every step from there on is dead, changing
nothing (fluid.wgsl:62-88)."""
import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi, scenes
from tests.helpers import box_scene, records
from tests.test_gpu_parity import SCENES, assert_same_float, compare_grid, compare_state

pytestmark = pytest.mark.gpu


def test_overflow_is_reported():
    print("[failures] overflow", flush=True)
    rng = np.random.default_rng(3)
    P = records(rng.uniform(0.1, 0.9, (500, 3)))
    s = box_scene(P, side=32, capacity=8)
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s)
    sim.phase(abi.PHASE_BEGIN)
    orc.phase(abi.PHASE_BEGIN)
    with pytest.raises(abi.NatrayError) as e:
        sim.phase(abi.PHASE_BIN)
    assert e.value.code == abi.E_OVERFLOW
    assert orc.phase(abi.PHASE_BIN) == abi.E_OVERFLOW
    compare_state("overflow", sim.state(), orc.state())
    assert sim.state().overflow == 1 and sim.state().n_blocks > 8
    compare_grid("overflow", sim.grid(), orc.grid(), cf_fields=(), acc=False)
    sim.close()


@pytest.mark.parametrize("name", ["impact_sgs_32k", "impact_pipelined_32k", "impact_pipelined_32k_mem"])
def test_unconverged_steps_match_the_oracle(name):
    print(f"[failures] CG capped at 2 iterations, {name}", flush=True)
    s = SCENES[name]()
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    sim.run(3, iters=400)
    orc.run(3, 400)
    a = sim.run(4, iters=2)
    b = orc.run(4, 2)
    compare_state("capped", a, b)
    assert a.unconverged == 4 and a.solve_iters == 2 and a.iters_max == 2
    assert_same_float("capped particles", sim.particles(), orc.particles())
    sim.close()


def test_stop_time_truncates_then_steps_are_dead():
    print("[failures] stop time", flush=True)
    s = SCENES["impact_sgs_32k"]()
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    st = sim.run(2, iters=400)
    stop = st.t_ns + st.free_dt_ns // 3  # inside the next step
    a = sim.run(1, iters=400, stop_ns=stop)
    orc.run(2, 400)
    b = orc.run(1, 400, stop)
    compare_state("truncated", a, b)
    assert a.t_ns == stop and 0 < a.dt_ns < a.free_dt_ns
    P = sim.particles()
    dead = sim.run(3, iters=400, stop_ns=stop)
    compare_state("dead", dead, orc.run(3, 400, stop))
    assert dead.dt_ns == 0 and dead.t_ns == stop and dead.step == a.step
    assert_same_float("dead particles", sim.particles(), P)
    assert_same_float("dead particles oracle", orc.particles(), P)
    sim.close()


@pytest.mark.parametrize("name", ["impact_sgs_32k", "impact_pipelined_32k"])
def test_solve_launches_give_the_same_bits(name):
    """The plain solve launch (the default, one graph per step) and the cooperative one, over
    the same steps from the same state, with graphs and kernel by kernel."""
    print(f"[failures] plain and cooperative solve launches, {name}", flush=True)
    s = SCENES[name]()
    out = []
    for mode, graph in ((abi.SOLVE_PLAIN, True), (abi.SOLVE_COOPERATIVE, True), (abi.SOLVE_PLAIN, False)):
        sim = Sim.from_scene(s)
        sim.set_solve_launch(mode)
        sim.set_graph(graph)
        st = sim.run(6, iters=400)
        out.append((sim.particles(), st))
        sim.close()
    for P, st in out[1:]:
        compare_state("launch", st, out[0][1])
        assert_same_float("launch particles", P, out[0][0])
    assert out[0][1].solve_iters > 5


def test_plain_solve_reports_workgroups_not_resident():
    """A plain launch with a workgroup that is not resident, forced from the host: the last
    workgroup stays out, the others' bounded waits for its chunk sums and halo words run out and
    abort the solve, and the step reports NATRAY_E_SOLVER with a message naming the cooperative
    launch. The handle then runs again in either mode."""
    print("[failures] plain solve residency check", flush=True)
    s = SCENES["impact_sgs_32k"]()
    sim = Sim.from_scene(s)
    sim.run(2, iters=400)
    sim.set_solve_launch(abi.SOLVE_PLAIN_TEST_CHECK)
    with pytest.raises(abi.NatrayError) as e:
        sim.run(1, iters=400)
    assert e.value.code == abi.E_SOLVER and "COOPERATIVE" in str(e.value)
    for mode in (abi.SOLVE_COOPERATIVE, abi.SOLVE_PLAIN):
        sim.set_solve_launch(mode)
        st = sim.run(2, iters=400)
        assert st.solve_iters > 5 and st.unconverged == 0
    sim.close()
