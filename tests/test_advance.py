"""The kept host API on the GPU: Sim::advance (src/sim.rs:498-522) with its batch sizing and CG
budget adaptation (src/sim.rs:636-645), and one step, against the oracle.

advance() sizes batches from the measured wall time per step, so its batches differ from run
to run. The library logs every natray_sim_run call it makes (natray_sim_run_log); the test
checks the log against the reference's rules (at most MAX_BATCH = 256 steps and the budget
left per batch, no batch once the stop time is reached, the budget chain 64 ->
natray_next_iters(...)) and replays exactly those batches on the oracle, which must end bit
for bit where the GPU did.
"""
import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi
from tests.test_gpu_parity import SCENES, assert_same_float, compare_state, say

pytestmark = pytest.mark.gpu

FIRST_ITERS = 64  # Sim::new's CG budget (src/sim.rs:480)


def replay(orc, log, start_iters, tag):
    """Runs the logged batches on the oracle; checks each batch's outcome and the budget chain."""
    iters = start_iters
    for k, rec in enumerate(log):
        assert rec["iters"] == iters, (tag, k, rec, iters)
        b = orc.run(rec["n_steps"], iters=rec["iters"], stop_ns=rec["stop_ns"])
        assert (b.t_ns, b.iters_max, b.unconverged) == (rec["t_ns"], rec["iters_max"], rec["unconverged"]), (tag, k, rec, b.as_dict())
        iters = int(abi.lib().natray_next_iters(iters, b.iters_max, b.unconverged))
    return iters


def check_batches(log, stop_ns, budget, taken):
    assert sum(r["taken"] for r in log) == taken
    left = budget
    for r in log:
        assert r["stop_ns"] == stop_ns
        assert 1 <= r["n_steps"] <= min(left, 256)
        assert r["taken"] <= r["n_steps"]
        left -= r["taken"]
    # Every batch but the last ended short of the stop time: advance stops once it is reached.
    for r in log[:-1]:
        assert r["t_ns"] < stop_ns


def test_advance_replays_on_the_oracle():
    s = SCENES["impact_sgs_32k"]()
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    say("[advance] 32k impact to t = 0.5 ms, budget 200")
    t_target, budget = 0.5e-3, 200
    taken = sim.advance(t_target, budget)
    st = sim.state()
    stop = int(round(t_target * 1e9))
    log = sim.run_log()
    say(f"[advance] {taken} steps in {len(log)} batches: {[r['n_steps'] for r in log]}, budgets {[r['iters'] for r in log]}")
    assert st.t_ns == stop, (st.t_ns, stop)  # the last step lands on the stop time
    assert taken == st.step and taken > 5
    check_batches(log, stop, budget, taken)
    assert log[0]["n_steps"] == 1  # no step time measured yet: one step (src/sim.rs:513)
    iters = replay(orc, log, FIRST_ITERS, "advance 1")
    compare_state("advance", st, orc.state())
    assert_same_float("advance particles", sim.particles(), orc.particles())

    say("[advance] on to t = 1 ms with a budget of 3 steps")
    n0 = len(log)
    taken2 = sim.advance(1.0e-3, 3)
    log2 = sim.run_log()[n0:]
    assert taken2 == 3, taken2
    check_batches(log2, 1_000_000, 3, taken2)
    iters = replay(orc, log2, iters, "advance 2")
    compare_state("advance 2", sim.state(), orc.state())
    assert_same_float("advance 2 particles", sim.particles(), orc.particles())

    say("[advance] one step() with the adapted budget")
    a = sim.step()
    rec = sim.run_log()[-1]
    assert (rec["n_steps"], rec["iters"], rec["stop_ns"]) == (1, iters, 0xFFFFFFFF)
    b = orc.run(1, iters=iters, stop_ns=0xFFFFFFFF)
    compare_state("step", a, b)
    assert_same_float("step particles", sim.particles(), orc.particles())

    say("[advance] a target already reached takes no step")
    assert sim.advance(0.5e-3, 10) == 0
    sim.close()
    orc.close()


def test_budget_doubles_after_an_unconverged_batch():
    """A batch whose steps run out of CG iterations doubles the budget (src/sim.rs:639-640);
    the next batch runs with it, and both end where the oracle's same batches do."""
    s = SCENES["impact_sgs_32k"]()
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    say("[advance] starved budget")
    sim.run(2, iters=400)
    orc.run(2, iters=400, stop_ns=0xFFFFFFFF)
    a = sim.run(1, iters=3)  # an explicit budget of 3: the step runs out of iterations
    orc.run(1, iters=3, stop_ns=0xFFFFFFFF)
    assert a.unconverged == 1 and a.solve_iters == 3, a.as_dict()
    log = sim.run_log()
    assert [r["iters"] for r in log] == [400, 3]
    # The adaptation follows the handle's own budget: 64 -> next_iters after the first batch
    # -> doubled after the starved one.
    want = int(abi.lib().natray_next_iters(FIRST_ITERS, log[0]["iters_max"], log[0]["unconverged"]))
    want = int(abi.lib().natray_next_iters(want, 3, 1))
    b = sim.step()
    assert sim.run_log()[-1]["iters"] == want, (sim.run_log()[-1], want)
    c = orc.run(1, iters=want, stop_ns=0xFFFFFFFF)
    compare_state("after doubling", b, c)
    assert_same_float("after doubling particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()
