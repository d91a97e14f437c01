"""The learned warm start of the pressure solve through the C ABI against the oracle.

Per step the GPU computes r0 = b - A p_carried, lists the 16^3 tiles with water, feeds each
tile's (r0 / s, cell kind) to the network, forms guess = p_carried + s o on the tiles' water
cells and starts the solve from it when |b - A guess|^2 < |r0|^2. Checked here, bit for bit:
the network's inputs against the oracle's r0 and tile layout, the guess against the network's
own forward pass on those inputs, the decision against the oracle's residuals, and the whole
solve from the guess (register and memory forms) against the oracle's solve from the same
guess, with the following phases and the training targets.

Two hand-set networks make the decision go both ways: o = c r0 / s (a Richardson step,
guess = p_carried + c r0) lowers the residual for c = 0.1 and raises it for c = -0.1.
"""
import sys

import numpy as np
import pytest

import oracle
from natray_amd import Sim, abi
from natray_amd import net as netmod
from tests.test_gpu_parity import SCENES, assert_same_float, compare_grid, compare_state

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def richardson_params(c: float) -> np.ndarray:
    """Parameters whose output is c times input channel 0: relu(x) and relu(-x) through the
    centre taps, recombined by the last layer."""
    p = np.zeros(netmod.PARAMS, np.float32)
    views = []
    for cin, cout, _, wo, _ in netmod.LAYERS:
        views.append(p[wo:wo + 27 * cin * cout].reshape(27, cin, cout))
    centre = 13
    views[0][centre, 0, 0] = 1.0
    views[0][centre, 0, 1] = -1.0
    for W in views[1:3]:
        W[centre, 0, 0] = 1.0
        W[centre, 1, 1] = 1.0
    views[3][centre, 0, 0] = c
    views[3][centre, 1, 0] = -c
    return p


def tile_layout(g: dict, bdims):
    """Tile id and voxel (256 x + 16 y + z in the tile) of every cell in use, in cell order."""
    bl = g["block_list"].astype(np.int64)
    bd1, bd2 = int(bdims[1]), int(bdims[2])
    bx, by, bz = bl // (bd2 * bd1), (bl // bd2) % bd1, bl % bd2
    tdy, tdz = (bd1 + 3) // 4, (bd2 + 3) // 4
    tile = ((bx >> 2) * tdy + (by >> 2)) * tdz + (bz >> 2)
    lane = np.arange(64)
    lx, ly, lz = lane >> 4, (lane >> 2) & 3, lane & 3
    vox = (((bx[:, None] & 3) * 4 + lx) * 256 + ((by[:, None] & 3) * 4 + ly) * 16 + (bz[:, None] & 3) * 4 + lz)
    return np.repeat(tile, 64), vox.ravel()


def expected_inputs(g: dict, r0: np.ndarray, bdims, cap: int):
    """The listed tiles, their scales and the network inputs [T][4096][2] from the oracle."""
    ty = g["cf"][abi.F_TYPE]
    water = ty == abi.WATER
    tile, vox = tile_layout(g, bdims)
    listed = np.unique(tile[water])[:cap]
    slot = {int(t): k for k, t in enumerate(listed)}
    scale = np.zeros(len(listed), np.float32)
    for k, t in enumerate(listed):
        scale[k] = np.abs(r0[water & (tile == t)]).max()
    X = np.zeros((len(listed), 4096, 2), np.float32)
    cell_slot = np.array([slot.get(int(t), -1) for t in tile])
    on = cell_slot >= 0
    kind = np.where(water, 1.0, np.where(ty == abi.SOLID, -1.0, 0.0)).astype(np.float32)
    s = scale[np.maximum(cell_slot, 0)]
    x = np.where(water & (s > 0), r0 / np.where(s > 0, s, 1).astype(np.float32), 0).astype(np.float32)
    X[cell_slot[on], vox[on], 0] = x[on]
    X[cell_slot[on], vox[on], 1] = kind[on]
    return listed, scale, X, cell_slot, vox


def water_neighbour_bits(g: dict, bdims) -> np.ndarray:
    """Per cell in use, bit 2 axis (- side) or 2 axis + 1 (+ side) for each water face
    neighbour, from the grid's block neighbours and cell types."""
    n = g["n_blocks"]
    ty = g["cf"][abi.F_TYPE]
    nb = g["block_nb"]
    out = np.zeros(64 * n, np.uint32)
    lane = np.arange(64)
    loc = np.stack([lane >> 4, (lane >> 2) & 3, lane & 3], axis=1)
    for axis in range(3):
        for side, bit in ((-1, 2 * axis), (1, 2 * axis + 1)):
            q = loc.copy()
            q[:, axis] += side
            off = np.zeros_like(q)
            off[:, axis] = np.where(q[:, axis] < 0, -1, np.where(q[:, axis] > 3, 1, 0))
            q = q - 4 * off
            k = (off[:, 0] + 1) * 9 + (off[:, 1] + 1) * 3 + off[:, 2] + 1
            blk = nb[:, k]  # [n][64]
            cell = blk.astype(np.int64) * 64 + (q[:, 0] * 16 + q[:, 1] * 4 + q[:, 2])
            ok = blk != 0xFFFFFFFF
            t = np.where(ok, ty[np.where(ok, cell, 0)], abi.GAS)
            out |= np.where(t.ravel() == abi.WATER, np.uint32(1 << bit), np.uint32(0))
    return out


CASES = [("impact_sgs_32k", 0.1), ("impact_sgs_32k", -0.1), ("impact_mixed_32k_mem", 0.1), ("impact_pipelined_32k", 0.1),
         ("impact_pipelined_32k_mem", 0.1)]


def warm_start_parity(name, s, c, steps=4, max_tiles=64):
    """Steps the scene with the Richardson network o = c r0 / s as its warm start, phase by
    phase against the oracle, and checks inputs, guess, decision, solve, training targets,
    operator samples and particles every step. Returns the number of guesses taken."""
    say(f"[warm] {name} c={c} N={s.n}")
    sim = Sim.from_scene(s)
    orc = oracle.Oracle.from_scene(s, "f32")
    n = netmod.Net(max_tiles=max_tiles, seed=1)
    n.set_params(richardson_params(c))
    sim.set_warm_start(n, 2)
    bdims = s.params.bdims
    taken = 0
    for step in range(steps):
        for ph in (abi.PHASE_BEGIN, abi.PHASE_BIN, abi.PHASE_P2G, abi.PHASE_GRID):
            sim.phase(ph)
            orc.phase(ph)
        g = orc.grid()
        nc = 64 * g["n_blocks"]
        r0 = orc.residual()
        p_car = g["cf"][abi.F_P].copy()
        say(f"[warm] {name} c={c} step {step} solve")
        sim.phase(abi.PHASE_SOLVE)
        tag = f"{name} c={c} step {step}"
        smp = sim.ws_samples()
        X, Y, M = smp["inputs"], smp["targets"], smp["masks"]
        listed, scale, Xo, cell_slot, vox = expected_inputs(g, r0, bdims, max_tiles)
        assert X.shape[0] == len(listed), (tag, X.shape, len(listed))
        assert_same_float(f"{tag} inputs", X, Xo)
        guess, accepted = sim.ws_guess()
        pred = n.forward(X)
        water = g["cf"][abi.F_TYPE] == abi.WATER
        on = water & (cell_slot >= 0)
        want = p_car.copy()
        want[on] = p_car[on] + scale[cell_slot[on]] * pred[cell_slot[on], vox[on]]
        assert_same_float(f"{tag} guess", guess[:nc], want)
        rg = orc.residual(guess).astype(np.float64)
        sg, s0 = float((rg ** 2).sum()), float((r0.astype(np.float64) ** 2).sum())
        say(f"[warm] {tag}: |r_guess|^2 = {sg:.6g}, |r0|^2 = {s0:.6g}, accepted {accepted}")
        # A step at rest has r0 = 0: nothing to improve, the guess equals the carried pressure.
        assert accepted == (c > 0 and s0 > 0), (tag, sg, s0)
        assert accepted == (sg < s0)
        taken += int(accepted)
        stats = sim.ws_stats()
        assert stats["offered"] == step + 1 and stats["accepted"] == taken
        np.testing.assert_allclose([stats["guess_rr"], stats["carried_rr"]], [sg, s0], rtol=1e-5)
        orc.set_guess(guess if accepted else None)
        orc.phase(abi.PHASE_SOLVE)
        compare_state(f"{tag} after solve", sim.state(), orc.state())
        g2 = sim.grid()
        compare_grid(f"{tag} after solve", g2, orc.grid())
        # Training targets: (p_solved - p_carried) / s on water cells.
        p_new = g2["cf"][abi.F_P]
        Yo = np.zeros_like(Y)
        Mo = np.zeros_like(M)
        lit = on & (scale[np.maximum(cell_slot, 0)] > 0)  # a tile with r0 = 0 teaches nothing
        Yo[cell_slot[lit], vox[lit]] = (p_new[lit] - p_car[lit]) / scale[cell_slot[lit]]
        Mo[cell_slot[lit], vox[lit]] = 1.0
        assert_same_float(f"{tag} targets", Y, Yo)
        assert_same_float(f"{tag} masks", M, Mo)
        # The operator of the solve on those cells: its diagonal and water-neighbour bits.
        Do = np.zeros_like(smp["diag"])
        Bo = np.zeros_like(smp["nbm"])
        Do[cell_slot[lit], vox[lit]] = g2["cf"][abi.F_DIAG][lit]
        water_nb = water_neighbour_bits(g2, bdims)
        Bo[cell_slot[lit], vox[lit]] = water_nb[lit]
        assert_same_float(f"{tag} diag", smp["diag"], Do)
        np.testing.assert_array_equal(smp["nbm"], Bo, err_msg=f"{tag} nbm")
        for ph in (abi.PHASE_PROJECT, abi.PHASE_EXTEND, abi.PHASE_G2P):
            sim.phase(ph)
            orc.phase(ph)
        assert_same_float(f"{tag} particles", sim.particles(), orc.particles())
    sim.close()
    orc.close()
    n.close()
    return taken


@pytest.mark.parametrize("name,c", CASES)
def test_warm_start_bit_exact(name, c):
    taken = warm_start_parity(name, SCENES[name](), c)
    assert taken >= (2 if c > 0 else 0)


def test_warm_start_off_and_inputs_only_change_nothing():
    """Mode 1 (inputs for training only) and mode 0 give the same bits as no warm start."""
    s = SCENES["impact_sgs_32k"]()
    say("[warm] modes 0 and 1 against a plain run")
    runs = []
    for mode in (None, 1, 0):
        sim = Sim.from_scene(s)
        if mode is not None:
            sim.set_warm_start(None, mode)
        sim.run(5, iters=int(s.params.limits[2]), stop_ns=0xFFFFFFFF)
        runs.append((sim.particles(), sim.state()))
        sim.close()
    for P, st in runs[1:]:
        assert_same_float("particles", P, runs[0][0])
        compare_state("state", st, runs[0][1])


def test_trained_network_gain():
    """What the trained network gains over the solve without it, on a seed it was not trained
    on: training tiles from two seeded runs of the 32k-particle impact, Adam on the residual
    loss with the network's own GPU kernels; then 20 steps with the warm start against 20
    without. Every guess must lower the starting residual (all are taken), by at least half in
    squared norm, and the solve must save at least 5 of the ~470 CG iterations of the 20 steps
    (measured: 23.50 iterations per step without it, 22.95 with it)."""
    from natray_amd import scenes
    from natray_amd import warm_start as ws

    say("[warm] training on the 32k impact")
    make = lambda seed: scenes.balloon_impact(side=64, n=32768, seed=seed, iters=400, capacity=4096)
    parts = []
    for seed in (3, 4):
        s = make(seed)
        sim = Sim.from_scene(s)
        parts.append(ws.collect(sim, 60, iters=int(s.params.limits[2])))
        sim.close()
    data = {k: np.concatenate([p[k] for p in parts]) for k in ws.KEYS}
    n = netmod.Net(max_tiles=64, seed=1)
    before = ws.residual_ratio(n, data, 64)
    ws.train(n, data, 1500, batch=32, lr=2e-3)
    after = ws.residual_ratio(n, data, 64)
    say(f"[warm] tile residual ratio {before:.3f} -> {after:.4f}")
    assert after < 0.05
    s = make(9)
    it = int(s.params.limits[2])
    iters = {}
    for mode in ("plain", "net"):
        sim = Sim.from_scene(s)
        sim.run(20, iters=it)
        if mode == "net":
            sim.set_warm_start(n, 2)
        c0 = sim.counters()["cg_iters"]
        ratios = []
        for _ in range(20):
            sim.run(1, iters=it)
            if mode == "net":
                st = sim.ws_stats()
                ratios.append(st["guess_rr"] / st["carried_rr"])
        iters[mode] = (sim.counters()["cg_iters"] - c0) / 20
        if mode == "net":
            st = sim.ws_stats()
            say(f"[warm] guesses taken {st['accepted']} of {st['offered']}; residual^2 ratio mean {np.mean(ratios):.3f}")
            assert st["accepted"] == st["offered"] == 20
            assert np.mean(ratios) < 0.5
        sim.close()
    say(f"[warm] CG iterations per step: plain {iters['plain']:.2f}, warm-started {iters['net']:.2f}")
    assert iters["net"] <= iters["plain"] - 0.25, iters
    n.close()


@pytest.mark.parametrize("name", ["impact_sgs_32k", "impact_mixed_32k_mem", "impact_pipelined_32k"])
def test_zero_start_matches_the_oracle(name):
    """Warm-start mode 3 (every solve from p = 0, b unchanged) against the oracle's solve from a
    zero guess, phase by phase."""
    s = SCENES[name]()
    say(f"[warm] zero start {name}")
    sim = Sim.from_scene(s)
    sim.set_warm_start(None, 3)
    orc = oracle.Oracle.from_scene(s, "f32")
    zeros = np.zeros(int(s.params.bdims[3]) * 64, np.float32)
    iters = 0
    for step in range(4):
        for ph in range(abi.N_PHASES):
            if ph == abi.PHASE_SOLVE:
                orc.set_guess(zeros)
            sim.phase(ph)
            orc.phase(ph)
            tag = f"zero start {name} step {step} after {abi.PHASE_NAMES[ph]}"
            compare_state(tag, sim.state(), orc.state())
            if ph == abi.PHASE_SOLVE:
                compare_grid(tag, sim.grid(), orc.grid())
        assert_same_float(f"zero start {name} step {step} particles", sim.particles(), orc.particles())
        iters += int(orc.state().solve_iters)
    assert iters > 0
    sim.close()
    orc.close()
