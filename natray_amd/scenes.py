"""Scene construction: the reference balloon's domain, particles and Params, and the synthetic
benchmark scenes.

Reference construction follows Sim::new (src/sim.rs:208-237,428-441) and the water fill of
balloon::build (src/balloon.rs:419-445). The balloon's inflated shape comes from the
membrane's inflation solve (src/balloon.rs:119-248), which belongs to the membrane and is not
part of this library; `SphereProfile` stands in for it (exact for the weightless preset, whose
balloon is round, src/balloon.rs:497-506).

The synthetic scenes are the benchmark workloads: a water sphere made of the N/8 cells nearest
the box centre, 8 particles per cell at uniform in-cell offsets from a splitmix64 stream, in a
unit box with solid walls; optionally a rigid capsule moving through it and a fixed spherical
mask in place of the membrane.
"""
from __future__ import annotations

import dataclasses as dc
import math

import numpy as np

from . import abi
from .params import SimParams

# src/sim.rs:13-47
SPRAY_CAP = 1_200_000
WET_RES = 1024
WET_HALF = 1.2
FRICTION = 0.8


@dc.dataclass
class Scene:
    """Everything natray_sim_create takes."""

    params: abi.Params
    particles: np.ndarray  # (N, 20) float32
    scene: abi.Scene
    init: abi.State | None = None
    meta: dict = dc.field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.particles.shape[0])


def next_multiple_of(x: int, m: int) -> int:
    return (x + m - 1) // m * m


# ---- reference construction ---------------------------------------------------------------


class Rng:
    """xorshift64 of src/balloon.rs:250-258."""

    def __init__(self, seed: int = 0x9E3779B97F4A7C15):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def next(self) -> float:
        M = 0xFFFFFFFFFFFFFFFF
        s = self.s
        s ^= (s << 13) & M
        s ^= s >> 7
        s ^= (s << 17) & M
        self.s = s
        return (s >> 11) / float(1 << 53)


# The membrane mesh is an icosphere of level 6, 10 * 4^6 + 2 vertices, and each vertex draws
# two numbers from the generator before the water does (src/balloon.rs:357-368).
MESH_DRAWS = 2 * (10 * 4**6 + 2)


@dc.dataclass
class SphereProfile:
    """A round balloon standing in for the inflated profile (src/balloon.rs Profile)."""

    radius: float
    neck_r: float

    @property
    def z_neck(self) -> float:
        return 2.0 * self.radius

    @property
    def z_centroid(self) -> float:
        return self.radius

    def z_top(self) -> float:
        return 2.0 * self.radius

    def r_at(self, z: float) -> float:
        q = self.radius * self.radius - (z - self.radius) ** 2
        return math.sqrt(q) if q > 0.0 else 0.0


def balloon_particles(p: SimParams, profile, bottom_y: float, radius_max: float) -> tuple[np.ndarray, float]:
    """The water fill of balloon::build (src/balloon.rs:419-445): a lattice with odd layers
    offset by half a spacing, jittered by +-0.15 spacing, inside r(z) - spacing/2."""
    d = p.balloon_diameter_cm * 0.01
    spacing = d / max(p.resolution, 16)
    rng = Rng()
    for _ in range(MESH_DRAWS):
        rng.next()
    out = []
    n = math.ceil(2.0 * radius_max / spacing) + 2
    nz = math.ceil(profile.z_neck / spacing) + 1
    half = int(n / 2)  # Rust integer division truncates toward zero
    for iy in range(nz):
        zl = (iy + 0.5) * spacing
        if zl > profile.z_neck - spacing:
            break
        rmax = profile.r_at(zl) - 0.5 * spacing
        if rmax <= 0.0:
            continue
        for ix in range(-half, half + 1):
            for iz in range(-half, half + 1):
                x = (ix + 0.5 * (iy % 2)) * spacing
                z = (iz + 0.5 * (iy % 2)) * spacing
                jx = (rng.next() - 0.5) * 0.3 * spacing
                jy = (rng.next() - 0.5) * 0.3 * spacing
                jz = (rng.next() - 0.5) * 0.3 * spacing
                if math.sqrt(x * x + z * z) < rmax:
                    out.append((x + jx, bottom_y + zl + jy, z + jz, 1.0))
    return np.asarray(out, dtype=np.float32).reshape(-1, 4), spacing


def pack_particles(xyzw: np.ndarray) -> np.ndarray:
    """Packs particles into 80-byte records with vel.w = 1 (src/sim.rs:227-236)."""
    P = np.zeros((xyzw.shape[0], abi.PARTICLE_FLOATS), dtype=np.float32)
    P[:, :4] = xyzw
    P[:, 7] = 1.0
    return P


def reference_scene(p: SimParams, capacity: int = abi.MAX_BLOCKS) -> Scene:
    """Sim::new's CPU half (src/sim.rs:208-237,428-441) with SphereProfile for the shape."""
    d = p.balloon_diameter_cm * 0.01
    prof = SphereProfile(radius=0.5 * d, neck_r=0.05 * d)
    bottom_y = p.balloon_height - prof.z_centroid
    xyzw, spacing = balloon_particles(p, prof, bottom_y, prof.radius)
    dx = 2.0 * spacing
    top = bottom_y + prof.z_top() + 0.6
    nxz = next_multiple_of(math.ceil(2.0 * WET_HALF / dx), 4)
    ny = next_multiple_of(math.ceil(top / dx) + 3, 4)
    half = np.float32(nxz * 0.5 * dx)
    U = abi.Params()
    U.origin[:] = [-half, np.float32(-3.0 * dx), -half, np.float32(dx)]
    U.dims[:] = [nxz, ny, nxz, xyzw.shape[0]]
    U.bdims[:] = [nxz // 4, ny // 4, nxz // 4, capacity]
    U.gravity[:] = [0.0, -p.gravity, 0.0, abi.SOUND]
    U.fluid[:] = [abi.RHO_WATER, abi.CFL, abi.DT_MAX, spacing]
    U.wet[:] = [WET_HALF, 0.0, 0.0, 0.0]
    U.limits[:] = [WET_RES, SPRAY_CAP, 0, 0]
    U.solve[:] = [abi.SOLVE_TOL, 0.0, 0.0, 0.0]
    # No rubber vertices and no bullets: the membrane and the bullets are not simulated here.
    U.mem[:] = [0, 2, 2, 0]
    U.rubber[:] = [0.0, 0.0, 0.0, p.rubber_thickness_mm * 1e-3]
    U.rubber2[:] = [FRICTION, p.rubber_break_stretch, 0.0, 0.0]
    U.knot[:] = [0.0, bottom_y + prof.z_neck, 0.0, prof.neck_r]
    sc = abi.Scene()
    sc.flags = abi.SCENE_REFERENCE
    sc.precond = abi.PRECOND_BLOCK_SGS
    sc.precision = abi.CG_FP32
    return Scene(U, pack_particles(xyzw), sc, None, {"workload": "reference balloon", "spacing": spacing, "dx": dx})


# ---- synthetic benchmark scenes -----------------------------------------------------------------

GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(seed: int, count: int) -> np.ndarray:
    """`count` outputs of splitmix64 seeded with `seed`, as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    return (splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def nearest_cells(side: int, n_cells: int, centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """The n_cells cells of a side^3 unit-box grid whose centres are nearest `centre`, ties
    broken by linear index; returned as (n, 3) int coordinates."""
    c = np.asarray(centre) * side
    r = (3.0 * n_cells / (4.0 * math.pi)) ** (1.0 / 3.0) + 3.0
    lo = np.maximum(np.floor(c - r).astype(int), 0)
    hi = np.minimum(np.ceil(c + r).astype(int) + 1, side)
    ax = [np.arange(lo[k], hi[k]) for k in range(3)]
    X, Y, Z = np.meshgrid(*ax, indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    # Squared distances times 4, exactly, in integers (centres at i + 1/2).
    cc = np.round(2.0 * c).astype(np.int64)
    d2 = (2 * X + 1 - cc[0]) ** 2 + (2 * Y + 1 - cc[1]) ** 2 + (2 * Z + 1 - cc[2]) ** 2
    lin = (X.astype(np.int64) * side + Y) * side + Z
    order = np.lexsort((lin, d2))[:n_cells]
    if len(order) < n_cells:
        raise ValueError("sphere does not fit the box")
    return np.stack([X[order], Y[order], Z[order]], axis=1)


def cells_within(side: int, radius: float, centre=(0.5, 0.5, 0.5)) -> int:
    """How many cells of a side^3 unit-box grid have their centre closer than `radius` (in
    box units) to `centre`: the cells a sphere of that radius fills at 8 particles per cell."""
    c2 = np.round(2.0 * np.asarray(centre) * side).astype(np.int64)
    i = 2 * np.arange(side, dtype=np.int64) + 1
    dx2, dy2, dz2 = (i - c2[0]) ** 2, (i - c2[1]) ** 2, (i - c2[2]) ** 2
    lim = (2.0 * radius * side) ** 2  # squared distances times 4, as in nearest_cells
    return int(sum(int(np.count_nonzero(dx2[x] + dy2[:, None] + dz2[None, :] < lim)) for x in range(side)))


def block_order(cells: np.ndarray, side: int) -> np.ndarray:
    """Sorts cells by 4^3 block (slot order), then by cell within the block, so particles of
    one block sit together in memory."""
    b = cells >> 2
    l = cells & 3
    bd = side // 4
    slot = (b[:, 0] * bd + b[:, 1]) * bd + b[:, 2]
    loc = l[:, 0] * 16 + l[:, 1] * 4 + l[:, 2]
    return cells[np.lexsort((loc, slot))]


def sphere_particles(side: int, n: int, seed: int, centre=(0.5, 0.5, 0.5)) -> np.ndarray:
    """n particles, 8 per cell, in the n/8 cells nearest the centre of a unit box of side^3
    cells, each at independent U[0,1)^3 in-cell offsets."""
    assert n % 8 == 0
    cells = block_order(nearest_cells(side, n // 8, centre), side)
    u = uniform01(seed, 3 * n).reshape(n, 3)
    base = np.repeat(cells, 8, axis=0).astype(np.float64)
    pos = (base + u) / side
    P = np.zeros((n, abi.PARTICLE_FLOATS), dtype=np.float32)
    P[:, :3] = pos.astype(np.float32)
    P[:, 3] = 1.0
    P[:, 7] = 1.0
    return P


def box_params(side: int, n: int, capacity: int, gravity: float = 9.81, cfl: float = 0.5, tol: float = 1e-6,
               iters: int = 200) -> abi.Params:
    """Params of a unit box of side^3 cells (dx = 1/side), particle spacing dx/2."""
    dx = np.float32(1.0 / side)
    U = abi.Params()
    U.origin[:] = [0.0, 0.0, 0.0, dx]
    U.dims[:] = [side, side, side, n]
    U.bdims[:] = [side // 4, side // 4, side // 4, capacity]
    U.gravity[:] = [0.0, -gravity, 0.0, abi.SOUND]
    U.fluid[:] = [abi.RHO_WATER, cfl, abi.DT_MAX, np.float32(0.5 * dx)]
    U.wet[:] = [WET_HALF, 0.0, 0.0, 0.0]
    U.limits[:] = [WET_RES, SPRAY_CAP, iters, 0xFFFFFFFF]
    U.solve[:] = [tol, 0.0, 0.0, 0.0]
    U.mem[:] = [0, 2, 2, 0]
    U.knot[:] = [0.0, 0.0, 0.0, 0.0]
    return U


def write_scene_file(s: Scene, path: str, bullets=None) -> None:
    """Writes the scene for natray_drive: b"NATRAYSC", u32 version 1, u32 particle count,
    Params, natray_scene, State, then the 80-byte particle records. With bullets = (list of
    abi.Bullet, patches [n][576][8]), a trailer follows: b"BULLETS\0", u32 n, the n 112-byte
    records and the patches."""
    import ctypes as C
    import struct

    init = s.init if s.init is not None else no_limits_state()
    with open(path, "wb") as f:
        f.write(b"NATRAYSC")
        f.write(struct.pack("<II", 1, s.n))
        f.write(bytes(C.string_at(C.addressof(s.params), C.sizeof(s.params))))
        f.write(bytes(C.string_at(C.addressof(s.scene), C.sizeof(s.scene))))
        f.write(bytes(C.string_at(C.addressof(init), C.sizeof(init))))
        f.write(np.ascontiguousarray(s.particles, np.float32).tobytes())
        if bullets:
            bl, pt = bullets
            pt = np.ascontiguousarray(pt, np.float32)
            assert pt.shape == (len(bl), abi.BULLET_PATCHES, 8)
            f.write(b"BULLETS\0")
            f.write(struct.pack("<I", len(bl)))
            for b in bl:
                f.write(bytes(b))
            f.write(pt.tobytes())


def no_limits_state() -> abi.State:
    """First State with no rubber or air step limit (neither is simulated)."""
    st = abi.State()
    st.mem_dt = np.finfo(np.float32).max
    st.cap_dt = np.finfo(np.float32).max
    st.air_dt = np.finfo(np.float32).max
    return st


def water_sphere(side: int = 64, n: int = 32768, seed: int = 1, fixed_dt: float = 1e-4, precond: int = abi.PRECOND_JACOBI,
                 tol: float = 1e-6, iters: int = 200, capacity: int | None = None) -> Scene:
    """The plain water sphere: n particles in a unit box of side^3 cells, solid walls, zero
    initial velocity, gravity 9.81, fixed dt, Jacobi CG with tolerance 1e-6 and at most 200
    iterations. With the defaults: 32,768 particles on 64^3, dt = 1e-4."""
    cap = capacity or min((side // 4) ** 3, abi.MAX_BLOCKS)
    U = box_params(side, n, cap, tol=tol, iters=iters)
    sc = abi.Scene()
    sc.flags = abi.SCENE_WALLS
    sc.precond = precond
    sc.precision = abi.CG_FP32
    sc.fixed_dt = fixed_dt
    P = sphere_particles(side, n, seed)
    r_eff = (3.0 * (n // 8) / (4.0 * math.pi)) ** (1.0 / 3.0) / side
    return Scene(U, P, sc, no_limits_state(),
                 {"workload": f"water sphere {n} particles {side}^3", "r_eff": r_eff, "dx": 1.0 / side})


def breakup_scene(side: int = 32, n_ball: int = 8 * 400, n_slow: int = 48, n_fast: int = 32, n_ground: int = 32,
                  seed: int = 9, fixed_dt: float = 1e-3, spray_cap: int = 4096) -> Scene:
    """Spray and wet ground (NATRAY_SCENE_SPRAY): a water ball in a solid-walled box plus lone
    particles, which leave the resolved flow at the first G2P (a lone particle's own weights
    give it a fill near 0.04): `n_slow` 0.10-0.18 m up falling at 15-25 m/s (drops of ~0.5 mm
    that land within ~10 steps), `n_fast` high up at 40-60 m/s sideways (the fine
    spray branch, a parcel of drops, with strong drag), and `n_ground` below 1.5 spacings
    (they wet the ground instead). Fixed dt; a small spray capacity."""
    base = water_sphere(side=side, n=n_ball, seed=seed, fixed_dt=fixed_dt)
    U, sc = base.params, base.scene
    dx = 1.0 / side
    spacing = 0.5 * dx
    rng = np.random.default_rng(seed)
    groups = []
    for count, ylo, yhi, vlo, vhi in ((n_slow, 0.10, 0.18, (-2.0, -25.0, -2.0), (2.0, -15.0, 2.0)),
                                      (n_fast, 0.3, 0.6, (40.0, -1.0, -5.0), (60.0, 1.0, 5.0)),
                                      (n_ground, 0.2 * spacing, 1.2 * spacing, (-0.5, -0.5, -0.5), (0.5, 0.0, 0.5))):
        P = np.zeros((count, abi.PARTICLE_FLOATS), np.float32)
        P[:, 0] = rng.uniform(0.25, 0.6, count)
        P[:, 1] = rng.uniform(ylo, yhi, count)
        P[:, 2] = rng.uniform(0.25, 0.75, count)
        P[:, 3] = 1.0
        P[:, 4:7] = rng.uniform(vlo, vhi, (count, 3))
        P[:, 7] = 1.0
        groups.append(P)
    P = np.concatenate([base.particles] + groups)
    U.dims[3] = P.shape[0]
    U.limits[1] = spray_cap
    sc.flags |= abi.SCENE_SPRAY
    return Scene(U, P, sc, no_limits_state(),
                 {"workload": f"breakup: water ball {n_ball} + {n_slow}/{n_fast}/{n_ground} lone particles {side}^3", "dx": dx})


def balloon_impact(side: int = 256, n: int = 2_097_152, seed: int = 2, speed: float = 300.0, radius: float = 0.01,
                   half_length: float = 0.015, precond: int = abi.PRECOND_BLOCK_SGS, precision: int = abi.CG_FP32,
                   refresh_every: int = 20, tol: float = 1e-6, iters: int = 1000, capacity: int = abi.MAX_BLOCKS,
                   cfl: float = 0.5, direction=(1.0, 0.0, 0.0), mask: bool = True) -> Scene:
    """The balloon impact: n particles (8 per cell) in a water sphere at the centre of a unit
    box of side^3 cells with solid walls; a rigid capsule of the given radius moving at `speed`
    along `direction` starts two cells short of the water on the -direction side and enters it
    as a moving solid; a fixed spherical mask at the water's radius stands in for the
    membrane; dt from CFL `cfl`. With the defaults: 2,097,152 particles on 256^3."""
    U = box_params(side, n, capacity, cfl=cfl, tol=tol, iters=iters)
    P = sphere_particles(side, n, seed)
    dx = 1.0 / side
    r_eff = (3.0 * (n // 8) / (4.0 * math.pi)) ** (1.0 / 3.0) * dx
    dvec = np.asarray(direction, dtype=np.float64)
    dvec = dvec / np.linalg.norm(dvec)
    centre = np.array([0.5, 0.5, 0.5])
    nose = centre - dvec * (r_eff + 2.0 * dx)
    b = nose - dvec * radius
    a = b - dvec * (2.0 * half_length)
    sc = abi.Scene()
    sc.flags = abi.SCENE_WALLS | abi.SCENE_CAPSULE | (abi.SCENE_MASK if mask else 0)
    sc.precond = precond
    sc.precision = precision
    sc.refresh_every = refresh_every if precision != abi.CG_FP32 else 0
    sc.capsule_a[:] = [a[0], a[1], a[2], radius]
    sc.capsule_b[:] = [b[0], b[1], b[2], 0.0]
    sc.capsule_v[:] = [speed * dvec[0], speed * dvec[1], speed * dvec[2], 0.0]
    sc.mask[:] = [0.5, 0.5, 0.5, r_eff]
    sc.mask_p = 0.0
    return Scene(U, P, sc, no_limits_state(),
                 {"workload": f"balloon impact {n} particles {side}^3", "r_eff": r_eff, "dx": dx})


def stated_radius_impact(side: int = 256, radius: float = 0.3, **kw) -> Scene:
    """The balloon impact with the sphere of the radius the workload states (0.3 on 256^3)
    filled at 8 particles per cell, instead of the stated particle count: every cell whose
    centre lies within the radius holds 8 particles (15.2M at the defaults)."""
    n = 8 * cells_within(side, radius)
    s = balloon_impact(side=side, n=n, **kw)
    s.meta["workload"] = f"balloon impact, stated radius {radius} at 8 particles per cell: {n} particles {side}^3"
    s.meta["radius"] = radius
    return s


def stated_radius_sphere(side: int = 64, radius: float = 0.25, **kw) -> Scene:
    """The water sphere with the radius the workload states (0.25 on 64^3) filled at 8 particles
    per cell (137k particles at the defaults) instead of the stated count."""
    n = 8 * cells_within(side, radius)
    s = water_sphere(side=side, n=n, **kw)
    s.meta["workload"] = f"water sphere, stated radius {radius} at 8 particles per cell: {n} particles {side}^3"
    s.meta["radius"] = radius
    return s


def batch_impacts(count: int = 64, seed: int = 4, n_lo: int = 16384, n_hi: int = 524288, side_lo: int = 64,
                  side_hi: int = 192, speed_lo: float = 200.0, speed_hi: float = 400.0, precision: int = abi.CG_MIXED,
                  iters: int = 1000, refresh_every: int = 20) -> list[Scene]:
    """The batched scenes: `count` independent balloon impacts. Scene k draws from its own
    seed (splitmix64 of `seed`, k-th value): a particle count log-uniform in [n_lo, n_hi]
    (multiple of 8), a grid side uniform in [side_lo, side_hi] (multiple of 4), a capsule
    speed uniform in [speed_lo, speed_hi] and a direction uniform on the sphere. FP32 state;
    the CG precision is `precision` (mixed, refresh every `refresh_every` = 20 iterations, by default). With the
    defaults: 64 scenes of 16k..512k particles on 64^3..192^3."""
    seeds = splitmix64(seed, count)
    out = []
    for k in range(count):
        u = uniform01(int(seeds[k]), 5)
        n = 8 * max(1, int(round(math.exp(math.log(n_lo) + u[0] * (math.log(n_hi) - math.log(n_lo))) / 8)))
        side = 4 * int(round((side_lo + u[1] * (side_hi - side_lo)) / 4))
        # The water ball and the capsule's start must fit inside the walls.
        while (3.0 * (n // 8) / (4.0 * math.pi)) ** (1.0 / 3.0) + 10 > 0.5 * side:
            side += 4
        speed = speed_lo + u[2] * (speed_hi - speed_lo)
        z = 2.0 * u[3] - 1.0
        phi = 2.0 * math.pi * u[4]
        rho = math.sqrt(max(0.0, 1.0 - z * z))
        direction = (rho * math.cos(phi), rho * math.sin(phi), z)
        cap = min((side // 4) ** 3, 8 * (n // 512) + 2048)
        s = balloon_impact(side=side, n=n, seed=int(seeds[k] & 0xFFFFFFFF), speed=speed, direction=direction,
                           precision=precision, iters=iters, capacity=cap, refresh_every=refresh_every)
        s.meta.update({"scene": k, "speed": speed, "direction": direction, "capacity": cap})
        out.append(s)
    return out


def membrane_shell(s: Scene, n_verts: int = 2000, gap_cells: float = 0.6, push: float = 2e-3, seed: int = 7,
                   extra_off_water: int = 8):
    """A stand-in for the balloon's rubber around a water sphere: n_verts vertices on a
    Fibonacci sphere `gap_cells` cells outside the water's radius, each with mass 1e-4 kg, its
    share of the sphere's area and a push of `push` newtons, small seeded velocities, plus
    `extra_off_water` vertices on a ring 0.06 m further out, away from the water. Returns (x [n][4], v [n][4],
    mass [n], push [n][2])."""
    dx = float(s.meta["dx"])
    r = float(s.meta["r_eff"]) + gap_cells * dx
    k = np.arange(n_verts, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * k / n_verts
    phi = math.pi * (3.0 - math.sqrt(5.0)) * k
    rho = np.sqrt(1.0 - z * z)
    pts = np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)
    x = np.zeros((n_verts + extra_off_water, 4), np.float32)
    x[:n_verts, :3] = 0.5 + r * pts
    theta = 2.0 * math.pi * np.arange(extra_off_water) / max(extra_off_water, 1)
    ring = r + 0.06
    x[n_verts:, 0] = 0.5 + ring * np.cos(theta)
    x[n_verts:, 1] = 0.5 + ring * np.sin(theta)
    x[n_verts:, 2] = 0.5
    x[:, 3] = 1.0
    u = uniform01(seed, 3 * (n_verts + extra_off_water)).reshape(-1, 3)
    v = np.zeros_like(x)
    v[:, :3] = (0.05 * (2.0 * u - 1.0)).astype(np.float32)
    mass = np.full(n_verts + extra_off_water, 1e-4, np.float32)
    area = 4.0 * math.pi * r * r / n_verts
    p = np.zeros((n_verts + extra_off_water, 2), np.float32)
    p[:, 0] = push
    p[:, 1] = area
    return x, v, mass, p


def bullet_surface(length: float, radius: float, nose: float, u_cg: float, slices: int = 32, sectors: int = 16):
    """A bullet's outline and surface in its body frame, as src/bullet.rs:58-124 cuts them: the
    outline (a cylinder of `radius` with a cone of length `nose` at the front, base at u = 0,
    tip at u = length) resampled at `slices` equal steps of arc length, each band between two
    outline points cut into `sectors` patches at the band's mean radius, and a disk of sectors
    at the flat base (centroids two thirds out). Returns (patches [576][8] float32: position
    relative to the centre of mass at u_cg and area, outward normal and 0; count; hull, the
    farthest surface point from the centre of mass)."""
    def profile(u):
        return radius if u <= length - nose else radius * (length - u) / nose
    fine = np.linspace(0.0, length, 4001)
    line = np.array([[u, profile(u)] for u in fine])
    arc = np.concatenate([[0.0], np.cumsum(np.hypot(np.diff(line[:, 0]), np.diff(line[:, 1])))])
    target = arc[-1] * np.arange(slices + 1) / slices
    pts = np.stack([np.interp(target, arc, line[:, 0]), np.interp(target, arc, line[:, 1])], axis=1)
    rows = []

    def ring(u, r, radial, axial, area):
        for j in range(sectors):
            phi = (j + 0.5) * 2.0 * math.pi / sectors
            c, sn = math.cos(phi), math.sin(phi)
            rows.append([r * c, r * sn, u - u_cg, area / sectors, radial * c, radial * sn, axial, 0.0])

    for (ua, ra), (ub, rb) in zip(pts[:-1], pts[1:]):
        seg = math.hypot(ub - ua, rb - ra)
        if seg <= 0.0:
            continue
        r = 0.5 * (ra + rb)
        ring(0.5 * (ua + ub), r, (ub - ua) / seg, -(rb - ra) / seg, 2.0 * math.pi * r * seg)
    ring(0.0, 2.0 / 3.0 * radius, 0.0, -1.0, math.pi * radius * radius)
    assert len(rows) <= abi.BULLET_PATCHES
    out = np.zeros((abi.BULLET_PATCHES, 8), np.float32)
    out[: len(rows)] = np.array(rows, np.float32)
    hull = float(np.max(np.hypot(pts[:, 0] - u_cg, pts[:, 1])))
    return out, len(rows), hull


def rotation_arc(n) -> tuple:
    """The shortest rotation taking +z to unit vector n, as (x, y, z, w)."""
    n = np.asarray(n, np.float64) / np.linalg.norm(n)
    c = np.cross([0.0, 0.0, 1.0], n)
    q = np.array([c[0], c[1], c[2], 1.0 + n[2]])
    return tuple(q / np.linalg.norm(q))


def bullet_shot(at=(0.25, 0.5, 0.5), direction=(1.0, 0.0, 0.0), speed: float = 300.0, arrive_ns: int = 0,
                mass: float = 0.0042, radius: float = 0.0045, length: float = 0.02, nose: float = 0.008,
                spin_rev_s: float = 3000.0, yaw: float = 0.02):
    """One bullet as the flight hands it over (the role of src/bullet.rs:127-144 with a flight
    of our own): at `at` moving along `direction` at `speed`, nose first, arriving at `arrive_ns`,
    spinning at `spin_rev_s` about its axis, with a transverse angular momentum of `yaw` times
    the spin's (so the axis cones). A solid of revolution of `mass`: moments of inertia of a
    cylinder of `radius` and `length`. Returns (abi.Bullet, patches [576][8])."""
    patches, count, hull = bullet_surface(length, radius, nose, 0.45 * length)
    ix = 0.5 * mass * radius * radius
    iy = mass * (3.0 * radius * radius + length * length) / 12.0
    n = np.asarray(direction, np.float64) / np.linalg.norm(direction)
    q = rotation_arc(n)
    side = np.cross(n, [0.0, 1.0, 0.0])
    side = side / np.linalg.norm(side) if np.linalg.norm(side) > 0 else np.array([1.0, 0.0, 0.0])
    spin = ix * 2.0 * math.pi * spin_rev_s
    h = spin * n + yaw * spin * side
    b = abi.Bullet()
    b.x[:] = [at[0], at[1], at[2], mass]
    b.v[:] = [speed * n[0], speed * n[1], speed * n[2], 0.0]
    b.q[:] = list(q)
    b.l[:] = [h[0], h[1], h[2], speed]
    b.shape[:] = [ix, iy, hull, radius]
    b.info[:] = [arrive_ns, count, 0, 0]
    return b, patches


def bullet_scene(side: int = 64, n: int = 32768, seed: int = 1) -> Scene:
    """A water sphere (32,768 particles on 64^3, solid walls, Jacobi CG) whose step follows the
    CFL and the bullets' pace (no fixed dt), for bullet_shot's bullets."""
    s = water_sphere(side=side, n=n, seed=seed, fixed_dt=0.0)
    s.meta["workload"] = f"bullet into a water sphere {n} particles {side}^3"
    return s
