/* CPU oracle of the natray water substep — TEST INFRASTRUCTURE ONLY (see natray_oracle.h).
 *
 * Compiled with -DREAL=float -DPFX=o32_ (FP32-faithful) and -DREAL=double -DPFX=o64_ (FP64).
 * Build with -ffp-contract=off and without -ffast-math: every f32 expression rounds the way
 * the WGSL source writes it, left to right.
 *
 * Each function below restates one reference kernel (file:line given at the function). The
 * deliberate, documented differences from the reference are:
 *   - block order: allocated blocks are listed in ascending domain-slot order, not in the
 *     order atomic claims happen to win (fluid.wgsl:109-125). The reference's order is
 *     nondeterministic; every result that does not depend on float summation order is the
 *     same for any order.
 *   - ranks inside a block: particle index order (the reference's atomicAdd order is
 *     nondeterministic; nothing downstream depends on it).
 *   - solver reductions over blocks: per-block tree as block_sum (fluid.wgsl:414-422), then
 *     fixed chunks of K = ceil(n/256) consecutive blocks (at most 256 chunks), each summed as a
 *     binary tree over the next power of two >= K slots, then the chunk sums go through the
 *     reference's total() (fluid.wgsl:426-439). For n <= 256 this is exactly the reference's
 *     total() over blocks in slot order.
 *   - the membrane coupling, the air boundary, the bullets and the spray are restated below
 *     and are off unless set (set_membrane, NATRAY_SCENE_AIR, set_bullets, NATRAY_SCENE_SPRAY):
 *     off, they behave as the reference without them (GAS pressure 0, their acc fields zero,
 *     carve, break_up and ground_touch no-ops, death only clearing vel.w). The rubber's contact
 *     with a bullet (bullet_hit) belongs to the membrane's own substeps and is always zero.
 *   - scene extensions (box walls, moving capsule solid, spherical mask, fixed dt, Jacobi
 *     preconditioner, mixed-precision CG) are defined here first and the HIP kernels follow.
 *
 * What pins this restatement (tests/test_oracle_*.py): brute-force NumPy goldens for P2G
 * fixed-point sums and G2P gathers on small inputs (tests/golden/), closed forms (B-spline
 * partition of unity, vanishing first moment, APIC affine reproduction, free-fall grid
 * velocity), binning invariants against brute force, the CG solution against a direct
 * scipy solve of the same matrix, symmetry and definiteness of the block-SGS
 * preconditioner probed with unit vectors, and the begin_step clock against a table; the
 * velocity extension by hand-worked means on built face patterns (test_oracle_extend.py);
 * the capsule's solid cells, its face velocity in `bounded`, its `carve` and the spherical
 * mask by float64 geometry and closed forms (test_oracle_capsule.py).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "natray_oracle.h"
/* The spray's exp, log, cos and pow: the kernels' own functions, so both give the same bits;
 * they are pinned against the C library by tests/test_oracle_spray.py. */
#include "natray_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif
#ifndef PFX
#define PFX o32_
#endif
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(PFX, name)

#define IS_F64 (sizeof(REAL) == 8)

static inline REAL rfloor(REAL x) { return IS_F64 ? (REAL)floor((double)x) : (REAL)floorf((float)x); }
static inline REAL sqrt_r(REAL x) { return IS_F64 ? (REAL)sqrt((double)x) : (REAL)sqrtf((float)x); }
static inline REAL rrint(REAL x) { return IS_F64 ? (REAL)rint((double)x) : (REAL)rintf((float)x); }
static inline REAL rmin(REAL a, REAL b) { return b < a ? b : a; }
static inline REAL rmax(REAL a, REAL b) { return a < b ? b : a; }
static inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline uint32_t umax(uint32_t a, uint32_t b) { return a < b ? b : a; }

typedef struct { int x, y, z; } i3;
typedef struct { REAL x, y, z; } r3;

#define GAS ((REAL)0)
#define WATER ((REAL)1)
#define SOLID ((REAL)2)
#define COVERED ((REAL)3)
#define RUBBER ((REAL)4)
#define NONE NATRAY_NONE
#define PENDING NATRAY_PENDING
#define TILE 343u

struct osim {
  natray_params U;
  natray_scene sc;
  natray_state st; /* integer members are authoritative; float members mirror the REALs below */
  double max_speed, rz, alpha, beta, b_norm, r_norm;
  int mixed;   /* dot products, norms and the convergence test in double */
  int refresh; /* recompute r = b - A p every sc.refresh_every iterations */
  uint32_t N, S, cap;
  REAL* P;               /* N x 20 */
  uint32_t* block_ptr;   /* S */
  uint32_t* slot_count;  /* S */
  uint32_t* block_list;  /* cap */
  uint32_t* block_nb;    /* cap x 27 */
  uint32_t* block_start; /* cap + 1 */
  uint32_t* sorted;      /* N */
  uint32_t* rank;        /* N */
  int32_t* acc;          /* 19 x cap x 64 */
  REAL* cf;              /* 11 x cap x 64 */
  REAL* bvec;            /* cap x 64: right-hand side kept for the true-residual refresh */
  double* part;          /* cap x 4: per-block partial sums */
  uint64_t cg_iters;
  REAL* guess;           /* cap x 64: a starting pressure for the next solve (set_guess) */
  int use_guess;
  /* The pipelined CG's vectors outside cf (cap x 64 each): z, q and m (see o_solve_pipelined). */
  REAL *pz, *pq, *pm;
  /* The membrane's vertices (set_membrane): positions and velocities (4 per vertex), masses,
   * push and area (2 per vertex); mem_water (4 per vertex) is what mem_sample leaves. */
  uint32_t mem_cap;
  REAL *mem_x, *mem_v, *mem_mass, *mem_push, *mem_water;
  /* The air boundary (NATRAY_SCENE_AIR): gas per cell [5][cap x 64] after the step and as
   * carried into the step's blocks, the last step's block list, each slot's index in it, and
   * the blocks whose air was active. */
  REAL *air_x, *air_tmp;
  /* Spray and wet ground (NATRAY_SCENE_SPRAY): spray_cap records of 12 (position and radius,
   * velocity and weight, 0 0 0 and the time made), where each landed (4), and the wet map
   * (wet_res^2 f32 bits of the first time water touched each ground cell). */
  REAL *spray, *landed;
  uint32_t* touch;
  uint32_t spray_cap, wet_res;
  uint32_t *prev_ptr, *block_list_prev, *block_active;
  /* Bullets (set_bullets): n_bullets of them; their patches [n][576][8] and cavity sections
   * [n][32][12] (bullet.wgsl:16-50). */
  struct obullet {
    REAL x[4], v[4], q[4], l[4], shape[4], reach[4];
    uint32_t info[4];
  } bullets[NATRAY_MAX_BULLETS];
  REAL* patches;
  REAL* sections;
  REAL* pose_log; /* [n][256][12]: the renderer's log of each bullet's pose per step */
  uint32_t n_bullets;
};

/* ---- Params / State accessors (sim_defs.wgsl:159-231) -------------------------------- */

static inline uint32_t cap_cells(const osim* s) { return s->U.bdims[3] * 64u; }
static inline REAL hh_(const osim* s) { return (REAL)s->U.origin[3]; }
/* dt() and now(): f32(ns) * 1e-9 (sim_defs.wgsl:167-173) */
static inline REAL dt_(const osim* s) { return (REAL)s->st.dt_ns * (REAL)1e-9; }
static inline REAL now_(const osim* s) { return (REAL)s->st.t_ns * (REAL)1e-9; }

static inline uint32_t block_slot(const osim* s, i3 b) {
  return ((uint32_t)b.x * s->U.bdims[1] + (uint32_t)b.y) * s->U.bdims[2] + (uint32_t)b.z;
}
static inline i3 slot_coord(const osim* s, uint32_t sl) {
  i3 r;
  r.x = (int)(sl / (s->U.bdims[2] * s->U.bdims[1]));
  r.y = (int)((sl / s->U.bdims[2]) % s->U.bdims[1]);
  r.z = (int)(sl % s->U.bdims[2]);
  return r;
}
static inline i3 local_coord(uint32_t i) {
  uint32_t l = i % 64u;
  i3 r = {(int)(l / 16u), (int)((l / 4u) % 4u), (int)(l % 4u)};
  return r;
}
static inline i3 cell_coord(const osim* s, uint32_t i) {
  i3 b = slot_coord(s, s->block_list[i / 64u]), l = local_coord(i);
  i3 r = {b.x * 4 + l.x, b.y * 4 + l.y, b.z * 4 + l.z};
  return r;
}
/* neighbour (sim_defs.wgsl:195-202) */
static inline uint32_t neighbour(const osim* s, uint32_t cell, int ox, int oy, int oz) {
  i3 l = local_coord(cell);
  int qx = l.x + ox, qy = l.y + oy, qz = l.z + oz;
  int nx = (qx >= 4) - (qx < 0), ny = (qy >= 4) - (qy < 0), nz = (qz >= 4) - (qz < 0);
  uint32_t b = s->block_nb[(cell / 64u) * 27u + (uint32_t)((nx + 1) * 9 + (ny + 1) * 3 + nz + 1)];
  if (b == NONE) return NONE;
  return b * 64u + (uint32_t)(qx - nx * 4) * 16u + (uint32_t)(qy - ny * 4) * 4u + (uint32_t)(qz - nz * 4);
}
static inline int32_t* accp(osim* s, uint32_t f, uint32_t cell) { return &s->acc[(size_t)f * cap_cells(s) + cell]; }
static inline int32_t accv(const osim* s, uint32_t f, uint32_t cell) { return s->acc[(size_t)f * cap_cells(s) + cell]; }
static inline REAL* cfp(osim* s, uint32_t f, uint32_t cell) { return &s->cf[(size_t)f * cap_cells(s) + cell]; }
static inline REAL cfv(const osim* s, uint32_t f, uint32_t cell) { return s->cf[(size_t)f * cap_cells(s) + cell]; }
static inline REAL cell_type(const osim* s, uint32_t cell) { return cell == NONE ? GAS : cfv(s, NATRAY_F_TYPE, cell); }

/* to_fp / from_fp (sim_defs.wgsl:217-222); WGSL round() is ties-to-even. */
static inline int32_t to_fp(REAL x, REAL scale) {
  REAL v = x * scale;
  v = rmin(rmax(v, (REAL)-2e9), (REAL)2e9);
  return (int32_t)rrint(v);
}
static inline REAL from_fp(int32_t x, REAL scale) { return (REAL)x / scale; }

/* ---- scene geometry (reference solids fluid.wgsl:305-314; extensions defined here) ----- */

static inline r3 cell_centre(const osim* s, i3 c) {
  REAL h = hh_(s);
  r3 p = {(REAL)s->U.origin[0] + ((REAL)c.x + (REAL)0.5) * h, (REAL)s->U.origin[1] + ((REAL)c.y + (REAL)0.5) * h,
          (REAL)s->U.origin[2] + ((REAL)c.z + (REAL)0.5) * h};
  return p;
}
static inline int is_ground(const osim* s, i3 c) {
  return (REAL)s->U.origin[1] + ((REAL)c.y + (REAL)0.5) * hh_(s) < (REAL)0;
}
static inline int in_knot(const osim* s, i3 c) {
  r3 q = cell_centre(s, c);
  REAL px = q.x - (REAL)s->U.knot[0], py = q.y - (REAL)s->U.knot[1], pz = q.z - (REAL)s->U.knot[2];
  REAL r = (REAL)s->U.knot[3] + (REAL)0.5 * hh_(s);
  return py > (REAL)0 && py < (REAL)2 * r && sqrt_r(px * px + pz * pz) < r;
}
static inline int in_walls(const osim* s, i3 c) {
  return c.x == 0 || c.y == 0 || c.z == 0 || c.x == (int)s->U.dims[0] - 1 || c.y == (int)s->U.dims[1] - 1 ||
         c.z == (int)s->U.dims[2] - 1;
}
/* Closest point on the capsule axis at time t, and the squared distance to it. */
static inline REAL capsule_d2(const osim* s, r3 p, REAL t, r3* on_out, r3* ab_out) {
  const natray_scene* sc = &s->sc;
  r3 a = {(REAL)sc->capsule_a[0] + (REAL)sc->capsule_v[0] * t, (REAL)sc->capsule_a[1] + (REAL)sc->capsule_v[1] * t,
          (REAL)sc->capsule_a[2] + (REAL)sc->capsule_v[2] * t};
  r3 ab = {(REAL)sc->capsule_b[0] - (REAL)sc->capsule_a[0], (REAL)sc->capsule_b[1] - (REAL)sc->capsule_a[1],
           (REAL)sc->capsule_b[2] - (REAL)sc->capsule_a[2]};
  REAL ll = (ab.x * ab.x + ab.y * ab.y) + ab.z * ab.z;
  REAL f = (((p.x - a.x) * ab.x + (p.y - a.y) * ab.y) + (p.z - a.z) * ab.z) / rmax(ll, (REAL)1e-20);
  f = rmin(rmax(f, (REAL)0), (REAL)1);
  r3 on = {a.x + f * ab.x, a.y + f * ab.y, a.z + f * ab.z};
  r3 d = {p.x - on.x, p.y - on.y, p.z - on.z};
  if (on_out) *on_out = on;
  if (ab_out) *ab_out = ab;
  return (d.x * d.x + d.y * d.y) + d.z * d.z;
}
static inline int in_capsule(const osim* s, i3 c) {
  REAL r = (REAL)s->sc.capsule_a[3];
  return capsule_d2(s, cell_centre(s, c), now_(s), NULL, NULL) < r * r;
}
static inline int in_mask(const osim* s, i3 c) {
  r3 p = cell_centre(s, c);
  REAL dx = p.x - (REAL)s->sc.mask[0], dy = p.y - (REAL)s->sc.mask[1], dz = p.z - (REAL)s->sc.mask[2];
  REAL r = (REAL)s->sc.mask[3];
  return (dx * dx + dy * dy) + dz * dz < r * r;
}
static inline int is_solid_cell(const osim* s, i3 c) {
  uint32_t fl = s->sc.flags;
  if ((fl & NATRAY_SCENE_GROUND) && is_ground(s, c)) return 1;
  if ((fl & NATRAY_SCENE_KNOT) && in_knot(s, c)) return 1;
  if ((fl & NATRAY_SCENE_WALLS) && in_walls(s, c)) return 1;
  if ((fl & NATRAY_SCENE_CAPSULE) && in_capsule(s, c)) return 1;
  return 0;
}
/* Velocity component d of the solid in `cell` (0 for the ground, knot and walls). */
static inline REAL solid_u(const osim* s, uint32_t cell, uint32_t d) {
  if (!(s->sc.flags & NATRAY_SCENE_CAPSULE) || cell == NONE) return (REAL)0;
  return in_capsule(s, cell_coord(s, cell)) ? (REAL)s->sc.capsule_v[d] : (REAL)0;
}

/* bounded (fluid.wgsl:360-365), with a moving solid's velocity in place of 0. */
static inline REAL bounded(const osim* s, REAL u, REAL hi, REAL lo, uint32_t hi_cell, uint32_t lo_cell, uint32_t d) {
  if (lo == SOLID && hi == SOLID) return (REAL)0;
  if (lo == SOLID) return rmax(u, solid_u(s, lo_cell, d));
  if (hi == SOLID) return rmin(u, solid_u(s, hi_cell, d));
  return u;
}
/* face_u (fluid.wgsl:368-371) */
static inline REAL face_u(const osim* s, uint32_t cell, uint32_t lo, uint32_t d) {
  if (cell == NONE) return (REAL)0;
  return bounded(s, cfv(s, NATRAY_F_U + d, cell), cfv(s, NATRAY_F_TYPE, cell), cell_type(s, lo), cell, lo, d);
}
/* pressure_of (fluid.wgsl:375-379) */
static inline REAL pressure_of(const osim* s, uint32_t cell) {
  if (cell == NONE) return (REAL)0;
  REAL p = cfv(s, NATRAY_F_P, cell);
  return cfv(s, NATRAY_F_TYPE, cell) == WATER ? rmax(p, (REAL)NATRAY_P_FLOOR) : p;
}

static inline uint32_t n_blocks_(const osim* s) { return umin(s->st.n_blocks, s->U.bdims[3]); }

static void sync_state_floats(osim* s) {
  s->st.max_speed = (float)s->max_speed;
  s->st.rz = (float)s->rz;
  s->st.alpha = (float)s->alpha;
  s->st.beta = (float)s->beta;
  s->st.b_norm = (float)s->b_norm;
  s->st.r_norm = (float)s->r_norm;
}

/* ---- begin_step (fluid.wgsl:31-89) --------------------------------------------------- */

static void o_begin_step(osim* s) {
  natray_state* st = &s->st;
  REAL vmax = (REAL)s->max_speed;
  uint32_t arrive = s->U.limits[3];
  /* The moving capsule counts as a bullet in flight (fluid.wgsl:39-50). */
  if (s->sc.flags & NATRAY_SCENE_CAPSULE) {
    REAL vx = (REAL)s->sc.capsule_v[0], vy = (REAL)s->sc.capsule_v[1], vz = (REAL)s->sc.capsule_v[2];
    vmax = rmax(vmax, sqrt_r((vx * vx + vy * vy) + vz * vz));
  }
  /* A bullet's own pace counts from the step it joins in (fluid.wgsl:36-50). */
  for (uint32_t k = 0; k < s->n_bullets; ++k) {
    const struct obullet* b = &s->bullets[k];
    if (b->info[2] == 1u) {
      vmax = rmax(vmax, b->l[3]);
    } else if (b->info[2] == 0u) {
      if (b->info[0] > st->t_ns)
        arrive = umin(arrive, b->info[0]);
      else
        vmax = rmax(vmax, sqrt_r((b->v[0] * b->v[0] + b->v[1] * b->v[1]) + b->v[2] * b->v[2]));
    }
  }
  REAL mem_dt = (REAL)st->mem_dt, air_dt = (REAL)st->air_dt;
  REAL dt_s = rmin((REAL)s->U.fluid[2], (REAL)st->cap_dt);
  if (vmax > (REAL)0) dt_s = rmin(dt_s, (REAL)s->U.fluid[1] * hh_(s) / vmax);
  st->dt_flow = (float)dt_s;
  st->mem_dt_seen = (float)mem_dt;
  st->air_dt_seen = (float)air_dt;
  dt_s = rmin(dt_s, rmin((REAL)s->U.mem[1] * mem_dt, (REAL)s->U.mem[2] * air_dt));
  if (st->free_dt_ns > 0u) dt_s = rmin(dt_s, (REAL)1.25 * (REAL)st->free_dt_ns * (REAL)1e-9);
  uint32_t free_ns = umax((uint32_t)(dt_s * (REAL)1e9), 20u);
  /* A fixed step is rounded to whole nanoseconds once, the same way in every precision. */
  if (s->sc.fixed_dt > 0.0f) free_ns = umax((uint32_t)floor((double)s->sc.fixed_dt * 1e9 + 0.5), 20u);
  uint32_t stop = s->U.limits[3];
  int live = st->t_ns < stop;
  uint32_t dt_ns = 0;
  if (live) {
    dt_ns = umin(free_ns, umin(stop, arrive) - st->t_ns);
    st->free_dt_ns = free_ns;
    st->t_ns += dt_ns;
    st->step += 1u;
    s->max_speed = 0.0;
    st->mem_dt = FLT_MAX;
    st->cap_dt = FLT_MAX;
    st->air_dt = FLT_MAX;
    st->mem_ticket = 0;
    st->n_blocks_prev = n_blocks_(s);
    st->n_blocks = 0;
  }
  st->dt_ns = dt_ns;
  sync_state_floats(s);
}

/* ---- binning: clear_blocks, mark_blocks/claim, scan_blocks, sort_particles -------------- */

static inline i3 own_cell(const osim* s, const REAL* p) {
  /* own_cell (fluid.wgsl:127-129): floor((p - origin) / h) */
  REAL h = hh_(s);
  i3 c = {(int)rfloor((p[0] - (REAL)s->U.origin[0]) / h), (int)rfloor((p[1] - (REAL)s->U.origin[1]) / h),
          (int)rfloor((p[2] - (REAL)s->U.origin[2]) / h)};
  return c;
}
static inline int in_domain(const osim* s, i3 c) {
  return c.x >= 0 && c.y >= 0 && c.z >= 0 && c.x < (int)s->U.dims[0] && c.y < (int)s->U.dims[1] &&
         c.z < (int)s->U.dims[2];
}
static inline int floor4(int v) { return v >> 2; } /* arithmetic shift, as (c - 1) >> 2 in WGSL */
static inline int iclamp(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---- membrane coupling (membrane.wgsl:151-256) ----------------------------------------- */

#define CONTACT ((REAL)0.05) /* rubber this deep in water is on it (membrane.wgsl:38) */

static inline uint32_t mem_count(const osim* s) { return umin(s->U.mem[0], s->mem_cap); }
static inline int on_water(const osim* s, uint32_t i) {
  return s->mem_mass[i] > (REAL)0 && s->mem_water[(size_t)i * 4 + 3] >= CONTACT;
}
/* cell_at (membrane.wgsl:156-162): the cell at c, NONE outside the domain or where no block of
 * this step holds it. */
static inline uint32_t cell_at(const osim* s, i3 c) {
  if (!in_domain(s, c)) return NONE;
  i3 b = {c.x >> 2, c.y >> 2, c.z >> 2};
  uint32_t blk = s->block_ptr[block_slot(s, b)];
  if (blk >= n_blocks_(s)) return NONE;
  return blk * 64u + (uint32_t)(c.x & 3) * 16u + (uint32_t)(c.y & 3) * 4u + (uint32_t)(c.z & 3);
}

/* mark_rubber (membrane.wgsl:164-180): claim the blocks of cells c-2..c+3 around rubber off
 * the water. */
static void o_mark_rubber(osim* s) {
  int bmx = (int)s->U.bdims[0] - 1, bmy = (int)s->U.bdims[1] - 1, bmz = (int)s->U.bdims[2] - 1;
  for (uint32_t i = 0; i < mem_count(s); ++i) {
    if (s->mem_push[(size_t)i * 2 + 1] <= (REAL)0 || on_water(s, i)) continue;
    i3 c = own_cell(s, &s->mem_x[(size_t)i * 4]);
    int lx = iclamp(floor4(c.x - 2), 0, bmx), ly = iclamp(floor4(c.y - 2), 0, bmy), lz = iclamp(floor4(c.z - 2), 0, bmz);
    int hx = iclamp(floor4(c.x + 3), 0, bmx), hy = iclamp(floor4(c.y + 3), 0, bmy), hz = iclamp(floor4(c.z + 3), 0, bmz);
    for (int x = lx; x <= hx; ++x)
      for (int y = ly; y <= hy; ++y)
        for (int z = lz; z <= hz; ++z) {
          i3 b = {x, y, z};
          s->block_ptr[block_slot(s, b)] = PENDING;
        }
  }
}

static void o_bin(osim* s) {
  if (s->st.dt_ns == 0u) return;
  const uint32_t cap = cap_cells(s);
  /* clear_blocks (fluid.wgsl:91-107): free last step's blocks, zero their transfer sums. */
  uint32_t nprev = s->st.n_blocks_prev;
  for (uint32_t b = 0; b < nprev; ++b) {
    uint32_t sl = s->block_list[b];
    s->block_ptr[sl] = NONE;
    s->slot_count[sl] = 0;
    if (s->air_x) {
      s->block_list_prev[b] = sl;
      s->prev_ptr[sl] = b;
    }
  }
  /* keep_blocks (air.wgsl:271-277): the blocks whose air was active are claimed again. */
  if (s->air_x)
    for (uint32_t b = 0; b < nprev; ++b)
      if (s->block_active[b]) s->block_ptr[s->block_list_prev[b]] = PENDING;
  for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f) memset(&s->acc[(size_t)f * cap], 0, (size_t)nprev * 64u * 4u);
  /* mark_blocks (fluid.wgsl:131-150): claim every block the stencil (cells c-1..c+2) touches,
   * take a rank in the own block. Claims are recorded as PENDING and given indices in slot
   * order below. */
  int bmx = (int)s->U.bdims[0] - 1, bmy = (int)s->U.bdims[1] - 1, bmz = (int)s->U.bdims[2] - 1;
  for (uint32_t i = 0; i < s->N; ++i) {
    const REAL* p = &s->P[(size_t)i * 20];
    if (p[7] < (REAL)0.5) continue;
    i3 c = own_cell(s, p);
    if (!in_domain(s, c)) continue; /* the reference indexes out of bounds here; such particles are not binned */
    int lx = iclamp(floor4(c.x - 1), 0, bmx), ly = iclamp(floor4(c.y - 1), 0, bmy), lz = iclamp(floor4(c.z - 1), 0, bmz);
    int hx = iclamp(floor4(c.x + 2), 0, bmx), hy = iclamp(floor4(c.y + 2), 0, bmy), hz = iclamp(floor4(c.z + 2), 0, bmz);
    for (int x = lx; x <= hx; ++x)
      for (int y = ly; y <= hy; ++y)
        for (int z = lz; z <= hz; ++z) {
          i3 b = {x, y, z};
          s->block_ptr[block_slot(s, b)] = PENDING;
        }
    i3 ob = {c.x >> 2, c.y >> 2, c.z >> 2};
    s->rank[i] = s->slot_count[block_slot(s, ob)]++;
  }
  o_mark_rubber(s);
  /* Block indices in slot order; overflowing claims are dropped (claim, fluid.wgsl:117-122). */
  uint32_t n = 0, bcap = s->U.bdims[3];
  for (uint32_t sl = 0; sl < s->S; ++sl) {
    if (s->block_ptr[sl] != PENDING) continue;
    if (n >= bcap) {
      s->st.overflow = 1u;
      s->block_ptr[sl] = NONE;
    } else {
      s->block_list[n] = sl;
      s->block_ptr[sl] = n;
    }
    ++n;
  }
  s->st.n_blocks = n;
  /* scan_blocks (fluid.wgsl:157-187): exclusive prefix sum of the blocks' own counts. */
  uint32_t nb = n_blocks_(s), run = 0;
  for (uint32_t b = 0; b < nb; ++b) {
    s->block_start[b] = run;
    run += s->slot_count[s->block_list[b]];
  }
  s->block_start[nb] = run;
  /* sort_particles (fluid.wgsl:189-196) */
  for (uint32_t i = 0; i < s->N; ++i) {
    const REAL* p = &s->P[(size_t)i * 20];
    if (p[7] < (REAL)0.5) continue;
    i3 c = own_cell(s, p);
    if (!in_domain(s, c)) continue;
    i3 ob = {c.x >> 2, c.y >> 2, c.z >> 2};
    uint32_t b = s->block_ptr[block_slot(s, ob)];
    if (b >= PENDING) continue;
    s->sorted[s->block_start[b] + s->rank[i]] = i;
  }
}

/* ---- p2g (fluid.wgsl:198-303) ---------------------------------------------------------- */

/* axis_weights (fluid.wgsl:221-223) */
static inline void axis_weights(REAL fx, REAL w[3]) {
  w[0] = (REAL)0.5 * ((REAL)1.5 - fx) * ((REAL)1.5 - fx);
  w[1] = (REAL)0.75 - (fx - (REAL)1) * (fx - (REAL)1);
  w[2] = (REAL)0.5 * (fx - (REAL)0.5) * (fx - (REAL)0.5);
}

/* splat (fluid.wgsl:233-252): one lattice of one particle into the block's tile. */
static void splat(int32_t* tile, const osim* s, const REAL x[3], const REAL o[3], i3 corner, uint32_t fa, REAL val,
                  const REAL grad[3], uint32_t fw, REAL scale) {
  int base[3];
  REAL fx[3], wx[3], wy[3], wz[3], hg[3];
  for (int k = 0; k < 3; ++k) {
    base[k] = (int)rfloor(x[k] - o[k] - (REAL)0.5);
    fx[k] = x[k] - o[k] - (REAL)base[k];
    hg[k] = grad[k] * hh_(s);
  }
  axis_weights(fx[0], wx);
  axis_weights(fx[1], wy);
  axis_weights(fx[2], wz);
  uint32_t t0 = (uint32_t)(base[0] - corner.x) * 49u + (uint32_t)(base[1] - corner.y) * 7u + (uint32_t)(base[2] - corner.z);
  for (uint32_t a = 0; a < 3; ++a)
    for (uint32_t b = 0; b < 3; ++b)
      for (uint32_t c = 0; c < 3; ++c) {
        REAL weight = wx[a] * wy[b] * wz[c];
        uint32_t t = t0 + a * 49u + b * 7u + c;
        REAL node = val + ((hg[0] * ((REAL)a - fx[0]) + hg[1] * ((REAL)b - fx[1])) + hg[2] * ((REAL)c - fx[2]));
        tile[fa * TILE + t] += to_fp(weight * node, scale);
        tile[fw * TILE + t] += to_fp(weight, (REAL)NATRAY_FP_W);
      }
}

/* tile_cell (fluid.wgsl:210-217) */
static inline uint32_t tile_cell(const uint32_t links[27], uint32_t t) {
  int rx = (int)(t / 49u) - 1, ry = (int)((t / 7u) % 7u) - 1, rz = (int)(t % 7u) - 1;
  int nx = (rx + 4) / 4 - 1, ny = (ry + 4) / 4 - 1, nz = (rz + 4) / 4 - 1;
  uint32_t b = links[(nx + 1) * 9 + (ny + 1) * 3 + nz + 1];
  if (b == NONE) return NONE;
  return b * 64u + (uint32_t)(rx - nx * 4) * 16u + (uint32_t)(ry - ny * 4) * 4u + (uint32_t)(rz - nz * 4);
}

static void o_p2g(osim* s) {
  if (s->st.dt_ns == 0u) return;
  uint32_t nb = n_blocks_(s);
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t b = 0; b < nb; ++b) {
    int32_t tile[8 * TILE];
    uint32_t own[64];
    uint32_t links[27];
    i3 bc = slot_coord(s, s->block_list[b]);
    for (uint32_t li = 0; li < 27; ++li) {
      i3 q = {bc.x + (int)(li / 9u) - 1, bc.y + (int)((li / 3u) % 3u) - 1, bc.z + (int)(li % 3u) - 1};
      uint32_t idx = NONE;
      if (q.x >= 0 && q.y >= 0 && q.z >= 0 && q.x < (int)s->U.bdims[0] && q.y < (int)s->U.bdims[1] &&
          q.z < (int)s->U.bdims[2]) {
        idx = s->block_ptr[block_slot(s, q)];
        if (idx >= PENDING) idx = NONE;
      }
      links[li] = idx;
      s->block_nb[b * 27u + li] = idx;
    }
    uint32_t count = s->block_start[b + 1] - s->block_start[b];
    if (count == 0u) {
      for (uint32_t li = 0; li < 64; ++li) *accp(s, NATRAY_A_N, b * 64u + li) = 0;
      continue;
    }
    memset(tile, 0, sizeof tile);
    memset(own, 0, sizeof own);
    i3 corner = {bc.x * 4 - 1, bc.y * 4 - 1, bc.z * 4 - 1};
    uint32_t first = s->block_start[b];
    REAL h = hh_(s);
    for (uint32_t k = 0; k < count; ++k) {
      const REAL* p = &s->P[(size_t)s->sorted[first + k] * 20];
      REAL x[3] = {(p[0] - (REAL)s->U.origin[0]) / h, (p[1] - (REAL)s->U.origin[1]) / h, (p[2] - (REAL)s->U.origin[2]) / h};
      static const REAL ou[3] = {0, 0.5, 0.5}, ov[3] = {0.5, 0, 0.5}, ow[3] = {0.5, 0.5, 0}, op[3] = {0.5, 0.5, 0.5};
      static const REAL zero[3] = {0, 0, 0};
      splat(tile, s, x, ou, corner, NATRAY_A_MOM(0), p[4], &p[8], NATRAY_A_WT(0), (REAL)NATRAY_FP_M);
      splat(tile, s, x, ov, corner, NATRAY_A_MOM(1), p[5], &p[12], NATRAY_A_WT(1), (REAL)NATRAY_FP_M);
      splat(tile, s, x, ow, corner, NATRAY_A_MOM(2), p[6], &p[16], NATRAY_A_WT(2), (REAL)NATRAY_FP_M);
      splat(tile, s, x, op, corner, NATRAY_A_PC, p[11] * (REAL)1e-6, zero, NATRAY_A_WC, (REAL)NATRAY_FP_P);
      i3 c = own_cell(s, p);
      own[(uint32_t)(c.x - bc.x * 4) * 16u + (uint32_t)(c.y - bc.y * 4) * 4u + (uint32_t)(c.z - bc.z * 4)] += 1u;
    }
    for (uint32_t t = 0; t < TILE; ++t) {
      uint32_t cell = tile_cell(links, t);
      if (cell == NONE) continue;
      for (uint32_t f = 0; f < 8u; ++f) {
        int32_t v = tile[f * TILE + t];
        if (v != 0) {
          int32_t* dst = accp(s, f, cell);
#pragma omp atomic
          *dst += v;
        }
      }
    }
    for (uint32_t li = 0; li < 64; ++li) *accp(s, NATRAY_A_N, b * 64u + li) = (int32_t)own[li];
  }
}

/* ---- grid_prep (fluid.wgsl:319-356) ----------------------------------------------------- */

/* mem_splat (membrane.wgsl:182-216): a vertex with area adds its area (mm^2) and area times
 * velocity to its own cell; a vertex on the water also spreads its push and area over the 27
 * cells around it with the particles' weights. */
static void o_mem_splat(osim* s) {
  if (s->st.dt_ns == 0u) return;
  REAL h = hh_(s);
  for (uint32_t i = 0; i < mem_count(s); ++i) {
    if (s->mem_mass[i] == (REAL)0) continue;
    const REAL* pa = &s->mem_push[(size_t)i * 2];
    if (pa[1] <= (REAL)0) continue;
    const REAL* p = &s->mem_x[(size_t)i * 4];
    uint32_t own = cell_at(s, own_cell(s, p));
    if (own != NONE) {
      REAL a = pa[1] * (REAL)1e6;
      *accp(s, NATRAY_A_MW, own) += to_fp(a, (REAL)NATRAY_FP_A);
      for (uint32_t d = 0; d < 3; ++d) *accp(s, NATRAY_A_MV + d, own) += to_fp(a * s->mem_v[(size_t)i * 4 + d], (REAL)NATRAY_FP_A);
    }
    if (!on_water(s, i)) continue;
    REAL x[3], w[3][3];
    int base[3];
    for (int k = 0; k < 3; ++k) {
      x[k] = (p[k] - (REAL)s->U.origin[k]) / h - (REAL)0.5;
      REAL bf = rfloor(x[k] - (REAL)0.5);
      base[k] = (int)bf;
      axis_weights(x[k] - bf, w[k]);
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          i3 cc = {base[0] + a, base[1] + b, base[2] + c};
          uint32_t cell = cell_at(s, cc);
          if (cell == NONE) continue;
          REAL wt = w[0][a] * w[1][b] * w[2][c];
          *accp(s, NATRAY_A_MF, cell) += to_fp(wt * pa[0], (REAL)NATRAY_FP_F);
          *accp(s, NATRAY_A_MA, cell) += to_fp(wt * pa[1] * (REAL)1e6, (REAL)NATRAY_FP_A);
        }
  }
}

/* sample_grid (membrane.wgsl:218-237) */
static void sample_grid(const osim* s, const REAL x[3], uint32_t fu, uint32_t fw, REAL out[2]) {
  REAL w[3][3];
  int base[3];
  for (int k = 0; k < 3; ++k) {
    REAL bf = rfloor(x[k] - (REAL)0.5);
    base[k] = (int)bf;
    axis_weights(x[k] - bf, w[k]);
  }
  REAL sx = 0, sy = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        i3 cc = {base[0] + a, base[1] + b, base[2] + c};
        uint32_t cell = cell_at(s, cc);
        if (cell == NONE) continue;
        REAL wt = w[0][a] * w[1][b] * w[2][c] * from_fp(accv(s, fw, cell), (REAL)NATRAY_FP_W);
        sx = sx + wt * cfv(s, fu, cell);
        sy = sy + wt;
      }
  out[0] = sx;
  out[1] = sy;
}

/* mem_sample (membrane.wgsl:239-254): the face velocity around each vertex, weighted by the
 * water each face carries, and how full of water the cells around it are. */
static void o_mem_sample(osim* s) {
  if (s->st.dt_ns == 0u) return;
  REAL h = hh_(s);
  for (uint32_t i = 0; i < mem_count(s); ++i) {
    if (s->mem_mass[i] == (REAL)0) continue;
    const REAL* p = &s->mem_x[(size_t)i * 4];
    REAL x[3], v[3], r[2];
    for (int k = 0; k < 3; ++k) x[k] = (p[k] - (REAL)s->U.origin[k]) / h;
    for (uint32_t d = 0; d < 3; ++d) {
      REAL xs[3];
      for (uint32_t k = 0; k < 3; ++k) xs[k] = k == d ? x[k] : x[k] - (REAL)0.5;
      sample_grid(s, xs, NATRAY_F_U + d, NATRAY_A_WT(d), r);
      v[d] = r[1] > (REAL)1e-6 ? r[0] / r[1] : (REAL)0;
    }
    REAL xc[3] = {x[0] - (REAL)0.5, x[1] - (REAL)0.5, x[2] - (REAL)0.5};
    sample_grid(s, xc, NATRAY_F_TYPE, NATRAY_A_WC, r);
    REAL* mw = &s->mem_water[(size_t)i * 4];
    mw[0] = v[0];
    mw[1] = v[1];
    mw[2] = v[2];
    mw[3] = r[1] / (REAL)8;
  }
}

/* ---- the air boundary (air.wgsl:85-148) ------------------------------------------------ */

#define GAMMA ((REAL)1.4f)
#define P_ATM ((REAL)101325.0f)
#define P_VAPOUR ((REAL)2339.0f)
#define RHO_ATM ((REAL)1.204f)
#define RHO_VAPOUR ((REAL)0.0278f)

typedef struct { REAL rho, mx, my, mz, e; } gas;
static inline gas gas_at_rest(REAL rho, REAL p) { gas g = {rho, 0, 0, 0, p / (GAMMA - (REAL)1)}; return g; }
static inline REAL gas_p(gas g) { return (GAMMA - (REAL)1) * (g.e - (REAL)0.5 * ((g.mx * g.mx + g.my * g.my) + g.mz * g.mz) / g.rho); }
static inline gas read_gas(const REAL* air, uint32_t capc, uint32_t cell) {
  gas g = {air[cell], air[(size_t)capc + cell], air[2 * (size_t)capc + cell], air[3 * (size_t)capc + cell], air[4 * (size_t)capc + cell]};
  return g;
}
static inline void write_gas(REAL* air, uint32_t capc, uint32_t cell, gas g) {
  air[cell] = g.rho;
  air[(size_t)capc + cell] = g.mx;
  air[2 * (size_t)capc + cell] = g.my;
  air[3 * (size_t)capc + cell] = g.mz;
  air[4 * (size_t)capc + cell] = g.e;
}

/* air_carry (air.wgsl:85-98): each cell gets its gas of last step, or the ambient air where its
 * block was not allocated then. */
static void o_air_carry(osim* s) {
  if (s->st.dt_ns == 0u || !s->air_x) return;
  uint32_t nb = n_blocks_(s), capc = cap_cells(s), nprev = s->st.n_blocks_prev;
  for (uint32_t b = 0; b < nb; ++b) {
    uint32_t sl = s->block_list[b], o = s->prev_ptr[sl];
    int carried = o < nprev && s->block_list_prev[o] == sl;
    for (uint32_t li = 0; li < 64; ++li) {
      gas g = carried ? read_gas(s->air_x, capc, o * 64u + li) : gas_at_rest(RHO_ATM, P_ATM);
      write_gas(s->air_tmp, capc, b * 64u + li, g);
    }
    s->block_active[b] = 0;
  }
}

/* air_fill (air.wgsl:100-148): water and solid cells hold no gas; a cell without gas takes the
 * mean of its neighbours' (in the order -x, +x, -y, +y, -z, +z) or vapour; GAS and RUBBER cells
 * get the gas's gauge pressure, a RUBBER cell the lowest of its GAS neighbours'. */
static void o_air_fill(osim* s) {
  if (s->st.dt_ns == 0u || !s->air_x) return;
  uint32_t nc = n_blocks_(s) * 64u, capc = cap_cells(s);
  for (uint32_t cell = 0; cell < nc; ++cell) {
    REAL ty = cfv(s, NATRAY_F_TYPE, cell);
    if (ty == WATER || ty == SOLID) {
      gas z = {0, 0, 0, 0, 0};
      write_gas(s->air_x, capc, cell, z);
      continue;
    }
    uint32_t nbs[6];
    for (int k = 0; k < 6; ++k) {
      int e[3] = {0, 0, 0};
      e[k / 2] = (k % 2 == 1) ? 1 : -1;
      nbs[k] = neighbour(s, cell, e[0], e[1], e[2]);
    }
    gas g = read_gas(s->air_tmp, capc, cell);
    if (g.rho <= (REAL)0) {
      gas sum = {0, 0, 0, 0, 0};
      REAL n = 0;
      for (int k = 0; k < 6; ++k) {
        if (nbs[k] == NONE) continue;
        gas q = read_gas(s->air_tmp, capc, nbs[k]);
        if (q.rho > (REAL)0) {
          sum.rho = sum.rho + q.rho;
          sum.mx = sum.mx + q.mx;
          sum.my = sum.my + q.my;
          sum.mz = sum.mz + q.mz;
          sum.e = sum.e + q.e;
          n += (REAL)1;
        }
      }
      if (n > (REAL)0) {
        gas m = {sum.rho / n, sum.mx / n, sum.my / n, sum.mz / n, sum.e / n};
        g = m;
      } else {
        g = gas_at_rest(RHO_VAPOUR, P_VAPOUR);
      }
    }
    write_gas(s->air_x, capc, cell, g);
    if (ty == GAS) *cfp(s, NATRAY_F_P, cell) = gas_p(g) - P_ATM;
    if (ty == RUBBER) {
      REAL p = (REAL)3.0e38f;
      for (int k = 0; k < 6; ++k) {
        if (nbs[k] == NONE || cfv(s, NATRAY_F_TYPE, nbs[k]) != GAS) continue;
        gas q = read_gas(s->air_tmp, capc, nbs[k]);
        if (q.rho > (REAL)0) p = rmin(p, gas_p(q));
      }
      *cfp(s, NATRAY_F_P, cell) = (p < (REAL)3.0e38f ? p : gas_p(g)) - P_ATM;
    }
  }
}

static void o_grid_prep(osim* s) {
  if (s->st.dt_ns == 0u) return;
  uint32_t nc = n_blocks_(s) * 64u;
  REAL dt = dt_(s);
#pragma omp parallel for schedule(static)
  for (uint32_t cell = 0; cell < nc; ++cell) {
    REAL ty = GAS;
    i3 c = cell_coord(s, cell);
    if (is_solid_cell(s, c)) {
      ty = SOLID;
    } else if (accv(s, NATRAY_A_N, cell) > 0) {
      ty = WATER;
    }
    *cfp(s, NATRAY_F_TYPE, cell) = ty;
    for (uint32_t d = 0; d < 3; ++d) {
      REAL w = from_fp(accv(s, NATRAY_A_WT(d), cell), (REAL)NATRAY_FP_W);
      REAL u = 0;
      if (w > (REAL)1e-6) u = from_fp(accv(s, NATRAY_A_MOM(d), cell), (REAL)NATRAY_FP_M) / w + dt * (REAL)s->U.gravity[d];
      *cfp(s, NATRAY_F_U + d, cell) = u;
    }
    REAL wc = from_fp(accv(s, NATRAY_A_WC, cell), (REAL)NATRAY_FP_W);
    REAL ma = from_fp(accv(s, NATRAY_A_MA, cell), (REAL)NATRAY_FP_A);
    if (ty == GAS && accv(s, NATRAY_A_MW, cell) > 0) {
      ty = RUBBER;
      *cfp(s, NATRAY_F_TYPE, cell) = ty;
    }
    REAL p = 0;
    if (ty == WATER && wc > (REAL)1e-6) {
      p = from_fp(accv(s, NATRAY_A_PC, cell), (REAL)NATRAY_FP_P) / wc * (REAL)1e6;
    } else if (ty != WATER && ty != SOLID && ma > (REAL)0) {
      ty = COVERED;
      *cfp(s, NATRAY_F_TYPE, cell) = ty;
      p = from_fp(accv(s, NATRAY_A_MF, cell), (REAL)NATRAY_FP_F) / ma * (REAL)1e6;
    } else if (ty != WATER && ty != SOLID && (s->sc.flags & NATRAY_SCENE_MASK) && in_mask(s, c)) {
      /* The fixed spherical mask stands in for the membrane: the water there is held at
       * the mask's pressure, as behind the rubber. */
      ty = COVERED;
      *cfp(s, NATRAY_F_TYPE, cell) = ty;
      p = (REAL)s->sc.mask_p;
    }
    *cfp(s, NATRAY_F_P, cell) = p;
  }
}

/* ---- reductions (fluid.wgsl:412-439 and the chunk fold above) ---------------------------- */

/* block_sum's tree over one block's 64 values (fluid.wgsl:414-422). */
static double tree64(double v[64], int dbl) {
  if (dbl) {
    for (uint32_t off = 32; off > 0; off >>= 1)
      for (uint32_t li = 0; li < off; ++li) v[li] += v[li + off];
    return v[0];
  }
  REAL r[64];
  for (int i = 0; i < 64; ++i) r[i] = (REAL)v[i];
  for (uint32_t off = 32; off > 0; off >>= 1)
    for (uint32_t li = 0; li < off; ++li) r[li] += r[li + off];
  return (double)r[0];
}

/* Sum of component k of the n block partials: chunks of K = ceil(n/CHUNKS_MAX) consecutive
 * blocks, each summed as a binary tree over W = the next power of two >= K slots (slot i = block
 * lo + i, zero past the chunk's last block; v[i] += v[i + off] for off = W/2 .. 1), then the
 * reference's total() (fluid.wgsl:426-439) over the chunk sums. For n <= CHUNKS_MAX (K = 1)
 * this is the reference's total() over the blocks in slot order. */
#define CHUNKS_MAX 256u
static uint32_t chunk_width(uint32_t K) {
  uint32_t w = 1u;
  while (w < K) w <<= 1;
  return w;
}
static double total(const osim* s, uint32_t n, int k, int dbl) {
  if (n == 0) return 0.0;
  uint32_t K = (n + CHUNKS_MAX - 1u) / CHUNKS_MAX, m = (n + K - 1u) / K, W = chunk_width(K);
  if (dbl) {
    double chunk[CHUNKS_MAX], red[256], v[CHUNKS_MAX];
    for (uint32_t j = 0; j < m; ++j) {
      for (uint32_t i = 0; i < W; ++i) v[i] = j * K + i < umin(n, j * K + K) ? s->part[(j * K + i) * 4 + k] : 0.0;
      for (uint32_t off = W / 2u; off > 0; off >>= 1)
        for (uint32_t i = 0; i < off; ++i) v[i] += v[i + off];
      chunk[j] = v[0];
    }
    for (uint32_t li = 0; li < 256; ++li) {
      double acc = 0;
      for (uint32_t i = li; i < m; i += 256) acc += chunk[i];
      red[li] = acc;
    }
    for (uint32_t off = 128; off > 0; off >>= 1)
      for (uint32_t li = 0; li < off; ++li) red[li] += red[li + off];
    return red[0];
  }
  REAL rc[CHUNKS_MAX], rr[256], v[CHUNKS_MAX];
  for (uint32_t j = 0; j < m; ++j) {
    for (uint32_t i = 0; i < W; ++i) v[i] = j * K + i < umin(n, j * K + K) ? (REAL)s->part[(j * K + i) * 4 + k] : (REAL)0;
    for (uint32_t off = W / 2u; off > 0; off >>= 1)
      for (uint32_t i = 0; i < off; ++i) v[i] += v[i + off];
    rc[j] = v[0];
  }
  for (uint32_t li = 0; li < 256; ++li) {
    REAL acc = 0;
    for (uint32_t i = li; i < m; i += 256) acc += rc[i];
    rr[li] = acc;
  }
  for (uint32_t off = 128; off > 0; off >>= 1)
    for (uint32_t li = 0; li < off; ++li) rr[li] += rr[li + off];
  return (double)rr[0];
}

/* converged (fluid.wgsl:441-443) */
static int converged(const osim* s, double rr) {
  if (s->mixed) return sqrt(rr) <= (double)s->U.solve[0] * (s->b_norm < 1.0 ? 1.0 : s->b_norm);
  return sqrt_r((REAL)rr) <= (REAL)s->U.solve[0] * rmax((REAL)s->b_norm, (REAL)1);
}

/* ---- precondition (fluid.wgsl:384-410) and the Jacobi alternative ------------------------- */

static void precondition(const osim* s, const REAL r[64], const REAL diag[64], const int water[64], REAL z[64]) {
  if (s->sc.precond == NATRAY_PRECOND_JACOBI) {
    for (int li = 0; li < 64; ++li) z[li] = water[li] ? r[li] / diag[li] : (REAL)0;
    return;
  }
  REAL zl[64];
  for (int li = 0; li < 64; ++li) zl[li] = 0;
  for (uint32_t k = 0; k < 5u; ++k) {
    /* One colour at a time; cells of one colour read only the other colour. */
    for (uint32_t li = 0; li < 64; ++li) {
      uint32_t lx = li / 16u, ly = (li / 4u) % 4u, lz = li % 4u;
      uint32_t colour = (lx + ly + lz) % 2u;
      if (!(water[li] && colour == k % 2u)) continue;
      REAL sum = r[li];
      if (lx > 0 && water[li - 16]) sum += zl[li - 16];
      if (lx < 3 && water[li + 16]) sum += zl[li + 16];
      if (ly > 0 && water[li - 4]) sum += zl[li - 4];
      if (ly < 3 && water[li + 4]) sum += zl[li + 4];
      if (lz > 0 && water[li - 1]) sum += zl[li - 1];
      if (lz < 3 && water[li + 1]) sum += zl[li + 1];
      zl[li] = sum / diag[li];
    }
  }
  for (int li = 0; li < 64; ++li) z[li] = zl[li];
}

/* product of two REALs in the reduction's precision */
static inline double prod(const osim* s, REAL a, REAL b) { return s->mixed ? (double)a * (double)b : (double)(a * b); }

/* ---- solve_init + solve_start (fluid.wgsl:445-506) ---------------------------------------- */

/* The solve's row of water cell `cell` (solve_init, fluid.wgsl:450-489): the diagonal and the
 * right-hand side b, both from the particle-carried pressure, and the return value r = b - A p0
 * for the starting pressure p0 (the carried pressure, or a warm-start guess; the reference
 * always starts from the carried pressure). */
static REAL solve_row(const osim* s, uint32_t cell, const REAL* p0, REAL* dg_out, REAL* b_out) {
  REAL h = hh_(s), dt = dt_(s);
  REAL cs = (REAL)s->U.gravity[3] * dt;
  REAL kc = h * h / (cs * cs);
  REAL div = 0, off = 0, dg = kc, b = kc * cfv(s, NATRAY_F_P, cell);
  for (uint32_t d = 0; d < 3; ++d) {
    int e[3] = {0, 0, 0};
    e[d] = 1;
    uint32_t up = neighbour(s, cell, e[0], e[1], e[2]);
    uint32_t dn = neighbour(s, cell, -e[0], -e[1], -e[2]);
    div += face_u(s, up, cell, d) - face_u(s, cell, dn, d);
    for (int sg = -1; sg <= 1; sg += 2) {
      uint32_t nbc = neighbour(s, cell, sg * e[0], sg * e[1], sg * e[2]);
      REAL nt = cell_type(s, nbc);
      if (nt == SOLID) continue;
      dg += (REAL)1;
      if (nt == WATER) {
        off -= p0[nbc];
      } else {
        b += pressure_of(s, nbc);
      }
    }
  }
  div -= from_fp(accv(s, NATRAY_A_SRC, cell), (REAL)NATRAY_FP_Q) / (h * h);
  b -= (REAL)s->U.fluid[0] * h / dt * div;
  *dg_out = dg;
  *b_out = b;
  return b - dg * p0[cell] - off;
}

static void o_solve_init(osim* s) {
  uint32_t nb = n_blocks_(s);
  REAL* P = s->cf + (size_t)NATRAY_F_P * cap_cells(s);
  const REAL* p0 = s->use_guess ? s->guess : P;
#pragma omp parallel for schedule(static)
  for (uint32_t blk = 0; blk < nb; ++blk) {
    REAL r[64], diag[64], bb[64], z[64];
    int water[64];
    for (uint32_t li = 0; li < 64; ++li) {
      uint32_t cell = blk * 64u + li;
      water[li] = cfv(s, NATRAY_F_TYPE, cell) == WATER;
      diag[li] = 1;
      r[li] = 0;
      bb[li] = 0;
      if (!water[li]) continue;
      REAL dg, b;
      REAL rr = solve_row(s, cell, p0, &dg, &b);
      *cfp(s, NATRAY_F_DIAG, cell) = dg;
      *cfp(s, NATRAY_F_R, cell) = rr;
      s->bvec[cell] = b;
      diag[li] = dg;
      r[li] = rr;
      bb[li] = b;
    }
    precondition(s, r, diag, water, z);
    double v0[64], v1[64], v2[64];
    for (uint32_t li = 0; li < 64; ++li) {
      if (water[li]) *cfp(s, NATRAY_F_Z, blk * 64u + li) = z[li];
      v0[li] = prod(s, r[li], z[li]);
      v1[li] = prod(s, r[li], r[li]);
      v2[li] = prod(s, bb[li], bb[li]);
    }
    s->part[blk * 4 + 0] = tree64(v0, s->mixed);
    s->part[blk * 4 + 1] = tree64(v1, s->mixed);
    s->part[blk * 4 + 2] = tree64(v2, s->mixed);
    s->part[blk * 4 + 3] = 0;
  }
  if (p0 != P) {
    /* the iterate starts at the guess */
    for (uint32_t cell = 0; cell < nb * 64u; ++cell)
      if (cfv(s, NATRAY_F_TYPE, cell) == WATER) P[cell] = p0[cell];
  }
}

/* Returns whether the iterations run. */
static int o_solve_start(osim* s) {
  uint32_t nb = n_blocks_(s);
  int dbl = s->mixed;
  double rz = total(s, nb, 0, dbl), rr = total(s, nb, 1, dbl), bb = total(s, nb, 2, dbl);
  s->rz = rz;
  s->beta = 0;
  if (dbl) {
    s->b_norm = sqrt(bb);
    s->r_norm = sqrt(rr);
  } else {
    s->b_norm = (double)sqrt_r((REAL)bb);
    s->r_norm = (double)sqrt_r((REAL)rr);
  }
  s->st.solve_iters = 0;
  sync_state_floats(s);
  return rz > 0.0 && !converged(s, rr);
}

/* ---- one CG iteration: solve_q, solve_alpha, solve_update, solve_beta (fluid.wgsl:508-578) -- */

static inline REAL direction(const osim* s, uint32_t cell, uint32_t k, REAL beta) {
  return cfv(s, NATRAY_F_Z, cell) + beta * cfv(s, NATRAY_F_D + k, cell);
}

static void o_solve_q(osim* s) {
  uint32_t nb = n_blocks_(s);
  uint32_t old = s->st.solve_iters % 2u;
  REAL beta = (REAL)s->beta;
  /* Every block reads its neighbours' old direction field and writes only the new one. */
#pragma omp parallel for schedule(static)
  for (uint32_t blk = 0; blk < nb; ++blk) {
    double v[64];
    for (uint32_t li = 0; li < 64; ++li) {
      uint32_t cell = blk * 64u + li;
      v[li] = 0;
      if (cfv(s, NATRAY_F_TYPE, cell) != WATER) continue;
      REAL dn = direction(s, cell, old, beta);
      REAL q = cfv(s, NATRAY_F_DIAG, cell) * dn;
      for (uint32_t d = 0; d < 3; ++d) {
        int e[3] = {0, 0, 0};
        e[d] = 1;
        for (int k = -1; k <= 1; k += 2) {
          uint32_t nbc = neighbour(s, cell, k * e[0], k * e[1], k * e[2]);
          if (cell_type(s, nbc) == WATER) q -= direction(s, nbc, old, beta);
        }
      }
      *cfp(s, NATRAY_F_D + 1u - old, cell) = dn;
      *cfp(s, NATRAY_F_Q, cell) = q;
      v[li] = prod(s, dn, q);
    }
    s->part[blk * 4 + 0] = tree64(v, s->mixed);
  }
}

static void o_solve_alpha(osim* s) {
  double dq = total(s, n_blocks_(s), 0, s->mixed);
  if (s->mixed)
    s->alpha = dq > 0.0 ? s->rz / dq : 0.0;
  else
    s->alpha = dq > 0.0 ? (double)((REAL)s->rz / (REAL)dq) : 0.0;
  sync_state_floats(s);
}

/* r = b - A p at the water cells of block blk, accumulated in double (mixed precision's
 * true-residual refresh). */
static REAL true_residual(const osim* s, uint32_t cell) {
  double ap = (double)cfv(s, NATRAY_F_DIAG, cell) * (double)cfv(s, NATRAY_F_P, cell);
  for (uint32_t d = 0; d < 3; ++d) {
    int e[3] = {0, 0, 0};
    e[d] = 1;
    for (int k = -1; k <= 1; k += 2) {
      uint32_t nbc = neighbour(s, cell, k * e[0], k * e[1], k * e[2]);
      if (cell_type(s, nbc) == WATER) ap -= (double)cfv(s, NATRAY_F_P, nbc);
    }
  }
  return (REAL)((double)s->bvec[cell] - ap);
}

static void o_solve_update(osim* s) {
  uint32_t nb = n_blocks_(s);
  REAL a = (REAL)s->alpha;
  uint32_t newd = NATRAY_F_D + 1u - s->st.solve_iters % 2u;
  int refresh = s->refresh && (s->st.solve_iters + 1u) % s->sc.refresh_every == 0u;
#pragma omp parallel for schedule(static)
  for (uint32_t blk = 0; blk < nb; ++blk)
    for (uint32_t li = 0; li < 64; ++li) {
      uint32_t cell = blk * 64u + li;
      if (cfv(s, NATRAY_F_TYPE, cell) != WATER) continue;
      *cfp(s, NATRAY_F_P, cell) += a * cfv(s, newd, cell);
      *cfp(s, NATRAY_F_R, cell) = cfv(s, NATRAY_F_R, cell) - a * cfv(s, NATRAY_F_Q, cell);
    }
  if (refresh) {
#pragma omp parallel for schedule(static)
    for (uint32_t blk = 0; blk < nb; ++blk)
      for (uint32_t li = 0; li < 64; ++li) {
        uint32_t cell = blk * 64u + li;
        if (cfv(s, NATRAY_F_TYPE, cell) == WATER) *cfp(s, NATRAY_F_R, cell) = true_residual(s, cell);
      }
  }
#pragma omp parallel for schedule(static)
  for (uint32_t blk = 0; blk < nb; ++blk) {
    REAL r[64], diag[64], z[64];
    int water[64];
    for (uint32_t li = 0; li < 64; ++li) {
      uint32_t cell = blk * 64u + li;
      water[li] = cfv(s, NATRAY_F_TYPE, cell) == WATER;
      r[li] = water[li] ? cfv(s, NATRAY_F_R, cell) : (REAL)0;
      diag[li] = water[li] ? cfv(s, NATRAY_F_DIAG, cell) : (REAL)1;
    }
    precondition(s, r, diag, water, z);
    double v0[64], v1[64];
    for (uint32_t li = 0; li < 64; ++li) {
      if (water[li]) *cfp(s, NATRAY_F_Z, blk * 64u + li) = z[li];
      v0[li] = prod(s, r[li], z[li]);
      v1[li] = prod(s, r[li], r[li]);
    }
    s->part[blk * 4 + 0] = tree64(v0, s->mixed);
    s->part[blk * 4 + 1] = tree64(v1, s->mixed);
  }
}

/* Returns whether the solver keeps going. */
static int o_solve_beta(osim* s) {
  uint32_t nb = n_blocks_(s);
  int dbl = s->mixed;
  double rz = total(s, nb, 0, dbl), rr = total(s, nb, 1, dbl);
  s->st.solve_iters += 1u;
  if (dbl) {
    s->r_norm = sqrt(rr);
    s->beta = s->rz > 0.0 ? rz / s->rz : 0.0;
  } else {
    s->r_norm = (double)sqrt_r((REAL)rr);
    s->beta = s->rz > 0.0 ? (double)((REAL)rz / (REAL)s->rz) : 0.0;
  }
  s->rz = rz;
  int done = converged(s, rr);
  int go = 1;
  if (done || s->st.solve_iters >= s->U.limits[2]) {
    s->st.iters_max = umax(s->st.iters_max, s->st.solve_iters);
    s->st.unconverged += done ? 0u : 1u;
    go = 0;
  }
  sync_state_floats(s);
  return go;
}

/* ---- pipelined CG (NATRAY_CG_PIPELINED; not in the reference) ---------------------------------
 *
 * Ghysels and Vanroose's pipelined preconditioned CG (Parallel Computing 40, 2014, Algorithm 3):
 * the same Krylov iterates as o_solve_q .. o_solve_beta in exact arithmetic, with one
 * reduction per iteration, of (r, u), (w, u) and (r, r), where u = M^-1 r and w = A u are
 * carried by recurrences. Each iteration computes m = M^-1 w and n = A m, then
 *   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
 *   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
 *   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z,
 * with gamma = (r, u), delta = (w, u). Every refresh_every iterations the residual is replaced
 * (Cools et al., SIAM J. Sci. Comput. 40, 2018): r = b - A x in double, u = M^-1 r, w = A u,
 * s = A p, q = M^-1 s, z = A q. Dots, norms and the tolerance test are in double, vectors in
 * REAL. The fields follow the memory form of the GPU solve: x in P, r in R, u in Z, s in Q,
 * p in D, w in D + 1; z, q and m in pz, pq, pm. The iteration stops when (r, r) meets the
 * tolerance (tested from the second reduction on; the first is the start's) or after
 * limits.z updates.
 *
 * A replacement brings r, u, w, s, q and z back to what x and p give; made once r has fallen
 * to the rounding gap between the recurred and the true residual, it breaks the recurrences'
 * relations (alpha's denominator above assumes them). So the residual is replaced only while
 * |r| > PIPE_REPLACE tol max(|b|, 1): then the gap is small against r, and the gap that builds
 * up afterwards starts from residuals that small. */
#define PIPE_REPLACE 1000.0

/* (A v) at water cell `cell`: diag v minus the water neighbours' v, in the order -x, +x, -y,
 * +y, -z, +z (the operator of o_solve_q). */
static REAL apply_a(const osim* s, const REAL* v, uint32_t cell) {
  REAL q = cfv(s, NATRAY_F_DIAG, cell) * v[cell];
  for (uint32_t d = 0; d < 3; ++d) {
    int e[3] = {0, 0, 0};
    e[d] = 1;
    for (int k = -1; k <= 1; k += 2) {
      uint32_t nbc = neighbour(s, cell, k * e[0], k * e[1], k * e[2]);
      if (cell_type(s, nbc) == WATER) q -= v[nbc];
    }
  }
  return q;
}

static inline REAL* field(osim* s, uint32_t f) { return s->cf + (size_t)f * cap_cells(s); }

/* M^-1 of field `src` over block blk into dst (water cells only). */
static void block_precondition(osim* s, uint32_t blk, const REAL* src, REAL* dst) {
  REAL v[64], diag[64], z[64];
  int water[64];
  for (uint32_t li = 0; li < 64; ++li) {
    uint32_t cell = blk * 64u + li;
    water[li] = cfv(s, NATRAY_F_TYPE, cell) == WATER;
    v[li] = water[li] ? src[cell] : (REAL)0;
    diag[li] = water[li] ? cfv(s, NATRAY_F_DIAG, cell) : (REAL)1;
  }
  precondition(s, v, diag, water, z);
  for (uint32_t li = 0; li < 64; ++li)
    if (water[li]) dst[blk * 64u + li] = z[li];
}

/* m = M^-1 w and the block partials (r, u), (w, u), (r, r) of block blk. */
static void pipe_block_m(osim* s, uint32_t blk) {
  const REAL *R = field(s, NATRAY_F_R), *U = field(s, NATRAY_F_Z), *W = field(s, NATRAY_F_D + 1u);
  block_precondition(s, blk, W, s->pm);
  double v0[64], v1[64], v2[64];
  for (uint32_t li = 0; li < 64; ++li) {
    uint32_t cell = blk * 64u + li;
    int water = cfv(s, NATRAY_F_TYPE, cell) == WATER;
    REAL r = water ? R[cell] : (REAL)0, u = water ? U[cell] : (REAL)0, w = water ? W[cell] : (REAL)0;
    v0[li] = prod(s, r, u);
    v1[li] = prod(s, w, u);
    v2[li] = prod(s, r, r);
  }
  s->part[blk * 4 + 0] = tree64(v0, 1);
  s->part[blk * 4 + 1] = tree64(v1, 1);
  s->part[blk * 4 + 2] = tree64(v2, 1);
}

/* Returns whether to update: false once (r, r) meets the tolerance or limits.z updates ran. */
static int pipe_scalars(osim* s, uint32_t i, double g, double dl, double rr) {
  s->r_norm = sqrt(rr);
  if (i > 0) {
    int done = converged(s, rr);
    if (done || i >= s->U.limits[2]) {
      s->st.iters_max = umax(s->st.iters_max, i);
      s->st.unconverged += done ? 0u : 1u;
      s->rz = g;
      return 0;
    }
  }
  double beta = (i > 0 && s->rz > 0.0) ? g / s->rz : 0.0;
  double den = (i > 0 && s->alpha != 0.0) ? dl - beta * g / s->alpha : dl;
  s->alpha = den > 0.0 ? g / den : 0.0;
  s->beta = beta;
  s->rz = g;
  return 1;
}

static void o_solve_pipelined(osim* s) {
  uint32_t nb = n_blocks_(s), nc = nb * 64u;
  REAL *X = field(s, NATRAY_F_P), *R = field(s, NATRAY_F_R), *U = field(s, NATRAY_F_Z), *S = field(s, NATRAY_F_Q);
  REAL *Pd = field(s, NATRAY_F_D), *W = field(s, NATRAY_F_D + 1u);
  REAL *Z = s->pz, *Q = s->pq, *M = s->pm;
  /* w = A u; z, q, s, p start at zero. */
#pragma omp parallel for schedule(static)
  for (uint32_t cell = 0; cell < nc; ++cell) {
    if (cfv(s, NATRAY_F_TYPE, cell) != WATER) continue;
    W[cell] = apply_a(s, U, cell);
    Z[cell] = Q[cell] = S[cell] = Pd[cell] = 0;
  }
  for (uint32_t i = 0;; ++i) {
#pragma omp parallel for schedule(static)
    for (uint32_t blk = 0; blk < nb; ++blk) pipe_block_m(s, blk);
    double g = total(s, nb, 0, 1), dl = total(s, nb, 1, 1), rr = total(s, nb, 2, 1);
    s->st.solve_iters = i;
    int go = pipe_scalars(s, i, g, dl, rr);
    sync_state_floats(s);
    if (!go) break;
    REAL a = (REAL)s->alpha, b = (REAL)s->beta;
#pragma omp parallel for schedule(static)
    for (uint32_t cell = 0; cell < nc; ++cell) {
      if (cfv(s, NATRAY_F_TYPE, cell) != WATER) continue;
      REAL n = apply_a(s, M, cell);
      Z[cell] = n + b * Z[cell];
      Q[cell] = M[cell] + b * Q[cell];
      S[cell] = W[cell] + b * S[cell];
      Pd[cell] = U[cell] + b * Pd[cell];
      X[cell] = X[cell] + a * Pd[cell];
      R[cell] = R[cell] - a * S[cell];
      U[cell] = U[cell] - a * Q[cell];
      W[cell] = W[cell] - a * Z[cell];
    }
    s->cg_iters++;
    if (s->refresh && (i + 1u) % s->sc.refresh_every == 0u &&
        s->r_norm > PIPE_REPLACE * (double)s->U.solve[0] * (s->b_norm < 1.0 ? 1.0 : s->b_norm)) {
#pragma omp parallel for schedule(static)
      for (uint32_t cell = 0; cell < nc; ++cell)
        if (cfv(s, NATRAY_F_TYPE, cell) == WATER) R[cell] = true_residual(s, cell);
#pragma omp parallel for schedule(static)
      for (uint32_t blk = 0; blk < nb; ++blk) block_precondition(s, blk, R, U);
#pragma omp parallel for schedule(static)
      for (uint32_t cell = 0; cell < nc; ++cell) {
        if (cfv(s, NATRAY_F_TYPE, cell) != WATER) continue;
        W[cell] = apply_a(s, U, cell);
        S[cell] = apply_a(s, Pd, cell);
      }
#pragma omp parallel for schedule(static)
      for (uint32_t blk = 0; blk < nb; ++blk) block_precondition(s, blk, S, Q);
#pragma omp parallel for schedule(static)
      for (uint32_t cell = 0; cell < nc; ++cell)
        if (cfv(s, NATRAY_F_TYPE, cell) == WATER) Z[cell] = apply_a(s, Q, cell);
    }
  }
}

static void o_solve(osim* s) {
  if (s->st.dt_ns == 0u) return;
  o_solve_init(s);
  s->use_guess = 0;
  int go = o_solve_start(s) && s->U.limits[2] > 0u;
  if (s->sc.precision == NATRAY_CG_PIPELINED) {
    if (go) o_solve_pipelined(s);
    return;
  }
  while (go) {
    o_solve_q(s);
    o_solve_alpha(s);
    o_solve_update(s);
    go = o_solve_beta(s);
    s->cg_iters++;
  }
}

/* ---- project (fluid.wgsl:580-598) ------------------------------------------------------------ */

static void o_project(osim* s) {
  if (s->st.dt_ns == 0u) return;
  uint32_t nc = n_blocks_(s) * 64u;
  REAL k = dt_(s) / ((REAL)s->U.fluid[0] * hh_(s));
#pragma omp parallel for schedule(static)
  for (uint32_t cell = 0; cell < nc; ++cell) {
    REAL ty = cfv(s, NATRAY_F_TYPE, cell);
    for (uint32_t d = 0; d < 3; ++d) {
      int e[3] = {0, 0, 0};
      e[d] = 1;
      uint32_t nbc = neighbour(s, cell, -e[0], -e[1], -e[2]);
      REAL nt = cell_type(s, nbc);
      REAL u = bounded(s, cfv(s, NATRAY_F_U + d, cell), ty, nt, cell, nbc, d);
      if (ty != SOLID && nt != SOLID && (ty == WATER || nt == WATER)) u -= k * (pressure_of(s, cell) - pressure_of(s, nbc));
      *cfp(s, NATRAY_F_U + d, cell) = u;
    }
  }
}

/* ---- extend_1, extend_2 (fluid.wgsl:600-674) --------------------------------------------------- */

static inline int face_set(const osim* s, uint32_t cell, uint32_t d) {
  if (cell == NONE) return 0;
  int e[3] = {0, 0, 0};
  e[d] = 1;
  return cfv(s, NATRAY_F_TYPE, cell) == WATER || cell_type(s, neighbour(s, cell, -e[0], -e[1], -e[2])) == WATER;
}

static void o_extend_1(osim* s) {
  uint32_t nc = n_blocks_(s) * 64u;
#pragma omp parallel for schedule(static)
  for (uint32_t cell = 0; cell < nc; ++cell) {
    uint32_t done = 0;
    for (uint32_t d = 0; d < 3; ++d) {
      REAL u = cfv(s, NATRAY_F_U + d, cell);
      if (face_set(s, cell, d)) {
        done |= 1u << d;
      } else {
        REAL sum = 0, n = 0;
        for (int k = 0; k < 6; ++k) {
          int e[3] = {0, 0, 0};
          e[k / 2] = (k % 2 == 1) ? 1 : -1;
          uint32_t nbc = neighbour(s, cell, e[0], e[1], e[2]);
          if (face_set(s, nbc, d)) {
            sum += cfv(s, NATRAY_F_U + d, nbc);
            n += (REAL)1;
          }
        }
        if (n > (REAL)0) {
          u = sum / n;
          done |= 1u << d;
        }
      }
      *cfp(s, NATRAY_F_EXT + d, cell) = u;
    }
    *cfp(s, NATRAY_F_EXT_SET, cell) = (REAL)done;
  }
}

static inline int ext_set(const osim* s, uint32_t cell, uint32_t d) {
  return cell != NONE && ((((uint32_t)cfv(s, NATRAY_F_EXT_SET, cell)) >> d) & 1u) != 0u;
}

static void o_extend_2(osim* s) {
  uint32_t nc = n_blocks_(s) * 64u;
#pragma omp parallel for schedule(static)
  for (uint32_t cell = 0; cell < nc; ++cell) {
    REAL ty = cfv(s, NATRAY_F_TYPE, cell);
    for (uint32_t d = 0; d < 3; ++d) {
      REAL u = cfv(s, NATRAY_F_U + d, cell);
      if (ext_set(s, cell, d)) {
        u = cfv(s, NATRAY_F_EXT + d, cell);
      } else {
        REAL sum = 0, n = 0;
        for (int k = 0; k < 6; ++k) {
          int e[3] = {0, 0, 0};
          e[k / 2] = (k % 2 == 1) ? 1 : -1;
          uint32_t nbc = neighbour(s, cell, e[0], e[1], e[2]);
          if (ext_set(s, nbc, d)) {
            sum += cfv(s, NATRAY_F_EXT + d, nbc);
            n += (REAL)1;
          }
        }
        if (n > (REAL)0) u = sum / n;
      }
      int e[3] = {0, 0, 0};
      e[d] = 1;
      uint32_t lo = neighbour(s, cell, -e[0], -e[1], -e[2]);
      *cfp(s, NATRAY_F_U + d, cell) = bounded(s, u, ty, cell_type(s, lo), cell, lo, d);
    }
  }
}

/* ---- g2p (fluid.wgsl:767-901) ------------------------------------------------------------------ */

typedef struct { REAL s, g[3], w; } gathered;

/* gather (fluid.wgsl:780-802) */
static gathered gather(const osim* s, const REAL* tg, const REAL x[3], const REAL o[3], i3 corner, uint32_t fa, uint32_t fw) {
  int base[3];
  REAL fx[3], wx[3], wy[3], wz[3];
  for (int k = 0; k < 3; ++k) {
    base[k] = (int)rfloor(x[k] - o[k] - (REAL)0.5);
    fx[k] = x[k] - o[k] - (REAL)base[k];
  }
  axis_weights(fx[0], wx);
  axis_weights(fx[1], wy);
  axis_weights(fx[2], wz);
  uint32_t t0 = (uint32_t)(base[0] - corner.x) * 49u + (uint32_t)(base[1] - corner.y) * 7u + (uint32_t)(base[2] - corner.z);
  gathered r = {0, {0, 0, 0}, 0};
  for (uint32_t a = 0; a < 3; ++a)
    for (uint32_t b = 0; b < 3; ++b)
      for (uint32_t c = 0; c < 3; ++c) {
        REAL weight = wx[a] * wy[b] * wz[c];
        uint32_t t = t0 + a * 49u + b * 7u + c;
        REAL u = weight * tg[fa * TILE + t];
        r.s += u;
        r.g[0] += u * ((REAL)a - fx[0]);
        r.g[1] += u * ((REAL)b - fx[1]);
        r.g[2] += u * ((REAL)c - fx[2]);
        r.w += weight * tg[fw * TILE + t];
      }
  REAL h = hh_(s);
  r.g[0] *= h;
  r.g[1] *= h;
  r.g[2] *= h;
  return r;
}

/* carve for the moving capsule (the role of bullet.wgsl:395-424): a particle that ends a step
 * inside the capsule goes onto its surface and takes at least the capsule's normal speed. */
static void carve(const osim* s, REAL x[3], REAL v[3]) {
  REAL r = (REAL)s->sc.capsule_a[3];
  r3 p = {x[0], x[1], x[2]}, on, ab;
  REAL d2 = capsule_d2(s, p, now_(s), &on, &ab);
  if (d2 >= r * r) return;
  REAL d[3] = {x[0] - on.x, x[1] - on.y, x[2] - on.z};
  if (d2 < (REAL)1e-14) {
    /* On the axis: push out along a fixed direction perpendicular to it. */
    REAL ax = ab.x, ay = ab.y, az = ab.z;
    /* cross(ab, (0, 1, 0)) or, for an axis along y, cross(ab, (1, 0, 0)) */
    if (ax * ax + az * az > (REAL)1e-20) {
      d[0] = -az; d[1] = 0; d[2] = ax;
    } else {
      d[0] = 0; d[1] = az; d[2] = -ay;
    }
    d2 = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
  }
  REAL len = sqrt_r(d2);
  REAL dir[3] = {d[0] / len, d[1] / len, d[2] / len};
  x[0] = on.x + dir[0] * r;
  x[1] = on.y + dir[1] * r;
  x[2] = on.z + dir[2] * r;
  REAL wall = ((REAL)s->sc.capsule_v[0] * dir[0] + (REAL)s->sc.capsule_v[1] * dir[1]) + (REAL)s->sc.capsule_v[2] * dir[2];
  REAL vn = (v[0] * dir[0] + v[1] * dir[1]) + v[2] * dir[2];
  REAL push = rmax(wall - vn, (REAL)0);
  v[0] += dir[0] * push;
  v[1] += dir[1] * push;
  v[2] += dir[2] * push;
}

/* g2p_one (fluid.wgsl:806-861). Returns the speed, 0 for a particle that died. */
/* ---- spray and wet ground (fluid.wgsl:676-765, sim_defs.wgsl:233-240; NATRAY_SCENE_SPRAY) ----- */

#define RHO_AIR 1.2          /* sim_defs.wgsl:101 */
#define MU_AIR 1.8e-5        /* sim_defs.wgsl:102 */
#define EVAPORATION 1.3e-10  /* fluid.wgsl:684 */
#define WE_CRIT 12.0         /* fluid.wgsl:718 */
#define PI_C 3.14159265358979 /* common.wgsl:1 */
#define BREAKUP_DROPS 4u     /* fluid.wgsl:719 */

static inline REAL r_exp(REAL x) { return IS_F64 ? (REAL)nm_exp_d((double)x) : (REAL)nm_expf((float)x); }
static inline REAL r_log(REAL x) { return IS_F64 ? (REAL)nm_log_d((double)x) : (REAL)nm_logf((float)x); }
static inline REAL r_cos(REAL x) { return IS_F64 ? (REAL)nm_cos_d((double)x) : (REAL)nm_cosf((float)x); }
static inline REAL r_pow(REAL x, REAL y) {
  if (!IS_F64) return (REAL)nm_powf((float)x, (float)y);
  return x == 0 ? (y > 0 ? (REAL)0 : (REAL)HUGE_VAL) : (REAL)nm_exp_d((double)y * nm_log_d((double)x));
}

/* pcg, hash3, rnd, rnd_gauss (common.wgsl:3-22) */
static inline uint32_t pcg(uint32_t v) {
  uint32_t t = v * 747796405u + 2891336453u;
  uint32_t w = ((t >> ((t >> 28u) + 4u)) ^ t) * 277803737u;
  return (w >> 22u) ^ w;
}
static inline uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) { return pcg(a ^ pcg(b ^ pcg(c))); }
static inline REAL rnd(uint32_t* seed) {
  *seed = pcg(*seed);
  return (REAL)(*seed >> 8u) * (REAL)(1.0 / 16777216.0);
}
static inline REAL rnd_gauss(uint32_t* seed) {
  REAL u1 = rmax(rnd(seed), (REAL)1e-7);
  REAL u2 = rnd(seed);
  return sqrt_r((REAL)-2 * r_log(u1)) * r_cos((REAL)2 * (REAL)PI_C * u2);
}

/* now() (sim_defs.wgsl:171-173) */
static inline REAL now_r(const osim* s) { return (REAL)s->st.t_ns * (REAL)1e-9; }

/* ground_touch (sim_defs.wgsl:233-240): the first time water touched the ground cell under p,
 * as the f32 bits of now() (an atomic minimum). The cell index is clamped to the map, which
 * only matters for map sides that are not powers of two. */
static void ground_touch(osim* s, REAL px, REAL pz) {
  REAL hw = (REAL)s->U.wet[0];
  uint32_t res = s->wet_res;
  REAL ux = (px + hw) / ((REAL)2 * hw), uz = (pz + hw) / ((REAL)2 * hw);
  if (ux < (REAL)0 || uz < (REAL)0 || ux >= (REAL)1 || uz >= (REAL)1 || res == 0u) return;
  uint32_t cx = umin((uint32_t)(ux * (REAL)res), res - 1u), cz = umin((uint32_t)(uz * (REAL)res), res - 1u);
  float t = (float)now_r(s);
  uint32_t bits;
  memcpy(&bits, &t, 4);
  uint32_t* w = &s->touch[(size_t)cz * res + cx];
  uint32_t cur = __atomic_load_n(w, __ATOMIC_RELAXED);
  while (bits < cur && !__atomic_compare_exchange_n(w, &cur, bits, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

/* break_up (fluid.wgsl:720-765): a particle that leaves the resolved flow becomes four drops
 * and a parcel of finer drops, in five consecutive spray slots taken with an atomic add. */
static void break_up(osim* s, const REAL x[3], const REAL v[3], uint32_t id) {
  uint32_t first = __atomic_fetch_add(&s->st.spray_count, BREAKUP_DROPS + 1u, __ATOMIC_RELAXED);
  uint32_t cap = s->spray_cap;
  if (first >= cap) return;
  uint32_t seed = hash3(s->st.step, id, 0xb4eau);
  REAL hs = (REAL)s->U.fluid[3];
  REAL vol = hs * hs * hs;
  REAL speed = sqrt_r((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
  /* breakup_diameter (fluid.wgsl:713-716) */
  REAL bd = ((REAL)WE_CRIT * (REAL)NATRAY_SIGMA_W) / ((REAL)RHO_AIR * rmax(speed * speed, (REAL)1e-4));
  REAL d_med = rmin((REAL)0.55 * hs, (REAL)0.5 * bd);
  REAL dd[BREAKUP_DROPS];
  REAL used = 0;
  for (uint32_t k = 0; k < BREAKUP_DROPS; ++k) {
    dd[k] = rmin(rmax(d_med * r_exp((REAL)0.5 * rnd_gauss(&seed)), (REAL)2e-5), (REAL)1.24 * hs);
    used += (REAL)0.5236 * dd[k] * dd[k] * dd[k];
  }
  REAL df = (REAL)0.5 * d_med;
  int fine = df < (REAL)1e-4;
  REAL fit = r_pow(vol / used, (REAL)1 / (REAL)3);
  REAL scale = fine ? rmin((REAL)1, fit) : fit;
  REAL tnow = now_r(s);
  for (uint32_t k = 0; k < BREAKUP_DROPS; ++k) {
    uint32_t slot = first + k;
    if (slot >= cap) return;
    REAL dk = dd[k] * scale;
    REAL spread = sqrt_r((REAL)NATRAY_SIGMA_W / ((REAL)s->U.fluid[0] * dk)) + (REAL)0.05 * speed;
    REAL g0 = rnd_gauss(&seed);
    REAL g1 = rnd_gauss(&seed);
    REAL g2 = rnd_gauss(&seed);
    REAL w0 = g0 + (REAL)1e-6, w1 = g1 + (REAL)1e-6, w2 = g2 + (REAL)1e-6;
    REAL len = sqrt_r((w0 * w0 + w1 * w1) + w2 * w2);
    REAL u[3] = {w0 / len, w1 / len, w2 / len};
    REAL o0 = rnd(&seed);
    REAL o1 = rnd(&seed);
    REAL o2 = rnd(&seed);
    REAL rv = rnd(&seed);
    REAL* rec = s->spray + (size_t)slot * 12;
    rec[0] = x[0] + (o0 - (REAL)0.5) * hs;
    rec[1] = x[1] + (o1 - (REAL)0.5) * hs;
    rec[2] = x[2] + (o2 - (REAL)0.5) * hs;
    rec[3] = (REAL)0.5 * dk;
    for (int c = 0; c < 3; ++c) rec[4 + c] = v[c] + u[c] * spread * rv;
    rec[7] = 1;
    rec[8] = rec[9] = rec[10] = 0;
    rec[11] = tnow;
  }
  uint32_t slot = first + BREAKUP_DROPS;
  if (slot >= cap) return;
  REAL n = (vol - used * scale * scale * scale) / ((REAL)0.5236 * df * df * df);
  REAL* rec = s->spray + (size_t)slot * 12;
  rec[0] = x[0];
  rec[1] = x[1];
  rec[2] = x[2];
  rec[3] = (fine && n >= (REAL)2) ? (REAL)0.5 * df : (REAL)0;
  rec[4] = v[0];
  rec[5] = v[1];
  rec[6] = v[2];
  rec[7] = rmax(n, (REAL)2);
  rec[8] = rec[9] = rec[10] = 0;
  rec[11] = tnow;
}

/* drag_cd (fluid.wgsl:676-679) */
static REAL drag_cd(REAL re) {
  if (re < (REAL)1e-4) return (REAL)24 / (REAL)1e-4;
  return (REAL)24 / re * ((REAL)1 + (REAL)0.15 * r_pow(re, (REAL)0.687)) +
         (REAL)0.42 / ((REAL)1 + (REAL)42500 * r_pow(re, (REAL)-1.16));
}

/* ---- bullets (bullet.wgsl:85-424): the water's side of the reference's bullet model ---------------
 *
 * One bullet at a time, as a workgroup of 256 does it: "thread" li sums patches li, li + 256,
 * ... in order, the 256 partial sums go through the reference's tree (bullet.wgsl:290-298), and
 * the body is stepped from the root. The rubber's impulses (bullet_hit) are zero: rubber_contact
 * belongs to the membrane's own substeps. Every vector expression follows the WGSL order. */

#define B_SEC NATRAY_BULLET_SECTIONS
#define R_HAND ((REAL)1.5)
#define LOGVINOVICH_K ((REAL)2)
#define CP_WAKE ((REAL)-0.4)
#define CF_SKIN ((REAL)0.003)

static inline r3 v3(REAL x, REAL y, REAL z) { r3 r = {x, y, z}; return r; }
static inline r3 vxyz(const REAL* a) { return v3(a[0], a[1], a[2]); }
static inline r3 vadd(r3 a, r3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline r3 vsub(r3 a, r3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline r3 vmul(r3 a, REAL k) { return v3(a.x * k, a.y * k, a.z * k); }
static inline r3 vdiv(r3 a, REAL k) { return v3(a.x / k, a.y / k, a.z / k); }
static inline r3 vadds(r3 a, REAL k) { return v3(a.x + k, a.y + k, a.z + k); }
static inline REAL vdot(r3 a, r3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline r3 vcross(r3 a, r3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline REAL vlen(r3 a) { return sqrt_r(vdot(a, a)); }
static inline r3 vnorm(r3 a) { return vdiv(a, vlen(a)); }
/* qrot, qmul (bullet.wgsl:91-98); q = (x, y, z, w) */
static inline r3 qrot(const REAL q[4], r3 v) {
  r3 qv = vxyz(q);
  r3 t = vmul(vcross(qv, v), (REAL)2);
  return vadd(vadd(v, vmul(t, q[3])), vcross(qv, t));
}
static inline void qmul(const REAL a[4], const REAL b[4], REAL out[4]) {
  r3 r = vadd(vadd(vmul(vxyz(b), a[3]), vmul(vxyz(a), b[3])), vcross(vxyz(a), vxyz(b)));
  REAL w = a[3] * b[3] - vdot(vxyz(a), vxyz(b));
  out[0] = r.x, out[1] = r.y, out[2] = r.z, out[3] = w;
}
static inline REAL r_sin(REAL x) { return IS_F64 ? (REAL)nm_sin_d((double)x) : (REAL)nm_sinf((float)x); }
static inline REAL r_asin(REAL x) { return IS_F64 ? (REAL)nm_asin_d((double)x) : (REAL)nm_asinf((float)x); }
static inline REAL* section_of(const osim* s, uint32_t bi, uint32_t k) { return &s->sections[((size_t)bi * B_SEC + k) * 12]; }
/* section_r2 (bullet.wgsl:152-154): c.x + b.w max(t - a.w, 0) */
static inline REAL section_r2(const REAL* sc, REAL t) { return sc[8] + sc[7] * rmax(t - sc[3], (REAL)0); }
static inline REAL hand2_of(REAL hh) { return R_HAND * R_HAND * hh * hh; }

/* cp_patch (bullet.wgsl:100-111) */
static REAL cp_patch(REAL sn) {
  if (sn < (REAL)0) return CP_WAKE * sn * sn;
  REAL beta = r_asin(rmin(sn, (REAL)1));
  if (beta < (REAL)PI_C / (REAL)12) return (REAL)2.13 * sn * sn;
  REAL x = beta / (REAL)PI_C - (REAL)0.25;
  return (REAL)0.5 + (REAL)1.81 * x - (REAL)2 * x * x;
}

/* sample_pressure (bullet.wgsl:113-134) */
static void sample_pressure(const osim* s, const REAL x[3], int gas_only, REAL out[2]) {
  REAL w[3][3];
  int base[3];
  for (int k = 0; k < 3; ++k) {
    REAL bf = rfloor(x[k] - (REAL)0.5);
    base[k] = (int)bf;
    axis_weights(x[k] - bf, w[k]);
  }
  REAL sx = 0, sy = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        i3 cc = {base[0] + a, base[1] + b, base[2] + c};
        uint32_t cell = cell_at(s, cc);
        if (cell == NONE) continue;
        REAL ty = cfv(s, NATRAY_F_TYPE, cell);
        if (ty == SOLID || (gas_only && ty == WATER)) continue;
        REAL wt = w[0][a] * w[1][b] * w[2][c];
        sx = sx + wt * cfv(s, NATRAY_F_P, cell);
        sy = sy + wt;
      }
  out[0] = sx;
  out[1] = sy;
}

/* push_water (bullet.wgsl:136-150) */
static void push_water(osim* s, const REAL X[3], uint32_t dim, REAL j, REAL sw) {
  if (sw <= (REAL)1e-6 || j == (REAL)0) return;
  REAL w[3][3];
  int base[3];
  for (int k = 0; k < 3; ++k) {
    REAL lx = X[k] - ((uint32_t)k == dim ? (REAL)0 : (REAL)0.5);
    REAL bf = rfloor(lx - (REAL)0.5);
    base[k] = (int)bf;
    axis_weights(lx - bf, w[k]);
  }
  REAL hs = (REAL)s->U.fluid[3];
  REAL kk = j / ((REAL)s->U.fluid[0] * hs * hs * hs * sw);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        i3 cc = {base[0] + a, base[1] + b, base[2] + c};
        uint32_t cell = cell_at(s, cc);
        if (cell == NONE || accv(s, NATRAY_A_WT(dim), cell) <= 0) continue;
        *accp(s, NATRAY_A_BU + dim, cell) += to_fp(kk * w[0][a] * w[1][b] * w[2][c], (REAL)NATRAY_FP_M);
      }
}

/* in_cavity (bullet.wgsl:156-170) */
static int in_cavity(const osim* s, uint32_t bi, r3 p, REAL t) {
  REAL hand2 = hand2_of(hh_(s));
  for (uint32_t k = 0; k < B_SEC; ++k) {
    const REAL* sc = section_of(s, bi, k);
    if (sc[7] <= (REAL)0 || section_r2(sc, t) >= hand2) continue;
    r3 a = vxyz(sc), ab = vsub(vxyz(sc + 4), a);
    REAL f = vdot(vsub(p, a), ab) / rmax(vdot(ab, ab), (REAL)1e-20);
    if (f < (REAL)0 || f > (REAL)1) continue;
    r3 d = vsub(p, vadd(a, vmul(ab, f)));
    if (vdot(d, d) < section_r2(sc, t)) return 1;
  }
  return 0;
}

/* bullet_step (bullet.wgsl:179-377) for bullet bi */
static void bullet_step_one(osim* s, uint32_t bi) {
  struct obullet* B = &s->bullets[bi];
  const struct obullet b = *B;
  const REAL step = dt_(s), t_end = now_(s), hh = hh_(s);
  const r3 origin = v3((REAL)s->U.origin[0], (REAL)s->U.origin[1], (REAL)s->U.origin[2]);
  const r3 bx = vxyz(b.x), bv = vxyz(b.v);
  uint32_t state = b.info[2];
  if (state == 0u && s->st.t_ns - s->st.dt_ns >= b.info[0]) state = 1u;
  REAL pf;
  {
    r3 ahead = vadd(bx, vmul(vnorm(vadds(bv, (REAL)1e-9)), b.shape[2] + (REAL)3 * hh));
    r3 xa = vdiv(vsub(ahead, origin), hh);
    REAL xs[3] = {xa.x - (REAL)0.5, xa.y - (REAL)0.5, xa.z - (REAL)0.5}, sp[2];
    sample_pressure(s, xs, 0, sp);
    pf = sp[1] > (REAL)1e-6 ? sp[0] / sp[1] : (REAL)0;
  }
  REAL* pose = &s->pose_log[((size_t)bi * NATRAY_BULLET_LOG + s->st.step % NATRAY_BULLET_LOG) * 12];
  if (state != 1u) {
    B->info[2] = state;
    pose[0] = pose[1] = pose[2] = pose[3] = 0;
    return;
  }
  const REAL conj[4] = {-b.q[0], -b.q[1], -b.q[2], b.q[3]};
  const r3 lb = qrot(conj, vxyz(b.l));
  const r3 w = qrot(b.q, v3(lb.x / b.shape[1], lb.y / b.shape[1], lb.z / b.shape[0]));
  const REAL rho = (REAL)s->U.fluid[0];
  const r3 vhat = vnorm(vadds(bv, (REAL)1e-9));
  REAL red_f[256][4], red_t[256][4], red_c[256][4];
  for (uint32_t li = 0; li < 256; ++li) {
    r3 f = v3(0, 0, 0), tq = v3(0, 0, 0), cen = v3(0, 0, 0);
    REAL cav_x = 0, cav_y = 0, cen_w = (REAL)3.0e38;
    for (uint32_t k = li; k < b.info[1]; k += 256u) {
      const REAL* pt = &s->patches[((size_t)bi * NATRAY_BULLET_PATCHES + k) * 8];
      r3 r = qrot(b.q, vxyz(pt));
      r3 p = vadd(bx, r);
      r3 n = qrot(b.q, vxyz(pt + 4));
      REAL area = pt[3];
      r3 Xv = vdiv(vsub(p, origin), hh);
      REAL X[3] = {Xv.x, Xv.y, Xv.z}, uw[3], sw[3], sg[2];
      for (uint32_t dim = 0; dim < 3; ++dim) {
        REAL xs[3] = {X[0] - (dim == 0 ? (REAL)0 : (REAL)0.5), X[1] - (dim == 1 ? (REAL)0 : (REAL)0.5),
                      X[2] - (dim == 2 ? (REAL)0 : (REAL)0.5)};
        sample_grid(s, xs, NATRAY_F_U + dim, NATRAY_A_WT(dim), sg);
        uw[dim] = sg[1] > (REAL)1e-6 ? sg[0] / sg[1] : (REAL)0;
        sw[dim] = sg[1];
      }
      REAL xc[3] = {X[0] - (REAL)0.5, X[1] - (REAL)0.5, X[2] - (REAL)0.5};
      sample_grid(s, xc, NATRAY_F_TYPE, NATRAY_A_WC, sg);
      REAL fill = rmin(rmax(sg[1] / (REAL)8, (REAL)0), (REAL)1);
      if (in_cavity(s, bi, p, t_end - step)) fill = 0;
      REAL gs[2];
      sample_pressure(s, xc, 1, gs);
      REAL p_gas = gs[1] > (REAL)1e-6 ? gs[0] / gs[1] : (REAL)0;
      r3 vp = vadd(bv, vcross(w, r));
      r3 rel = vsub(vp, v3(uw[0], uw[1], uw[2]));
      REAL sp = vlen(rel);
      REAL wn = vdot(rel, n);
      REAL qd = (REAL)0.5 * rho * sp * sp;
      REAL p_water = rmax((REAL)NATRAY_P_FLOOR, pf + qd * cp_patch(wn / rmax(sp, (REAL)1e-9)));
      r3 fw = vmul(n, -fill * p_water * area);
      r3 rt = vsub(rel, vmul(n, wn));
      REAL rtl = vlen(rt);
      if (rtl > (REAL)1e-9) fw = vsub(fw, vmul(vmul(rt, fill * CF_SKIN * (REAL)0.5 * rho * rtl), area));
      r3 fp = vsub(fw, vmul(n, ((REAL)1 - fill) * rmax(p_gas, (REAL)NATRAY_P_FLOOR) * area));
      f = vadd(f, fp);
      tq = vadd(tq, vcross(r, fp));
      push_water(s, X, 0, -fw.x * step, sw[0]);
      push_water(s, X, 1, -fw.y * step, sw[1]);
      push_water(s, X, 2, -fw.z * step, sw[2]);
      if (wn > (REAL)0 && fill > (REAL)0) {
        REAL face = fill * area * rmax(vdot(n, vhat), (REAL)0);
        cav_x = cav_x + -vdot(fw, vhat);
        cav_y = cav_y + face;
        cen = vadd(cen, vmul(p, face));
        cen_w = rmin(cen_w, vdot(r, vhat));
      }
    }
    red_f[li][0] = f.x, red_f[li][1] = f.y, red_f[li][2] = f.z, red_f[li][3] = cav_x;
    red_t[li][0] = tq.x, red_t[li][1] = tq.y, red_t[li][2] = tq.z, red_t[li][3] = cav_y;
    red_c[li][0] = cen.x, red_c[li][1] = cen.y, red_c[li][2] = cen.z, red_c[li][3] = cen_w;
  }
  for (uint32_t off = 128u; off > 0u; off >>= 1u)
    for (uint32_t li = 0; li < off; ++li)
      for (int c = 0; c < 4; ++c) {
        red_f[li][c] = red_f[li][c] + red_f[li + off][c];
        red_t[li][c] = red_t[li][c] + red_t[li + off][c];
        red_c[li][c] = c < 3 ? red_c[li][c] + red_c[li + off][c] : rmin(red_c[li][c], red_c[li + off][c]);
      }
  const REAL hand2 = hand2_of(hh);
  {
    const REAL mass = b.x[3];
    const r3 hit = v3(0, 0, 0), hit_l = v3(0, 0, 0);
    const r3 force = vadd(vxyz(red_f[0]), vmul(v3((REAL)s->U.gravity[0], (REAL)s->U.gravity[1], (REAL)s->U.gravity[2]), mass));
    const REAL drag = red_f[0][3], face = red_t[0][3];
    const r3 v = vadd(bv, vdiv(vadd(vmul(force, step), hit), mass));
    const r3 x = vadd(bx, vmul(v, step));
    const r3 lnew = vadd(vadd(vxyz(b.l), vmul(vxyz(red_t[0]), step)), hit_l);
    const r3 axis0 = qrot(b.q, v3(0, 0, 1));
    const REAL spin_extra = vdot(lnew, axis0) * ((REAL)1 / b.shape[0] - (REAL)1 / b.shape[1]) * step;
    const REAL cone = vlen(lnew) / b.shape[1] * step;
    REAL qs[4] = {0, 0, r_sin((REAL)0.5 * spin_extra), r_cos((REAL)0.5 * spin_extra)}, qn[4];
    qmul(b.q, qs, qn);
    if (cone > (REAL)1e-12) {
      r3 ax = vmul(vnorm(lnew), r_sin((REAL)0.5 * cone));
      REAL qa[4] = {ax.x, ax.y, ax.z, r_cos((REAL)0.5 * cone)}, t[4];
      qmul(qa, qn, t);
      memcpy(qn, t, sizeof t);
    }
    const REAL ql = sqrt_r(((qn[0] * qn[0] + qn[1] * qn[1]) + qn[2] * qn[2]) + qn[3] * qn[3]);
    for (int k = 0; k < 4; ++k) qn[k] = qn[k] / ql;
    const REAL roll = b.v[3] + vdot(lnew, axis0) / b.shape[0] * step;
    uint32_t info[4] = {b.info[0], b.info[1], 1u, b.info[3]};
    const REAL xa[3] = {x.x, x.y, x.z};
    for (int k = 0; k < 3; ++k)
      if (xa[k] < (REAL)s->U.origin[k] || xa[k] > (REAL)s->U.origin[k] + (REAL)s->U.dims[k] * hh) info[2] = 2u;
    const r3 xc = vdiv(vsub(x, origin), hh);
    const int cb[3] = {(int)rfloor(xc.x) >> 2, (int)rfloor(xc.y) >> 2, (int)rfloor(xc.z) >> 2};
    int near = 0;
    for (int k = 0; k < 27; ++k) {
      i3 nb = {cb[0] + k / 9 - 1, cb[1] + (k / 3) % 3 - 1, cb[2] + k % 3 - 1};
      if (nb.x >= 0 && nb.y >= 0 && nb.z >= 0 && nb.x < (int)s->U.bdims[0] && nb.y < (int)s->U.bdims[1] &&
          nb.z < (int)s->U.bdims[2] && s->block_ptr[block_slot(s, nb)] < n_blocks_(s))
        near = 1;
    }
    const REAL pace = near ? vlen(v) : vlen(v) * (REAL)s->U.fluid[1] / (REAL)4;
    if (drag > (REAL)0 && face > (REAL)0) {
      const r3 mid = vdiv(vxyz(red_c[0]), face);
      const r3 a = vadd(mid, vmul(vhat, red_c[0][3] - vdot(vsub(mid, bx), vhat)));
      const r3 e = vadd(a, vmul(v, step));
      REAL* sn = section_of(s, bi, info[3] % B_SEC);
      sn[0] = a.x, sn[1] = a.y, sn[2] = a.z, sn[3] = t_end - (REAL)0.5 * step;
      sn[4] = e.x, sn[5] = e.y, sn[6] = e.z, sn[7] = (REAL)2 * sqrt_r(drag / ((REAL)PI_C * rho * LOGVINOVICH_K));
      sn[8] = face / (REAL)PI_C, sn[9] = 0, sn[10] = 0, sn[11] = 0;
      info[3] += 1u;
    }
    REAL reach = 0;
    for (uint32_t k = 0; k < B_SEC; ++k) {
      const REAL* sk = section_of(s, bi, k);
      if (sk[7] <= (REAL)0 || section_r2(sk, t_end) >= hand2) continue;
      reach = rmax(reach, rmax(vlen(vsub(vxyz(sk), x)), vlen(vsub(vxyz(sk + 4), x))));
    }
    B->x[0] = x.x, B->x[1] = x.y, B->x[2] = x.z;
    B->v[0] = v.x, B->v[1] = v.y, B->v[2] = v.z, B->v[3] = roll;
    for (int k = 0; k < 4; ++k) B->q[k] = qn[k];
    B->l[0] = lnew.x, B->l[1] = lnew.y, B->l[2] = lnew.z, B->l[3] = pace;
    B->reach[0] = x.x, B->reach[1] = x.y, B->reach[2] = x.z, B->reach[3] = reach > (REAL)0 ? reach + R_HAND * hh : (REAL)0;
    for (int k = 0; k < 4; ++k) B->info[k] = info[k];
    float tbits;
    memcpy(&tbits, &s->st.t_ns, 4);
    pose[0] = x.x, pose[1] = x.y, pose[2] = x.z, pose[3] = (REAL)tbits;
    for (int k = 0; k < 4; ++k) pose[4 + k] = qn[k];
    pose[8] = v.x, pose[9] = v.y, pose[10] = v.z, pose[11] = roll;
  }
  /* Sections thinner than the grid push the water out of the way as they open. */
  for (uint32_t li = 0; li < B_SEC; ++li) {
    const REAL* sk = section_of(s, bi, li);
    REAL r2 = section_r2(sk, t_end);
    if (!(sk[7] > (REAL)0 && r2 < hand2)) continue;
    r3 a = vxyz(sk), e = vxyz(sk + 4);
    REAL qv = (REAL)PI_C * sk[7] * vlen(vsub(e, a));
    r3 m = vdiv(vsub(vmul(vadd(a, e), (REAL)0.5), origin), hh);
    REAL x[3] = {m.x - (REAL)0.5, m.y - (REAL)0.5, m.z - (REAL)0.5}, wt[3][3];
    int base[3];
    for (int k = 0; k < 3; ++k) {
      REAL bf = rfloor(x[k] - (REAL)0.5);
      base[k] = (int)bf;
      axis_weights(x[k] - bf, wt[k]);
    }
    for (int a3 = 0; a3 < 3; ++a3)
      for (int b3 = 0; b3 < 3; ++b3)
        for (int c3 = 0; c3 < 3; ++c3) {
          i3 cc = {base[0] + a3, base[1] + b3, base[2] + c3};
          uint32_t cell = cell_at(s, cc);
          if (cell == NONE) continue;
          *accp(s, NATRAY_A_SRC, cell) += to_fp(qv * wt[0][a3] * wt[1][b3] * wt[2][c3], (REAL)NATRAY_FP_Q);
        }
  }
}

/* bullet_step for every bullet, then bullet_apply (bullet.wgsl:379-386). */
static void o_bullets(osim* s) {
  if (s->st.dt_ns == 0u || !s->n_bullets) return;
  for (uint32_t bi = 0; bi < s->n_bullets; ++bi) bullet_step_one(s, bi);
  uint32_t nc = n_blocks_(s) * 64u;
  for (uint32_t cell = 0; cell < nc; ++cell)
    for (uint32_t d = 0; d < 3; ++d) {
      int32_t du = accv(s, NATRAY_A_BU + d, cell);
      if (du != 0) *cfp(s, NATRAY_F_U + d, cell) += from_fp(du, (REAL)NATRAY_FP_M);
    }
}

/* carve (bullet.wgsl:388-424) */
static void carve_cavity(const osim* s, REAL x[3], REAL v[3], uint32_t id) {
  const REAL t = now_(s), hand2 = hand2_of(hh_(s));
  r3 X = vxyz(x), V = vxyz(v);
  for (uint32_t bi = 0; bi < s->n_bullets; ++bi) {
    const REAL* reach = s->bullets[bi].reach;
    if (reach[3] <= (REAL)0 || vlen(vsub(X, vxyz(reach))) > reach[3]) continue;
    for (uint32_t k = 0; k < B_SEC; ++k) {
      const REAL* sc = section_of(s, bi, k);
      REAL r2 = section_r2(sc, t);
      if (sc[7] <= (REAL)0 || r2 >= hand2) continue;
      r3 a = vxyz(sc), ab = vsub(vxyz(sc + 4), a);
      REAL f = vdot(vsub(X, a), ab) / rmax(vdot(ab, ab), (REAL)1e-20);
      if (f < (REAL)0 || f > (REAL)1) continue;
      r3 on = vadd(a, vmul(ab, f));
      r3 d = vsub(X, on);
      if (vdot(d, d) >= r2) continue;
      if (vdot(d, d) < (REAL)1e-14) {
        uint32_t seed = hash3(id, s->st.step, k);
        REAL g0 = rnd_gauss(&seed);
        REAL g1 = rnd_gauss(&seed);
        REAL g2 = rnd_gauss(&seed);
        d = vcross(vadds(ab, (REAL)1e-9), v3(g0, g1, g2));
      }
      r3 dir = vnorm(d);
      REAL r = sqrt_r(r2);
      X = vadd(on, vmul(dir, r));
      REAL wall = sc[7] / ((REAL)2 * r);
      V = vadd(V, vmul(dir, rmax(wall - vdot(V, dir), (REAL)0)));
    }
  }
  x[0] = X.x, x[1] = X.y, x[2] = X.z;
  v[0] = V.x, v[1] = V.y, v[2] = V.z;
}

/* spray_update (fluid.wgsl:686-711): drops fly through still air (implicit drag), evaporate,
 * and land. Runs before g2p, so drops made in a step start moving in the next. */
static void o_spray_update(osim* s) {
  if (s->st.dt_ns == 0u || !s->spray) return;
  uint32_t n = umin(s->st.spray_count, s->spray_cap);
  REAL step = dt_(s), tnow = now_r(s);
  REAL g[3] = {(REAL)s->U.gravity[0], (REAL)s->U.gravity[1], (REAL)s->U.gravity[2]};
#pragma omp parallel for schedule(static)
  for (uint32_t i = 0; i < n; ++i) {
    REAL* rec = s->spray + (size_t)i * 12;
    REAL r = rec[3];
    if (r <= (REAL)0) continue;
    REAL sp = sqrt_r((rec[4] * rec[4] + rec[5] * rec[5]) + rec[6] * rec[6]);
    REAL re = (REAL)2 * r * sp * (REAL)RHO_AIR / (REAL)MU_AIR;
    REAL k = (REAL)3 * (REAL)RHO_AIR * drag_cd(re) * sp / ((REAL)8 * (REAL)s->U.fluid[0] * r);
    REAL den = (REAL)1 + k * step;
    REAL vn[3];
    for (int c = 0; c < 3; ++c) vn[c] = (rec[4 + c] + step * g[c]) / den;
    r = sqrt_r(rmax(r * r - (REAL)EVAPORATION * step, (REAL)0));
    if (r < (REAL)2.5e-6) r = 0;
    REAL pos[3];
    for (int c = 0; c < 3; ++c) pos[c] = rec[c] + step * vn[c];
    if (pos[1] <= r) {
      REAL* l = s->landed + (size_t)i * 4;
      l[0] = pos[0];
      l[1] = pos[2];
      l[2] = r * r_pow(rec[7], (REAL)1 / (REAL)3);
      l[3] = tnow;
      r = 0;
    }
    for (int c = 0; c < 3; ++c) {
      rec[c] = pos[c];
      rec[4 + c] = vn[c];
    }
    rec[3] = r;
  }
}

static REAL g2p_one(osim* s, const REAL* tg, uint32_t i, i3 corner, const uint32_t links[27]) {
  REAL* p = &s->P[(size_t)i * 20];
  REAL h = hh_(s);
  REAL X[3] = {(p[0] - (REAL)s->U.origin[0]) / h, (p[1] - (REAL)s->U.origin[1]) / h, (p[2] - (REAL)s->U.origin[2]) / h};
  static const REAL ou[3] = {0, 0.5, 0.5}, ov[3] = {0.5, 0, 0.5}, ow[3] = {0.5, 0.5, 0}, op[3] = {0.5, 0.5, 0.5};
  gathered gu = gather(s, tg, X, ou, corner, NATRAY_A_MOM(0), NATRAY_A_WT(0));
  gathered gv = gather(s, tg, X, ov, corner, NATRAY_A_MOM(1), NATRAY_A_WT(1));
  gathered gw = gather(s, tg, X, ow, corner, NATRAY_A_MOM(2), NATRAY_A_WT(2));
  gathered gp = gather(s, tg, X, op, corner, NATRAY_A_PC, NATRAY_A_WC);
  REAL v[3] = {gu.s, gv.s, gw.s};
  REAL m_near = gu.w + gv.w + gw.w;
  REAL step = dt_(s);
  REAL x[3] = {p[0] + step * v[0], p[1] + step * v[1], p[2] + step * v[2]};
  if (s->n_bullets) carve_cavity(s, x, v, i);
  if (s->sc.flags & NATRAY_SCENE_CAPSULE) carve(s, x, v);
  REAL spacing = (REAL)s->U.fluid[3];
  REAL lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = (REAL)s->U.origin[k] + (REAL)3.5 * h;
    hi[k] = (REAL)s->U.origin[k] + ((REAL)s->U.dims[k] - (REAL)4.5) * h;
  }
  int escaped = x[0] < lo[0] || x[2] < lo[2] || x[0] > hi[0] || x[2] > hi[2] || x[1] > hi[1];
  REAL fill = m_near / (REAL)24;
  int on_ground = x[1] < (REAL)1.5 * spacing;
  REAL r_thin = (REAL)0.75 * spacing;
  REAL t_thin = (REAL)3 * sqrt_r((REAL)s->U.fluid[0] * r_thin * r_thin * r_thin / (REAL)NATRAY_SIGMA_W);
  i3 cf_ = {(int)rfloor(X[0]), (int)rfloor(X[1]), (int)rfloor(X[2])};
  uint32_t tix = (uint32_t)(cf_.x - corner.x) * 49u + (uint32_t)(cf_.y - corner.y) * 7u + (uint32_t)(cf_.z - corner.z);
  uint32_t own = tile_cell(links, tix);
  int covered = own != NONE && accv(s, NATRAY_A_MA, own) > 0;
  if ((s->sc.flags & NATRAY_SCENE_MASK) && in_mask(s, cf_)) covered = 1;
  REAL thin_age = (fill < (REAL)0.45 && !on_ground && !covered) ? p[15] + step : (REAL)0;
  if (escaped || (!covered && fill < (REAL)0.2) || thin_age > t_thin) {
    if (s->spray) {
      /* Too thin for the grid: on the ground it soaks in, elsewhere it breaks up into spray. */
      if (on_ground)
        ground_touch(s, x[0], x[2]);
      else
        break_up(s, x, v, i);
    }
    p[4] = v[0];
    p[5] = v[1];
    p[6] = v[2];
    p[7] = 0;
    return 0;
  }
  REAL xx[3] = {x[0], x[1], x[2]}, vv[3] = {v[0], v[1], v[2]};
  if (xx[1] < (REAL)0.3 * spacing) {
    xx[1] = (REAL)0.3 * spacing;
    vv[1] = rmax(vv[1], (REAL)0);
  }
  if (s->spray && xx[1] < (REAL)0.7 * spacing) ground_touch(s, xx[0], xx[2]);
  REAL cs = (REAL)4 / (h * h);
  p[0] = xx[0];
  p[1] = xx[1];
  p[2] = xx[2];
  p[4] = vv[0];
  p[5] = vv[1];
  p[6] = vv[2];
  p[7] = 1;
  p[8] = gu.g[0] * cs;
  p[9] = gu.g[1] * cs;
  p[10] = gu.g[2] * cs;
  p[11] = gp.w > (REAL)1e-6 ? gp.s / gp.w : (REAL)0;
  p[12] = gv.g[0] * cs;
  p[13] = gv.g[1] * cs;
  p[14] = gv.g[2] * cs;
  p[15] = thin_age;
  p[16] = gw.g[0] * cs;
  p[17] = gw.g[1] * cs;
  p[18] = gw.g[2] * cs;
  p[19] = 0;
  return sqrt_r((vv[0] * vv[0] + vv[1] * vv[1]) + vv[2] * vv[2]);
}

static void o_g2p(osim* s) {
  if (s->st.dt_ns == 0u) return;
  uint32_t nb = n_blocks_(s);
  double fastest_all = 0.0;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : fastest_all)
  for (uint32_t b = 0; b < nb; ++b) {
    uint32_t count = s->block_start[b + 1] - s->block_start[b];
    if (count == 0u) continue;
    uint32_t links[27];
    REAL tg[8 * TILE];
    for (int li = 0; li < 27; ++li) links[li] = s->block_nb[b * 27u + (uint32_t)li];
    for (uint32_t t = 0; t < TILE; ++t) {
      uint32_t cell = tile_cell(links, t);
      int water = cell != NONE && cfv(s, NATRAY_F_TYPE, cell) == WATER;
      for (uint32_t d = 0; d < 3; ++d) {
        REAL u = 0, w = 0;
        if (cell != NONE) {
          u = cfv(s, NATRAY_F_U + d, cell);
          w = from_fp(accv(s, NATRAY_A_WT(d), cell), (REAL)NATRAY_FP_W);
        }
        tg[NATRAY_A_MOM(d) * TILE + t] = u;
        tg[NATRAY_A_WT(d) * TILE + t] = w;
      }
      tg[NATRAY_A_PC * TILE + t] = water ? pressure_of(s, cell) : (REAL)0;
      tg[NATRAY_A_WC * TILE + t] = water ? (REAL)1 : (REAL)0;
    }
    i3 bc = slot_coord(s, s->block_list[b]);
    i3 corner = {bc.x * 4 - 1, bc.y * 4 - 1, bc.z * 4 - 1};
    uint32_t first = s->block_start[b];
    REAL fastest = 0;
    for (uint32_t k = 0; k < count; ++k) fastest = rmax(fastest, g2p_one(s, tg, s->sorted[first + k], corner, links));
    if ((double)fastest > fastest_all) fastest_all = (double)fastest;
  }
  if (fastest_all > s->max_speed) s->max_speed = fastest_all;
  sync_state_floats(s);
}

/* ---- the plan (src/sim.rs:390-426, in-scope kernels only) ------------------------------------- */

int F(phase)(osim* s, int phase) {
  switch (phase) {
    case NATRAY_PHASE_BEGIN: o_begin_step(s); break;
    case NATRAY_PHASE_BIN: o_bin(s); break;
    case NATRAY_PHASE_P2G:
      o_p2g(s);
      o_air_carry(s);
      o_mem_splat(s);
      break;
    case NATRAY_PHASE_GRID:
      o_grid_prep(s);
      o_air_fill(s);
      o_bullets(s);
      break;
    case NATRAY_PHASE_SOLVE: o_solve(s); break;
    case NATRAY_PHASE_PROJECT: o_project(s); break;
    case NATRAY_PHASE_EXTEND:
      if (s->st.dt_ns != 0u) {
        o_extend_1(s);
        o_extend_2(s);
      }
      o_mem_sample(s);
      break;
    case NATRAY_PHASE_G2P:
      o_spray_update(s);
      o_g2p(s);
      break;
    default: return NATRAY_E_ARG;
  }
  return s->st.overflow ? NATRAY_E_OVERFLOW : NATRAY_OK;
}

int F(run)(osim* s, uint32_t n_steps, uint32_t iters, uint32_t stop_ns) {
  s->U.limits[2] = iters;
  s->U.limits[3] = stop_ns;
  s->st.iters_max = 0;
  s->st.unconverged = 0;
  for (uint32_t k = 0; k < n_steps; ++k)
    for (int ph = 0; ph < NATRAY_N_PHASES; ++ph) F(phase)(s, ph);
  return s->st.overflow ? NATRAY_E_OVERFLOW : NATRAY_OK;
}

/* ---- construction and access ---------------------------------------------------------------------- */

osim* F(create)(const natray_params* params, const float* particles, uint32_t n, const natray_scene* scene,
                const natray_state* init) {
  osim* s = (osim*)calloc(1, sizeof(osim));
  if (!s) return NULL;
  s->U = *params;
  if (scene) {
    s->sc = *scene;
  } else {
    memset(&s->sc, 0, sizeof s->sc);
    s->sc.flags = NATRAY_SCENE_REFERENCE;
  }
  s->mixed = IS_F64 || s->sc.precision == NATRAY_CG_MIXED || s->sc.precision == NATRAY_CG_PIPELINED;
  s->refresh = (s->sc.precision == NATRAY_CG_MIXED || s->sc.precision == NATRAY_CG_PIPELINED) && s->sc.refresh_every > 0u;
  s->N = n;
  s->S = params->bdims[0] * params->bdims[1] * params->bdims[2];
  s->cap = params->bdims[3];
  size_t cells = (size_t)s->cap * 64u;
  s->P = (REAL*)malloc((size_t)n * 20 * sizeof(REAL) + 1);
  s->block_ptr = (uint32_t*)malloc((size_t)s->S * 4);
  s->slot_count = (uint32_t*)calloc(s->S, 4);
  s->block_list = (uint32_t*)calloc((size_t)s->cap + 1, 4);
  s->block_nb = (uint32_t*)calloc((size_t)s->cap * 27 + 1, 4);
  s->block_start = (uint32_t*)calloc((size_t)s->cap + 2, 4);
  s->sorted = (uint32_t*)calloc((size_t)n + 1, 4);
  s->rank = (uint32_t*)calloc((size_t)n + 1, 4);
  s->acc = (int32_t*)calloc(NATRAY_ACC_FIELDS * cells + 1, 4);
  s->cf = (REAL*)calloc(NATRAY_CELL_FIELDS * cells + 1, sizeof(REAL));
  s->bvec = (REAL*)calloc(cells + 1, sizeof(REAL));
  s->part = (double*)calloc((size_t)s->cap * 4 + 4, sizeof(double));
  if (!s->P || !s->block_ptr || !s->slot_count || !s->block_list || !s->block_nb || !s->block_start || !s->sorted ||
      !s->rank || !s->acc || !s->cf || !s->bvec || !s->part) {
    F(destroy)(s);
    return NULL;
  }
  for (uint32_t i = 0; i < s->S; ++i) s->block_ptr[i] = NONE;
  if (s->sc.precision == NATRAY_CG_PIPELINED) {
    s->pz = (REAL*)calloc(cells + 1, sizeof(REAL));
    s->pq = (REAL*)calloc(cells + 1, sizeof(REAL));
    s->pm = (REAL*)calloc(cells + 1, sizeof(REAL));
    if (!s->pz || !s->pq || !s->pm) {
      F(destroy)(s);
      return NULL;
    }
  }
  if (s->sc.flags & NATRAY_SCENE_SPRAY) {
    /* src/sim.rs:252-254: spray and landed start zeroed, the wet map at u32::MAX. */
    s->spray_cap = params->limits[1];
    s->wet_res = params->limits[0];
    size_t ww = (size_t)s->wet_res * s->wet_res;
    s->spray = (REAL*)calloc((size_t)s->spray_cap * 12 + 1, sizeof(REAL));
    s->landed = (REAL*)calloc((size_t)s->spray_cap * 4 + 1, sizeof(REAL));
    s->touch = (uint32_t*)malloc((ww + 1) * 4);
    if (!s->spray || !s->landed || !s->touch) {
      F(destroy)(s);
      return NULL;
    }
    memset(s->touch, 0xff, (ww + 1) * 4);
  }
  if (s->sc.flags & NATRAY_SCENE_AIR) {
    s->air_x = (REAL*)calloc(5 * cells + 1, sizeof(REAL));
    s->air_tmp = (REAL*)calloc(5 * cells + 1, sizeof(REAL));
    s->prev_ptr = (uint32_t*)malloc(((size_t)s->S + 1) * 4);
    s->block_list_prev = (uint32_t*)calloc((size_t)s->cap + 1, 4);
    s->block_active = (uint32_t*)calloc((size_t)s->cap + 1, 4);
    if (!s->air_x || !s->air_tmp || !s->prev_ptr || !s->block_list_prev || !s->block_active) {
      F(destroy)(s);
      return NULL;
    }
    for (uint32_t i = 0; i < s->S; ++i) s->prev_ptr[i] = NONE;
  }
  for (size_t i = 0; i < (size_t)n * 20; ++i) s->P[i] = (REAL)particles[i];
  if (init) {
    F(write_state)(s, init);
  } else {
    /* src/sim.rs:240-241 */
    memset(&s->st, 0, sizeof s->st);
    float c_air = sqrtf(1.4f * 101325.0f / 1.204f);
    s->st.mem_dt = FLT_MAX;
    s->st.cap_dt = FLT_MAX;
    s->st.air_dt = 0.9f * params->origin[3] / (3.0f * c_air);
    s->max_speed = s->rz = s->alpha = s->beta = s->b_norm = s->r_norm = 0.0;
  }
  return s;
}

void F(destroy)(osim* s) {
  if (!s) return;
  free(s->P);
  free(s->block_ptr);
  free(s->slot_count);
  free(s->block_list);
  free(s->block_nb);
  free(s->block_start);
  free(s->sorted);
  free(s->rank);
  free(s->acc);
  free(s->cf);
  free(s->bvec);
  free(s->part);
  free(s->pz);
  free(s->pq);
  free(s->pm);
  free(s->guess);
  free(s->mem_x);
  free(s->mem_v);
  free(s->mem_mass);
  free(s->mem_push);
  free(s->mem_water);
  free(s->air_x);
  free(s->spray);
  free(s->landed);
  free(s->touch);
  free(s->patches);
  free(s->sections);
  free(s->pose_log);
  free(s->air_tmp);
  free(s->prev_ptr);
  free(s->block_list_prev);
  free(s->block_active);
  free(s);
}

void F(read_state)(const osim* s, natray_state* out) { *out = s->st; }
void F(write_state)(osim* s, const natray_state* in) {
  s->st = *in;
  s->max_speed = in->max_speed;
  s->rz = in->rz;
  s->alpha = in->alpha;
  s->beta = in->beta;
  s->b_norm = in->b_norm;
  s->r_norm = in->r_norm;
}
void F(read_particles)(const osim* s, float* dst) {
  for (size_t i = 0; i < (size_t)s->N * 20; ++i) dst[i] = (float)s->P[i];
}
void F(read_particles_f64)(const osim* s, double* dst) {
  for (size_t i = 0; i < (size_t)s->N * 20; ++i) dst[i] = (double)s->P[i];
}
void F(write_particles)(osim* s, const float* src) {
  for (size_t i = 0; i < (size_t)s->N * 20; ++i) s->P[i] = (REAL)src[i];
}
uint32_t F(n_blocks)(const osim* s) { return n_blocks_(s); }
void F(read_grid)(const osim* s, uint32_t* block_list, uint32_t* block_nb, uint32_t* block_start, int32_t* acc, float* cf) {
  uint32_t n = n_blocks_(s), nc = n * 64u;
  size_t cap = cap_cells(s);
  if (block_list) memcpy(block_list, s->block_list, (size_t)n * 4);
  if (block_nb) memcpy(block_nb, s->block_nb, (size_t)n * 27 * 4);
  if (block_start) memcpy(block_start, s->block_start, ((size_t)n + 1) * 4);
  if (acc)
    for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f) memcpy(&acc[(size_t)f * nc], &s->acc[f * cap], (size_t)nc * 4);
  if (cf)
    for (uint32_t f = 0; f < NATRAY_CELL_FIELDS; ++f)
      for (uint32_t c = 0; c < nc; ++c) cf[(size_t)f * nc + c] = (float)s->cf[f * cap + c];
}
void F(read_cf_f64)(const osim* s, double* cf) {
  uint32_t nc = n_blocks_(s) * 64u;
  size_t cap = cap_cells(s);
  for (uint32_t f = 0; f < NATRAY_CELL_FIELDS; ++f)
    for (uint32_t c = 0; c < nc; ++c) cf[(size_t)f * nc + c] = (double)s->cf[f * cap + c];
}
void F(write_grid)(osim* s, const int32_t* acc, const float* cf) {
  uint32_t nc = n_blocks_(s) * 64u;
  size_t cap = cap_cells(s);
  if (acc)
    for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f) memcpy(&s->acc[f * cap], &acc[(size_t)f * nc], (size_t)nc * 4);
  if (cf)
    for (uint32_t f = 0; f < NATRAY_CELL_FIELDS; ++f)
      for (uint32_t c = 0; c < nc; ++c) s->cf[f * cap + c] = (REAL)cf[(size_t)f * nc + c];
}
void F(read_bins)(const osim* s, uint32_t* block_ptr, uint32_t* sorted) {
  if (block_ptr) memcpy(block_ptr, s->block_ptr, (size_t)s->S * 4);
  if (sorted) memcpy(sorted, s->sorted, (size_t)s->N * 4);
}
void F(solver_scalars)(const osim* s, double out[8]) {
  out[0] = s->rz;
  out[1] = s->alpha;
  out[2] = s->beta;
  out[3] = s->b_norm;
  out[4] = s->r_norm;
  out[5] = (double)s->st.solve_iters;
  out[6] = s->max_speed;
  out[7] = (double)s->cg_iters;
}
/* The preconditioner of one block on its own, for checking it as an operator. */
void F(precondition)(const float* r, const float* diag, const int32_t* water, float* z, uint32_t mode) {
  osim tmp;
  memset(&tmp, 0, sizeof tmp);
  tmp.sc.precond = mode;
  REAL rr[64], dd[64], zz[64];
  int ww[64];
  for (int i = 0; i < 64; ++i) {
    rr[i] = (REAL)r[i];
    dd[i] = (REAL)diag[i];
    ww[i] = water[i] != 0;
  }
  precondition(&tmp, rr, dd, ww, zz);
  for (int i = 0; i < 64; ++i) z[i] = (float)zz[i];
}

void F(set_guess)(osim* s, const float* guess) {
  s->use_guess = 0;
  if (!guess) return;
  size_t nc = (size_t)cap_cells(s);
  if (!s->guess) s->guess = (REAL*)calloc(nc, sizeof(REAL));
  if (!s->guess) return;
  for (size_t i = 0; i < nc; ++i) s->guess[i] = (REAL)guess[i];
  s->use_guess = 1;
}

void F(residual)(osim* s, const float* p, float* r) {
  uint32_t nc = n_blocks_(s) * 64u;
  REAL* P = s->cf + (size_t)NATRAY_F_P * cap_cells(s);
  REAL* q = P;
  if (p) {
    q = (REAL*)malloc((size_t)cap_cells(s) * sizeof(REAL));
    if (!q) return;
    for (size_t i = 0; i < (size_t)cap_cells(s); ++i) q[i] = (REAL)p[i];
  }
  for (uint32_t cell = 0; cell < nc; ++cell) {
    REAL dg, b;
    r[cell] = cfv(s, NATRAY_F_TYPE, cell) == WATER ? (float)solve_row(s, cell, q, &dg, &b) : 0.0f;
  }
  if (q != P) free(q);
}

/* pack_pos (keys.wgsl:21-24) and pack_water (pack.wgsl:33-39): 21 bits per axis over the 8 m
 * box from (-4, -0.5, -4), the live flag in bit 31 of the second word. */
void F(pack_water)(const osim* s, uint32_t* out) {
  const REAL lo[3] = {(REAL)-4, (REAL)-0.5, (REAL)-4};
  for (uint32_t i = 0; i < s->N; ++i) {
    const REAL* p = &s->P[(size_t)i * 20];
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      REAL t = rmin(rmax((p[k] - lo[k]) / (REAL)8, (REAL)0), (REAL)1);
      q[k] = (uint32_t)(t * (REAL)2097151 + (REAL)0.5);
    }
    out[2 * i] = q[0] | (q[1] << 21);
    out[2 * i + 1] = (q[1] >> 11) | (q[2] << 10) | (p[7] > (REAL)0.5 ? 1u << 31 : 0u);
  }
}

int F(set_membrane)(osim* s, uint32_t n, const float* x, const float* v, const float* mass, const float* push) {
  if (n > s->mem_cap) {
    REAL* nx = (REAL*)calloc((size_t)n * 4, sizeof(REAL));
    REAL* nv = (REAL*)calloc((size_t)n * 4, sizeof(REAL));
    REAL* nm = (REAL*)calloc((size_t)n, sizeof(REAL));
    REAL* np = (REAL*)calloc((size_t)n * 2, sizeof(REAL));
    REAL* nw = (REAL*)calloc((size_t)n * 4, sizeof(REAL));
    if (!nx || !nv || !nm || !np || !nw) {
      free(nx);
      free(nv);
      free(nm);
      free(np);
      free(nw);
      return NATRAY_E_NOMEM;
    }
    free(s->mem_x);
    free(s->mem_v);
    free(s->mem_mass);
    free(s->mem_push);
    free(s->mem_water);
    s->mem_x = nx;
    s->mem_v = nv;
    s->mem_mass = nm;
    s->mem_push = np;
    s->mem_water = nw;
    s->mem_cap = n;
  }
  for (size_t k = 0; k < (size_t)n * 4; ++k) {
    s->mem_x[k] = (REAL)x[k];
    s->mem_v[k] = (REAL)v[k];
  }
  for (size_t k = 0; k < n; ++k) s->mem_mass[k] = (REAL)mass[k];
  for (size_t k = 0; k < (size_t)n * 2; ++k) s->mem_push[k] = (REAL)push[k];
  s->U.mem[0] = n;
  return NATRAY_OK;
}

void F(read_membrane)(const osim* s, float* water) {
  for (size_t k = 0; k < (size_t)mem_count(s) * 4; ++k) water[k] = (float)s->mem_water[k];
}

/* A whole keyframe (pack.wgsl:33-124, layout keys.wgsl:1-54): header (spray slots, live
 * slots, 0, 0), the water's and the membrane's 2 words each, the triangles' alive bits, the
 * spray's live bits, the live count before each bit word and 3 words per live slot. Returns the
 * size in words; writes `words` (if not null and max_words is enough) and spray_w[n][2]. */
static void pack_pos_(REAL px, REAL py, REAL pz, int flag, uint32_t out[2]) {
  const REAL lo[3] = {(REAL)-4, (REAL)-0.5, (REAL)-4}, p[3] = {px, py, pz};
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    REAL t = rmin(rmax((p[k] - lo[k]) / (REAL)8, (REAL)0), (REAL)1);
    q[k] = (uint32_t)(t * (REAL)2097151 + (REAL)0.5);
  }
  out[0] = q[0] | (q[1] << 21);
  out[1] = (q[1] >> 11) | (q[2] << 10) | (flag ? 1u << 31 : 0u);
}
uint32_t F(pack_keyframe)(const osim* s, uint32_t n_tris, const uint32_t* alive, uint32_t* words, uint32_t max_words,
                          float* spray_w) {
  const uint32_t n_verts = umin(s->U.mem[0], s->mem_cap);
  const uint32_t n = s->spray ? umin(s->st.spray_count, s->spray_cap) : 0u;
  const uint32_t nw = (n + 31u) / 32u;
  uint32_t live = 0;
  for (uint32_t i = 0; i < n; ++i) live += s->spray[(size_t)i * 12 + 3] > (REAL)0;
  const uint32_t total = 4u + 2u * (s->N + n_verts) + (n_tris + 31u) / 32u + 2u * nw + 3u * live;
  if (!words || max_words < total) return total;
  memset(words, 0, (size_t)total * 4);
  words[0] = n;
  words[1] = live;
  F(pack_water)(s, words + 4);
  const uint32_t mem_word = 4u + 2u * s->N, alive_word = mem_word + 2u * n_verts;
  for (uint32_t i = 0; i < n_verts; ++i) {
    const REAL* x = &s->mem_x[(size_t)i * 4];
    pack_pos_(x[0], x[1], x[2], 0, words + mem_word + 2u * i);
  }
  for (uint32_t t = 0; t < n_tris; ++t)
    if (alive[t]) words[alive_word + t / 32u] |= 1u << (t % 32u);
  const uint32_t spray_word = alive_word + (n_tris + 31u) / 32u;
  uint32_t before = 0, slot = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    uint32_t bits = 0;
    for (uint32_t k = 0; k < 32u && w * 32u + k < n; ++k)
      if (s->spray[(size_t)(w * 32u + k) * 12 + 3] > (REAL)0) bits |= 1u << k;
    words[spray_word + w] = bits;
    words[spray_word + nw + w] = before;
    before += (uint32_t)__builtin_popcount(bits);
  }
  for (uint32_t i = 0; i < n; ++i) {
    const REAL* r = &s->spray[(size_t)i * 12];
    if (spray_w) {
      spray_w[2 * i] = (float)r[7];
      spray_w[2 * i + 1] = (float)r[11];
    }
    if (!(r[3] > (REAL)0)) continue;
    uint32_t* e = words + spray_word + 2u * nw + 3u * slot++;
    pack_pos_(r[0], r[1], r[2], 0, e);
    float radius = (float)r[3];
    memcpy(&e[2], &radius, 4);
  }
  return total;
}

/* The spray's and the bullet's elementary functions on n floats (which: 0 exp, 1 log, 2 cos,
 * 3 pow(a, b), 4 sin, 5 asin), for the tests that pin them against the C library. */
void F(elementary)(int which, const float* a, const float* b, float* out, uint32_t n) {
  for (uint32_t k = 0; k < n; ++k)
    out[k] = which == 0 ? nm_expf(a[k]) : which == 1 ? nm_logf(a[k]) : which == 2 ? nm_cosf(a[k])
           : which == 3 ? nm_powf(a[k], b[k]) : which == 4 ? nm_sinf(a[k]) : nm_asinf(a[k]);
}

uint32_t F(read_spray)(const osim* s, uint32_t max_records, double* spray, double* landed) {
  if (!s->spray) return 0;
  uint32_t n = umin(umin(s->st.spray_count, s->spray_cap), max_records);
  for (size_t k = 0; k < (size_t)n * 12; ++k) spray[k] = (double)s->spray[k];
  for (size_t k = 0; k < (size_t)n * 4; ++k) landed[k] = (double)s->landed[k];
  return n;
}

int F(set_bullets)(osim* s, uint32_t n, const natray_bullet* bullets, const float* patches) {
  if (n > NATRAY_MAX_BULLETS) return NATRAY_E_ARG;
  if (n && !s->patches) {
    s->patches = (REAL*)calloc((size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_PATCHES * 8, sizeof(REAL));
    s->sections = (REAL*)calloc((size_t)NATRAY_MAX_BULLETS * B_SEC * 12, sizeof(REAL));
    s->pose_log = (REAL*)calloc((size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_LOG * 12, sizeof(REAL));
    if (!s->patches || !s->sections || !s->pose_log) return NATRAY_E_NOMEM;
  }
  for (uint32_t k = 0; k < n; ++k) {
    const natray_bullet* b = &bullets[k];
    struct obullet* o = &s->bullets[k];
    for (int c = 0; c < 4; ++c) {
      o->x[c] = (REAL)b->x[c], o->v[c] = (REAL)b->v[c], o->q[c] = (REAL)b->q[c], o->l[c] = (REAL)b->l[c];
      o->shape[c] = (REAL)b->shape[c], o->reach[c] = (REAL)b->reach[c], o->info[c] = b->info[c];
    }
  }
  if (n) {
    for (size_t i = 0; i < (size_t)n * NATRAY_BULLET_PATCHES * 8; ++i) s->patches[i] = (REAL)patches[i];
    memset(s->sections, 0, (size_t)NATRAY_MAX_BULLETS * B_SEC * 12 * sizeof(REAL));
    memset(s->pose_log, 0, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_LOG * 12 * sizeof(REAL));
  }
  if (!n && s->n_bullets)
    memset(&s->acc[(size_t)NATRAY_A_BU * cap_cells(s)], 0, (size_t)4 * cap_cells(s) * 4);
  s->n_bullets = n;
  s->U.mem[3] = n;
  return NATRAY_OK;
}

void F(read_bullets)(const osim* s, natray_bullet* bullets, double* sections) {
  for (uint32_t k = 0; k < s->n_bullets; ++k) {
    const struct obullet* o = &s->bullets[k];
    natray_bullet* b = &bullets[k];
    for (int c = 0; c < 4; ++c) {
      b->x[c] = (float)o->x[c], b->v[c] = (float)o->v[c], b->q[c] = (float)o->q[c], b->l[c] = (float)o->l[c];
      b->shape[c] = (float)o->shape[c], b->reach[c] = (float)o->reach[c], b->info[c] = o->info[c];
    }
  }
  if (sections)
    for (size_t i = 0; i < (size_t)s->n_bullets * B_SEC * 12; ++i) sections[i] = (double)s->sections[i];
}

void F(read_pose_log)(const osim* s, float* log) {
  for (size_t i = 0; i < (size_t)s->n_bullets * NATRAY_BULLET_LOG * 12; ++i) log[i] = (float)s->pose_log[i];
}

void F(read_wet)(const osim* s, uint32_t* touch) {
  if (!s->touch) return;
  memcpy(touch, s->touch, (size_t)s->wet_res * s->wet_res * 4);
}

void F(read_air)(const osim* s, float* out) {
  uint32_t nc = n_blocks_(s) * 64u, capc = cap_cells(s);
  if (!s->air_x) return;
  for (uint32_t f = 0; f < 5; ++f)
    for (uint32_t c = 0; c < nc; ++c) out[(size_t)f * nc + c] = (float)s->air_x[(size_t)f * capc + c];
}

void F(set_threads)(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
