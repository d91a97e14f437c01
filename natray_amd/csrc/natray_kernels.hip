// HIP kernels of the natray water substep for MI355X (gfx950, CDNA4, wave64).
//
// One step runs, in the reference's plan order (src/sim.rs:390-426):
//   k_begin                       begin_step                 shaders/fluid.wgsl:31-89
//   k_clear k_mark k_scan_* k_sort clear/mark/scan/sort      shaders/fluid.wgsl:91-196
//   k_p2g                         p2g                        shaders/fluid.wgsl:198-303
//   k_grid_prep                   grid_prep                  shaders/fluid.wgsl:305-356
//   k_solve (cooperative)         solve_init..solve_beta     shaders/fluid.wgsl:381-578
//   k_project                     project                    shaders/fluid.wgsl:580-598
//   k_extend1 k_extend2           extend_1, extend_2         shaders/fluid.wgsl:600-674
//   k_g2p                         g2p                        shaders/fluid.wgsl:767-901
//
// Mapping to the hardware:
//   - per-cell kernels run one 64-lane wavefront per 4^3 block, lane = cell (l = 16x+4y+z);
//     in-block reductions and the block-SGS preconditioner stay inside the wave (DPP/bpermute
//     shuffles, no workgroup barriers);
//   - P2G and G2P run one 256-thread workgroup per block with the 7^3 tile of 8 fields in LDS;
//     P2G accumulates there with LDS integer atomics and flushes each non-zero tile sum with one
//     global integer atomic, so grid sums are exact and order-independent;
//   - block allocation is a deterministic compaction of claimed domain slots in slot order (a
//     three-kernel scan), replacing the reference's CAS claims; particle ranks use wave-
//     aggregated atomics;
//   - the whole pressure solve is one persistent cooperative launch with a grid barrier, so the
//     GPU decides when to stop iterating without host round trips.
// Every float expression is written in the reference's order and compiled without
// contraction, so the kernels reproduce the CPU oracle bit for bit given the same inputs.
#include <hip/hip_runtime.h>
#include <float.h>
#include <mutex>
#include <type_traits>

#include "natray_dev.h"
#include "natray_lanes.h"
#include "natray_math.h"

namespace natray {

#define NONE NATRAY_NONE
#define PENDING NATRAY_PENDING
constexpr uint32_t TILE = 343;
constexpr float GAS = NATRAY_GAS, WATER = NATRAY_WATER, SOLID = NATRAY_SOLID, COVERED = NATRAY_COVERED,
                RUBBER = NATRAY_RUBBER;

// ---- launch policies ---------------------------------------------------------------------------

// A kernel's scene and its workgroup among the scene's: with One (a single scene, Dev by value)
// the launch's own; with Many (a batch of scenes in one launch) from the workgroup map.
struct View {
  const Dev& d;
  uint32_t wg, wgs;
};
__device__ __forceinline__ View view(const One& l) { return {l.d, blockIdx.x, gridDim.x}; }
__device__ __forceinline__ View view(const Many& l) {
  const WgMap m = l.map[blockIdx.x];
  return {l.scenes[m.scene], m.wg, m.wgs};
}

// ---- small helpers ---------------------------------------------------------------------------

__device__ __forceinline__ float fmin_(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float fmax_(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ uint32_t n_blocks(const Dev& d) {
  uint32_t n = d.st->n_blocks, c = d.U->bdims[3];
  return n < c ? n : c;
}
__device__ __forceinline__ float dt_of(const Dev& d) { return (float)d.st->dt_ns * 1e-9f; }
__device__ __forceinline__ float now_of(const Dev& d) { return (float)d.st->t_ns * 1e-9f; }

__device__ __forceinline__ uint32_t block_slot(const natray_params* U, int bx, int by, int bz) {
  return ((uint32_t)bx * U->bdims[1] + (uint32_t)by) * U->bdims[2] + (uint32_t)bz;
}
__device__ __forceinline__ void slot_coord(const natray_params* U, uint32_t s, int& x, int& y, int& z) {
  x = (int)(s / (U->bdims[2] * U->bdims[1]));
  y = (int)((s / U->bdims[2]) % U->bdims[1]);
  z = (int)(s % U->bdims[2]);
}
__device__ __forceinline__ void cell_coord(const Dev& d, uint32_t cell, int& x, int& y, int& z) {
  slot_coord(d.U, d.block_list[cell >> 6], x, y, z);
  uint32_t l = cell & 63u;
  x = x * 4 + (int)(l >> 4);
  y = y * 4 + (int)((l >> 2) & 3u);
  z = z * 4 + (int)(l & 3u);
}

// to_fp / from_fp (shaders/sim_defs.wgsl:217-222): ties-to-even rounding of the clamped product.
// The clamp is one v_med3_f32, which equals min(max(v, -2e9), 2e9) for every v but NaN.
__device__ __forceinline__ int32_t to_fp(float x, float scale) {
  return (int32_t)__builtin_rintf(__builtin_amdgcn_fmed3f(x * scale, -2e9f, 2e9f));
}
__device__ __forceinline__ float from_fp(int32_t x, float scale) { return (float)x / scale; }

__device__ __forceinline__ float& cfr(const Dev& d, uint32_t f, uint32_t cell) { return d.cf[(size_t)f * d.capc + cell]; }
__device__ __forceinline__ int32_t& accr(const Dev& d, uint32_t f, uint32_t cell) {
  return d.acc[(size_t)f * d.capc + cell];
}
__device__ __forceinline__ float cell_type(const Dev& d, uint32_t cell) {
  return cell == NONE ? GAS : cfr(d, NATRAY_F_TYPE, cell);
}

// The cell `o` away from local cell (lx, ly, lz) of the block whose 27 neighbours are `lk`
// (shaders/sim_defs.wgsl:195-202); offsets up to 4 in any axis stay inside the 27 blocks.
__device__ __forceinline__ uint32_t nbr(const uint32_t* lk, int lx, int ly, int lz, int ox, int oy, int oz) {
  int qx = lx + ox, qy = ly + oy, qz = lz + oz;
  int nx = (qx >= 4) - (qx < 0), ny = (qy >= 4) - (qy < 0), nz = (qz >= 4) - (qz < 0);
  uint32_t b = lk[(nx + 1) * 9 + (ny + 1) * 3 + nz + 1];
  if (b == NONE) return NONE;
  return b * 64u + (uint32_t)(qx - nx * 4) * 16u + (uint32_t)(qy - ny * 4) * 4u + (uint32_t)(qz - nz * 4);
}

// ---- scene geometry (the reference's ground and knot, plus the synthetic scenes' solids) ------

__device__ __forceinline__ void cell_centre(const Dev& d, int cx, int cy, int cz, float& px, float& py, float& pz) {
  float h = d.U->origin[3];
  px = d.U->origin[0] + ((float)cx + 0.5f) * h;
  py = d.U->origin[1] + ((float)cy + 0.5f) * h;
  pz = d.U->origin[2] + ((float)cz + 0.5f) * h;
}
__device__ __forceinline__ float capsule_d2(const Dev& d, float px, float py, float pz, float t, float* on, float* ab) {
  const natray_scene* sc = d.sc;
  float ax = sc->capsule_a[0] + sc->capsule_v[0] * t, ay = sc->capsule_a[1] + sc->capsule_v[1] * t,
        az = sc->capsule_a[2] + sc->capsule_v[2] * t;
  float bx = sc->capsule_b[0] - sc->capsule_a[0], by = sc->capsule_b[1] - sc->capsule_a[1],
        bz = sc->capsule_b[2] - sc->capsule_a[2];
  float ll = (bx * bx + by * by) + bz * bz;
  float f = (((px - ax) * bx + (py - ay) * by) + (pz - az) * bz) / fmax_(ll, 1e-20f);
  f = fmin_(fmax_(f, 0.0f), 1.0f);
  float ox = ax + f * bx, oy = ay + f * by, oz = az + f * bz;
  if (on) {
    on[0] = ox; on[1] = oy; on[2] = oz;
    ab[0] = bx; ab[1] = by; ab[2] = bz;
  }
  float dx = px - ox, dy = py - oy, dz = pz - oz;
  return (dx * dx + dy * dy) + dz * dz;
}
__device__ __forceinline__ bool in_capsule(const Dev& d, int cx, int cy, int cz) {
  float px, py, pz;
  cell_centre(d, cx, cy, cz, px, py, pz);
  float r = d.sc->capsule_a[3];
  return capsule_d2(d, px, py, pz, now_of(d), nullptr, nullptr) < r * r;
}
__device__ __forceinline__ bool in_mask(const Dev& d, int cx, int cy, int cz) {
  float px, py, pz;
  cell_centre(d, cx, cy, cz, px, py, pz);
  float dx = px - d.sc->mask[0], dy = py - d.sc->mask[1], dz = pz - d.sc->mask[2], r = d.sc->mask[3];
  return (dx * dx + dy * dy) + dz * dz < r * r;
}
// is_ground / in_knot (shaders/fluid.wgsl:305-314) and the scene's walls and capsule.
__device__ __forceinline__ bool is_solid_cell(const Dev& d, int cx, int cy, int cz) {
  const uint32_t fl = d.sc->flags;
  const natray_params* U = d.U;
  float h = U->origin[3];
  if ((fl & NATRAY_SCENE_GROUND) && U->origin[1] + ((float)cy + 0.5f) * h < 0.0f) return true;
  if (fl & NATRAY_SCENE_KNOT) {
    float px, py, pz;
    cell_centre(d, cx, cy, cz, px, py, pz);
    px -= U->knot[0];
    py -= U->knot[1];
    pz -= U->knot[2];
    float r = U->knot[3] + 0.5f * h;
    if (py > 0.0f && py < 2.0f * r && __builtin_sqrtf(px * px + pz * pz) < r) return true;
  }
  if ((fl & NATRAY_SCENE_WALLS) && (cx == 0 || cy == 0 || cz == 0 || cx == (int)U->dims[0] - 1 ||
                                    cy == (int)U->dims[1] - 1 || cz == (int)U->dims[2] - 1))
    return true;
  if ((fl & NATRAY_SCENE_CAPSULE) && in_capsule(d, cx, cy, cz)) return true;
  return false;
}
__device__ __forceinline__ float solid_u(const Dev& d, uint32_t cell, uint32_t dim) {
  if (!(d.sc->flags & NATRAY_SCENE_CAPSULE) || cell == NONE) return 0.0f;
  int x, y, z;
  cell_coord(d, cell, x, y, z);
  return in_capsule(d, x, y, z) ? d.sc->capsule_v[dim] : 0.0f;
}
// bounded (shaders/fluid.wgsl:360-365); a moving solid's velocity takes the place of 0.
__device__ __forceinline__ float bounded(const Dev& d, float u, float hi, float lo, uint32_t hi_cell, uint32_t lo_cell,
                                         uint32_t dim) {
  if (lo == SOLID && hi == SOLID) return 0.0f;
  if (lo == SOLID) return fmax_(u, solid_u(d, lo_cell, dim));
  if (hi == SOLID) return fmin_(u, solid_u(d, hi_cell, dim));
  return u;
}
// face_u (shaders/fluid.wgsl:368-371)
__device__ __forceinline__ float face_u(const Dev& d, uint32_t cell, uint32_t lo, uint32_t dim) {
  if (cell == NONE) return 0.0f;
  return bounded(d, cfr(d, NATRAY_F_U + dim, cell), cfr(d, NATRAY_F_TYPE, cell), cell_type(d, lo), cell, lo, dim);
}
// pressure_of (shaders/fluid.wgsl:375-379)
__device__ __forceinline__ float pressure_of(const Dev& d, uint32_t cell) {
  if (cell == NONE) return 0.0f;
  float p = cfr(d, NATRAY_F_P, cell);
  return cfr(d, NATRAY_F_TYPE, cell) == WATER ? fmax_(p, NATRAY_P_FLOOR) : p;
}

__device__ __forceinline__ uint32_t pack_key(const Dev& d, float px, float py, float pz, bool alive) {
  if (!alive) return NONE;
  const natray_params* U = d.U;
  float h = U->origin[3];
  int cx = (int)floorf((px - U->origin[0]) / h), cy = (int)floorf((py - U->origin[1]) / h),
      cz = (int)floorf((pz - U->origin[2]) / h);
  if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)U->dims[0] || cy >= (int)U->dims[1] || cz >= (int)U->dims[2]) return NONE;
  return (uint32_t)cx | ((uint32_t)cy << KEY_BITS) | ((uint32_t)cz << (2 * KEY_BITS));
}

// ---- handing out blocks ---------------------------------------------------------------------

// P2G and G2P hand out blocks to their workgroups as each finishes one, since blocks cost what
// their particles cost. The dealer has eight counters per kernel, Dev::bar[base + 32 x] for
// x = 0..7, 128 bytes apart, and counter x deals range x, the blocks [x n / 8, (x + 1) n / 8).
// Workgroup g (of G) starts on range x = g mod 8: its first block is that range's (g / 8)-th,
// with no atomic, and its later ones are the range's (G_x + i)-th for draws i = 0, 1, ... of
// counter x, where G_x workgroups start on range x. Once a range is dealt, wave 0 of the
// workgroup checks the other counters at once and draws from the next open range. Each draw
// happens one block ahead: lane 0 issues the atomic for the block after next, and wave 0
// reads its result during the next block, past that block's first barrier, so neither the
// atomic's latency nor a barrier of its own falls between two blocks. One counter for every
// draw serialised them on one address: at 5,700 blocks that took most of the kernels' time.
// k_begin resets the counters.
constexpr uint32_t DEAL_P2G = 256, DEAL_G2P = 512;

struct Dealer {
  uint32_t x;      // the range drawn from (the same in every lane of wave 0)
  uint32_t spent;  // bit x: range x is dealt (the same in every lane of wave 0)
  uint32_t ahead;  // lane 0: the result of the last draw, resolved during the next block
};

__device__ __forceinline__ uint32_t deal_lo(uint32_t n, uint32_t x) { return (uint32_t)((uint64_t)n * x / 8u); }
__device__ __forceinline__ uint32_t deal_size(uint32_t n, uint32_t x) { return deal_lo(n, x + 1u) - deal_lo(n, x); }
// Workgroups (of wgs) that start on range x.
__device__ __forceinline__ uint32_t deal_starts(uint32_t wgs, uint32_t x) { return wgs > x ? (wgs - x + 7u) / 8u : 0u; }

// The block after draw `j` of range s.x (wave 0, every lane): that range's block, else the
// next open range's, whose counters wave 0 checks at once; NONE once every block is dealt.
__device__ __forceinline__ uint32_t deal_resolve(const Dev& d, uint32_t base, uint32_t n, uint32_t wgs, Dealer& s,
                                                 uint32_t j) {
  const uint32_t lane = threadIdx.x;
  uint32_t x = s.x, spent = s.spent, b = NONE;
  for (;;) {
    if (j < deal_size(n, x)) {
      b = deal_lo(n, x) + j;
      break;
    }
    spent |= 1u << x;
    bool open = false;
    if (lane < 8u && !(spent >> lane & 1u))
      open = __hip_atomic_load(d.bar + base + 32u * lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) +
                 deal_starts(wgs, lane) <
             deal_size(n, lane);
    const uint32_t m = (uint32_t)__ballot(open) & 0xffu;
    spent |= ~m & 0xffu;
    if (m == 0u) break;
    const uint32_t sh = (x + 1u) & 7u;  // bit i of rot: range (x + 1 + i) mod 8
    const uint32_t rot = ((m >> sh) | (m << (8u - sh))) & 0xffu;
    x = (x + 1u + (uint32_t)__builtin_ctz(rot)) & 7u;
    uint32_t i = 0;
    if (lane == 0u) i = atomicAdd(d.bar + base + 32u * x, 1u);
    j = __builtin_amdgcn_readfirstlane(i) + deal_starts(wgs, x);
  }
  s.x = x;
  s.spent = spent;
  return b;
}

// The workgroup's first block: range wg mod 8's (wg / 8)-th, with no atomic. Call from all
// threads with the workgroup's index wg among wgs; returns the same block to all.
__device__ __forceinline__ uint32_t deal_first(const Dev& d, uint32_t base, uint32_t n, uint32_t wg, uint32_t wgs,
                                               Dealer& s, uint32_t* s_next) {
  if (threadIdx.x < 64u) {
    s.x = wg & 7u;
    s.spent = 0u;
    const uint32_t b = deal_resolve(d, base, n, wgs, s, wg >> 3);
    if (threadIdx.x == 0u) {
      if (b != NONE) s.ahead = atomicAdd(d.bar + base + 32u * s.x, 1u);
      *s_next = NONE;
      s_next[1] = b;
    }
  }
  __syncthreads();
  return s_next[1];
}

// During a block, after its first barrier (all threads call it): wave 0 resolves the draw
// made one block ago into the block after this one, and lane 0 draws again. The result waits
// in s_next[0] for deal_next.
__device__ __forceinline__ void deal_ahead(const Dev& d, uint32_t base, uint32_t n, uint32_t wgs, Dealer& s,
                                           uint32_t* s_next) {
  if (threadIdx.x < 64u) {
    const uint32_t b = deal_resolve(d, base, n, wgs, s, __builtin_amdgcn_readfirstlane(s.ahead) + deal_starts(wgs, s.x));
    if (threadIdx.x == 0u) {
      if (b != NONE) s.ahead = atomicAdd(d.bar + base + 32u * s.x, 1u);
      s_next[0] = b;
    }
  }
}

// The next block, once every thread is done with the current one.
__device__ __forceinline__ uint32_t deal_next(const uint32_t* s_next) {
  __syncthreads();
  return s_next[0];
}

// ---- begin_step (shaders/fluid.wgsl:31-89) -----------------------------------------------------

template <class L>
__global__ void k_begin(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (threadIdx.x != 0) return;
  natray_state* st = d.st;
  const natray_params* U = d.U;
  const natray_scene* sc = d.sc;
  float vmax = st->max_speed;
  uint32_t arrive = U->limits[3];
  if (sc->flags & NATRAY_SCENE_CAPSULE) {
    float vx = sc->capsule_v[0], vy = sc->capsule_v[1], vz = sc->capsule_v[2];
    vmax = fmax_(vmax, __builtin_sqrtf((vx * vx + vy * vy) + vz * vz));
  }
  // A bullet's own pace counts from the step it joins in, which starts exactly when it arrives.
  for (uint32_t k = 0; k < d.n_bullets; ++k) {
    const natray_bullet& b = d.bullets[k];
    if (b.info[2] == 1u) {
      vmax = fmax_(vmax, b.l[3]);
    } else if (b.info[2] == 0u) {
      if (b.info[0] > st->t_ns)
        arrive = arrive < b.info[0] ? arrive : b.info[0];
      else
        vmax = fmax_(vmax, __builtin_sqrtf((b.v[0] * b.v[0] + b.v[1] * b.v[1]) + b.v[2] * b.v[2]));
    }
  }
  float mem_dt = st->mem_dt, air_dt = st->air_dt;
  float dt_s = fmin_(U->fluid[2], st->cap_dt);
  if (vmax > 0.0f) dt_s = fmin_(dt_s, U->fluid[1] * U->origin[3] / vmax);
  st->dt_flow = dt_s;
  st->mem_dt_seen = mem_dt;
  st->air_dt_seen = air_dt;
  dt_s = fmin_(dt_s, fmin_((float)U->mem[1] * mem_dt, (float)U->mem[2] * air_dt));
  if (st->free_dt_ns > 0u) dt_s = fmin_(dt_s, 1.25f * (float)st->free_dt_ns * 1e-9f);
  uint32_t free_ns = (uint32_t)(dt_s * 1e9f);
  free_ns = free_ns < 20u ? 20u : free_ns;
  if (d.fixed_ns) free_ns = d.fixed_ns;
  uint32_t stop = U->limits[3];
  bool live = st->t_ns < stop;
  uint32_t dt_ns = 0;
  if (live) {
    uint32_t lim = (stop < arrive ? stop : arrive) - st->t_ns;
    dt_ns = free_ns < lim ? free_ns : lim;
    st->free_dt_ns = free_ns;
    st->t_ns += dt_ns;
    st->step += 1u;
    st->max_speed = 0.0f;
    st->mem_dt = FLT_MAX;
    st->cap_dt = FLT_MAX;
    st->air_dt = FLT_MAX;
    st->mem_ticket = 0;
    st->n_blocks_prev = n_blocks(d);
    st->n_blocks = 0;
  }
  st->dt_ns = dt_ns;
  for (uint32_t x = 0; x < 8u; ++x) {
    d.bar[DEAL_P2G + 32u * x] = 0u;
    d.bar[DEAL_G2P + 32u * x] = 0u;
  }
  d.bar[10] = 0u;
  d.bar[11] = 0u;
  d.bar[12] = 0u;
}


// Keys of particles written from the host.
template <class L>
__global__ void k_keys(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  uint32_t i = v.wg * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  const float4* p = reinterpret_cast<const float4*>(d.P) + (size_t)i * 5;
  float4 pos = p[0], vel = p[1];
  d.key[i] = pack_key(d, pos.x, pos.y, pos.z, !(vel.w < 0.5f));
}

// ---- binning (shaders/fluid.wgsl:91-196) ---------------------------------------------------------

// clear_blocks (fluid.wgsl:91-107): zero the transfer sums of last step's cells. The slots'
// block pointers and counts are reset by the scan itself.
template <class L>
__global__ void k_clear(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  uint32_t n4 = d.st->n_blocks_prev * 16u;  // int4 per field
  // With the air boundary, the last step's block list and each slot's place in it (read by
  // air_carry and keep_blocks; this step's list is built after this kernel).
  if (d.air_x)
    for (uint32_t b = v.wg * blockDim.x + threadIdx.x; b < d.st->n_blocks_prev; b += v.wgs * blockDim.x) {
      const uint32_t s = d.block_list[b];
      d.block_list_prev[b] = s;
      d.prev_ptr[s] = b;
    }
  // The transfer fields, with a membrane its coupling fields (A_MF, A_MA, A_MW, A_MV) and with
  // bullets theirs (A_BU, A_SRC).
  for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f) {
    const bool mem_f = f == NATRAY_A_MF || f == NATRAY_A_MA || f >= NATRAY_A_MW;
    const bool bullet_f = f >= NATRAY_A_BU && f <= NATRAY_A_SRC;
    if ((mem_f && !d.mem_cap) || (bullet_f && !d.n_bullets)) continue;
    int4* a = reinterpret_cast<int4*>(d.acc + (size_t)f * d.capc);
    for (uint32_t i = v.wg * blockDim.x + threadIdx.x; i < n4; i += v.wgs * blockDim.x) a[i] = make_int4(0, 0, 0, 0);
  }
}

// mark_blocks + claim (fluid.wgsl:109-150): mark the blocks of cells c-1..c+2 as claimed and
// take a rank in the own block. Each thread takes PER particles, 256 apart, so that their loads
// and atomics are in flight together: the kernel is a chain of dependent round trips (key,
// block slots, the claim), and fewer, fuller waves wait through fewer of them. Lanes of a wave
// with the same own block share one atomic; every group's atomic is issued before any result
// is used.
template <class L, uint32_t PER>
__global__ __launch_bounds__(256) void k_mark(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const natray_params* U = d.U;
  const uint32_t lane = lane_id();
  const uint32_t i0 = v.wg * 256u * PER + threadIdx.x;
  uint32_t key[PER], own[PER];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t i = i0 + 256u * k;
    key[k] = i < d.N ? d.key[i] : NONE;
  }
  const int bmx = (int)U->bdims[0] - 1, bmy = (int)U->bdims[1] - 1, bmz = (int)U->bdims[2] - 1;
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    own[k] = 0;
    if (key[k] == NONE) continue;
    const int cx = (int)(key[k] & KEY_MASK), cy = (int)((key[k] >> KEY_BITS) & KEY_MASK), cz = (int)(key[k] >> (2 * KEY_BITS));
    const int lx = max(0, min((cx - 1) >> 2, bmx)), ly = max(0, min((cy - 1) >> 2, bmy)), lz = max(0, min((cz - 1) >> 2, bmz));
    const int hx = max(0, min((cx + 2) >> 2, bmx)), hy = max(0, min((cy + 2) >> 2, bmy)), hz = max(0, min((cz + 2) >> 2, bmz));
    // Cells c-1..c+2 span one or two blocks per axis. All of their slots are read before any
    // is written, so the reads go out together instead of one round trip per slot.
    uint32_t slot[8], cur[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int x = lx + (q >> 2), y = ly + ((q >> 1) & 1), z = lz + (q & 1);
      slot[q] = (x <= hx && y <= hy && z <= hz) ? block_slot(U, x, y, z) : NONE;
      cur[q] = slot[q] != NONE ? d.block_ptr[slot[q]] : PENDING;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (cur[q] != PENDING) d.block_ptr[slot[q]] = PENDING;
    own[k] = block_slot(U, cx >> 2, cy >> 2, cz >> 2);
  }
  // Groups of lanes with the same own block: each lane's leader and its place in the group.
  uint32_t leader[PER], place[PER], base[PER];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const bool active = key[k] != NONE;
    unsigned long long todo = __ballot(active);
    leader[k] = 0;
    place[k] = 0;
    base[k] = 0;
    uint32_t size = 0;
    while (todo) {
      const int first = __ffsll((long long)todo) - 1;
      const uint32_t lslot = __shfl(own[k], first);
      const unsigned long long peers = __ballot(active && own[k] == lslot);
      if ((peers >> lane) & 1ull) {
        leader[k] = (uint32_t)first;
        place[k] = (uint32_t)__popcll(peers & ((1ull << lane) - 1ull));
      }
      if ((int)lane == first) size = (uint32_t)__popcll(peers);
      todo &= ~peers;
    }
    if (size) base[k] = atomicAdd(&d.slot_count[own[k]], size);
  }
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    const uint32_t got = __shfl(base[k], (int)leader[k]);
    if (key[k] != NONE) d.rank[i0 + 256u * k] = got + place[k];
  }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// scan_blocks, first pass: claimed slots and their particles per tile of SCAN_TILE slots.
template <class L>
__global__ __launch_bounds__(256) void k_scan_reduce(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  __shared__ uint32_t ws[2][4];
  const uint32_t base = v.wg * SCAN_TILE + threadIdx.x * 4u;
  uint4 bp = *reinterpret_cast<const uint4*>(d.block_ptr + base);
  uint4 sc = *reinterpret_cast<const uint4*>(d.slot_count + base);
  uint32_t f = (bp.x == PENDING) + (bp.y == PENDING) + (bp.z == PENDING) + (bp.w == PENDING);
  uint32_t c = sc.x + sc.y + sc.z + sc.w;
  f = wave_sum(f);
  c = wave_sum(c);
  if (lane_id() == 0) {
    ws[0][threadIdx.x >> 6] = f;
    ws[1][threadIdx.x >> 6] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    d.tile_sum[v.wg] = ws[0][0] + ws[0][1] + ws[0][2] + ws[0][3];
    d.tile_sum[d.ntiles + v.wg] = ws[1][0] + ws[1][1] + ws[1][2] + ws[1][3];
  }
}

// Inclusive scan of two counters across a 64-lane wave.
__device__ __forceinline__ void wave_scan2(uint32_t& a, uint32_t& b) {
  const uint32_t lane = lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t ta = __shfl_up(a, off), tb = __shfl_up(b, off);
    if (lane >= (uint32_t)off) {
      a += ta;
      b += tb;
    }
  }
}

// scan_blocks, second pass: exclusive prefix of the tile totals (one workgroup), the block
// count and the overflow flag (claim, fluid.wgsl:117-121).
template <class L>
__global__ __launch_bounds__(1024) void k_scan_tiles(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  __shared__ uint32_t wsum[2][16];
  const uint32_t nt = d.ntiles, t = threadIdx.x;
  const uint32_t per = (nt + 1023u) / 1024u;
  const uint32_t lo = min(t * per, nt), hi = min(lo + per, nt);
  uint32_t fa = 0, ca = 0;
  for (uint32_t j = lo; j < hi; ++j) {
    fa += d.tile_sum[j];
    ca += d.tile_sum[nt + j];
  }
  uint32_t fi = fa, ci = ca;
  wave_scan2(fi, ci);
  const uint32_t w = t >> 6;
  if (lane_id() == 63) {
    wsum[0][w] = fi;
    wsum[1][w] = ci;
  }
  __syncthreads();
  uint32_t fo = 0, co = 0;
  for (uint32_t k = 0; k < w; ++k) {
    fo += wsum[0][k];
    co += wsum[1][k];
  }
  uint32_t frun = fo + fi - fa, crun = co + ci - ca;
  for (uint32_t j = lo; j < hi; ++j) {
    d.tile_off[j] = frun;
    d.tile_off[nt + j] = crun;
    frun += d.tile_sum[j];
    crun += d.tile_sum[nt + j];
  }
  if (t == 1023) {
    uint32_t total_f = fo + fi, total_c = co + ci;
    uint32_t cap = d.U->bdims[3];
    d.st->n_blocks = total_f;
    if (total_f > cap) {
      d.st->overflow = 1u;
    } else {
      d.block_start[total_f] = total_c;
    }
  }
}

// scan_blocks, third pass: block indices in slot order, block starts, and every slot's pointer
// (NONE where unclaimed); resets the counts for the next step.
template <class L>
__global__ __launch_bounds__(256) void k_scan_scatter(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  __shared__ uint32_t ws[2][4];
  const uint32_t j = v.wg, base = j * SCAN_TILE + threadIdx.x * 4u;
  uint4 bp = *reinterpret_cast<const uint4*>(d.block_ptr + base);
  uint4 sc = *reinterpret_cast<const uint4*>(d.slot_count + base);
  uint32_t fl[4] = {bp.x == PENDING, bp.y == PENDING, bp.z == PENDING, bp.w == PENDING};
  uint32_t cn[4] = {sc.x, sc.y, sc.z, sc.w};
  uint32_t fa = fl[0] + fl[1] + fl[2] + fl[3], ca = cn[0] + cn[1] + cn[2] + cn[3];
  uint32_t fi = fa, ci = ca;
  wave_scan2(fi, ci);
  const uint32_t w = threadIdx.x >> 6;
  if (lane_id() == 63) {
    ws[0][w] = fi;
    ws[1][w] = ci;
  }
  __syncthreads();
  uint32_t fo = d.tile_off[j], co = d.tile_off[d.ntiles + j];
  for (uint32_t k = 0; k < w; ++k) {
    fo += ws[0][k];
    co += ws[1][k];
  }
  fo += fi - fa;
  co += ci - ca;
  const uint32_t cap = d.U->bdims[3];
  uint32_t out[4];
  for (int q = 0; q < 4; ++q) {
    uint32_t s = base + (uint32_t)q;
    out[q] = NONE;
    if (fl[q]) {
      if (fo < cap) {
        d.block_list[fo] = s;
        d.block_start[fo] = co;
        out[q] = fo;
      } else if (fo == cap) {
        d.block_start[cap] = co;
      }
      ++fo;
    }
    co += cn[q];
  }
  *reinterpret_cast<uint4*>(d.block_ptr + base) = make_uint4(out[0], out[1], out[2], out[3]);
  *reinterpret_cast<uint4*>(d.slot_count + base) = make_uint4(0, 0, 0, 0);
}

// sort_particles (fluid.wgsl:189-196)
template <class L, uint32_t PER>
__global__ __launch_bounds__(256) void k_sort(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  // PER particles per thread, 256 apart, each step's loads in flight together (as k_mark).
  const uint32_t i0 = v.wg * 256u * PER + threadIdx.x;
  uint32_t key[PER], b[PER];
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) key[k] = i0 + 256u * k < d.N ? d.key[i0 + 256u * k] : NONE;
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k) {
    b[k] = PENDING;
    if (key[k] == NONE) continue;
    const int cx = (int)(key[k] & KEY_MASK), cy = (int)((key[k] >> KEY_BITS) & KEY_MASK), cz = (int)(key[k] >> (2 * KEY_BITS));
    b[k] = d.block_ptr[block_slot(d.U, cx >> 2, cy >> 2, cz >> 2)];
  }
#pragma unroll
  for (uint32_t k = 0; k < PER; ++k)
    if (b[k] < PENDING) d.sorted[d.block_start[b[k]] + d.rank[i0 + 256u * k]] = i0 + 256u * k;
}

// ---- p2g (shaders/fluid.wgsl:198-303) --------------------------------------------------------------

// axis_weights (fluid.wgsl:221-223)
__device__ __forceinline__ void axis_weights(float fx, float w[3]) {
  w[0] = 0.5f * (1.5f - fx) * (1.5f - fx);
  w[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f);
  w[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
}

// tile_cell (fluid.wgsl:210-217)
__device__ __forceinline__ uint32_t tile_cell(const uint32_t* links, uint32_t t) {
  int rx = (int)(t / 49u) - 1, ry = (int)((t / 7u) % 7u) - 1, rz = (int)(t % 7u) - 1;
  int nx = (rx + 4) / 4 - 1, ny = (ry + 4) / 4 - 1, nz = (rz + 4) / 4 - 1;
  uint32_t b = links[(nx + 1) * 9 + (ny + 1) * 3 + nz + 1];
  if (b == NONE) return NONE;
  return b * 64u + (uint32_t)(rx - nx * 4) * 16u + (uint32_t)(ry - ny * 4) * 4u + (uint32_t)(rz - nz * 4);
}

// P2G's LDS tile: the 7^3 nodes a block's particles reach, x-major with padded strides, so
// that the atomics of a half-wave's stencils spread over more LDS banks. Each lattice's node
// holds a pair of 32-bit sums, the weight's first and then the momentum's (or the pressure's).
constexpr uint32_t P2G_SX = 60, P2G_SY = 8;
constexpr uint32_t P2G_TILE = 6u * P2G_SX + 6u * P2G_SY + 7u;
// The largest weight one particle adds to a node is rint(0.75^3 * FP_W) = 7077888, so the
// weight sums of a block of at most 606 particles stay below 2^32. Such a block adds each
// node's pair with one 64-bit atomic: the weight's sum in the low word never carries into the
// momentum's in the high word, which wraps like the 32-bit atomic it replaces.
constexpr uint32_t P2G_PACK_MAX = 606;
static_assert(NATRAY_A_MOM(1) == 2u && NATRAY_A_WT(1) == 3u && NATRAY_A_PC == NATRAY_A_WC + 1u,
              "P2G's flush maps each lattice's pair onto these fields");

// splat (fluid.wgsl:233-252) of one lattice (offset o = (OX, OY, OZ) / 2 cells) into the tile.
// The reference adds to_fp(weight * node, scale) and to_fp(weight, FP_W) per node. Here the
// powers of two move onto the factors: w24 = (wx * wy) * (wz * FP_W) and node' = node *
// (scale / FP_W), built from val and the gradient terms each scaled by scale / FP_W = 2^-8.
// Scaling by a power of two commutes with rounding in the normal range, so w24 is
// weight * FP_W and w24 * node' rounds to RN(weight * node) * scale, bit for bit. Where a
// subnormal occurs instead (a weight below 2^-126, a scaled term below 2^-126), both products
// are far below 0.5 and convert to 0: such a term changes a sum only while the sum stays within
// 2^25 of it, so after node's three additions it reaches at most 2^-51 and w24 * node' < 2^-27
// (and |node| < 2^100 bounds the other side). This saves two multiplications per node. A
// lattice without gradient (GRAD false: the pressure's, grad = 0) has node = val + (+-0) = val
// for every nonzero val, and +-0 converts to 0 like val = 0, so it splats val alone.
template <int OX, int OY, int OZ, bool GRAD, bool PACK>
__device__ __forceinline__ void splat(int* tile, const float x[3], const int corner[3], uint32_t lat, float val,
                                      const float grad[3], float scale, float h) {
  const float down = scale / NATRAY_FP_W;
  const float o[3] = {0.5f * OX, 0.5f * OY, 0.5f * OZ};
  int base[3];
  float fx[3], hg[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    base[k] = (int)floorf(x[k] - o[k] - 0.5f);
    fx[k] = x[k] - o[k] - (float)base[k];
    hg[k] = (grad[k] * h) * down;
  }
  float wx[3], wy[3], wz[3];
  axis_weights(fx[0], wx);
  axis_weights(fx[1], wy);
  axis_weights(fx[2], wz);
  float wz24[3], gx[3], gy[3], gz[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    wz24[k] = wz[k] * NATRAY_FP_W;
    gx[k] = hg[0] * ((float)k - fx[0]);
    gy[k] = hg[1] * ((float)k - fx[1]);
    gz[k] = hg[2] * ((float)k - fx[2]);
  }
  const float val8 = val * down;
  const int t0 = (base[0] - corner[0]) * (int)P2G_SX + (base[1] - corner[1]) * (int)P2G_SY + (base[2] - corner[2]);
  int* tp = tile + 2 * ((int)(lat * P2G_TILE) + t0);
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const float wxy = wx[a] * wy[b];
      const float gxy = gx[a] + gy[b];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float w24 = wxy * wz24[c];
        const float node = GRAD ? val8 + (gxy + gz[c]) : val8;
        const int32_t wq = (int32_t)__builtin_rintf(w24);
        const int32_t mq = (int32_t)__builtin_rintf(__builtin_amdgcn_fmed3f(w24 * node, -2e9f, 2e9f));
        int* q = tp + 2 * (a * (int)P2G_SX + b * (int)P2G_SY + c);
        if (PACK) {
          atomicAdd(reinterpret_cast<unsigned long long*>(q),
                    (unsigned long long)(uint32_t)mq << 32 | (unsigned long long)(uint32_t)wq);
        } else {
          atomicAdd(q, wq);
          atomicAdd(q + 1, mq);
        }
      }
    }
}

// Neighbour links of block b (fluid.wgsl:261-270), also stored for the later kernels.
__device__ __forceinline__ void block_links(const Dev& d, uint32_t b, uint32_t li, uint32_t* links) {
  const natray_params* U = d.U;
  int bx, by, bz;
  slot_coord(U, d.block_list[b], bx, by, bz);
  int qx = bx + (int)(li / 9u) - 1, qy = by + (int)((li / 3u) % 3u) - 1, qz = bz + (int)(li % 3u) - 1;
  uint32_t idx = NONE;
  if (qx >= 0 && qy >= 0 && qz >= 0 && qx < (int)U->bdims[0] && qy < (int)U->bdims[1] && qz < (int)U->bdims[2]) {
    idx = d.block_ptr[block_slot(U, qx, qy, qz)];
    if (idx >= PENDING) idx = NONE;
  }
  links[li] = idx;
  d.block_nb[b * 27u + li] = idx;
}

// Splats particle i, whose own cell is local cell `l` of the block at (bx, by, bz), into the
// tile: the four lattices of fluid.wgsl:286-289. Returns false (and adds nothing) for a
// particle that is not in that cell, which only a position written behind the library's back
// can cause.
template <bool PACK>
__device__ __forceinline__ bool splat_particle(const Dev& d, int* tile, uint32_t i, uint32_t l, int bx, int by, int bz,
                                               const int corner[3], float h) {
  const natray_params* U = d.U;
  const float4* p = reinterpret_cast<const float4*>(d.P) + (size_t)i * 5;
  const float4 pos = p[0], vel = p[1], c0 = p[2], c1 = p[3], c2 = p[4];
  const float x[3] = {(pos.x - U->origin[0]) / h, (pos.y - U->origin[1]) / h, (pos.z - U->origin[2]) / h};
  const int ox = (int)floorf(x[0]) - bx * 4, oy = (int)floorf(x[1]) - by * 4, oz = (int)floorf(x[2]) - bz * 4;
  if ((uint32_t)ox * 16u + (uint32_t)oy * 4u + (uint32_t)oz != l || (unsigned)ox > 3u || (unsigned)oy > 3u ||
      (unsigned)oz > 3u)
    return false;
  const float g0[3] = {c0.x, c0.y, c0.z}, g1[3] = {c1.x, c1.y, c1.z}, g2[3] = {c2.x, c2.y, c2.z};
  const float zero[3] = {0.0f, 0.0f, 0.0f};
  splat<0, 1, 1, true, PACK>(tile, x, corner, 0u, vel.x, g0, NATRAY_FP_M, h);
  splat<1, 0, 1, true, PACK>(tile, x, corner, 1u, vel.y, g1, NATRAY_FP_M, h);
  splat<1, 1, 0, true, PACK>(tile, x, corner, 2u, vel.z, g2, NATRAY_FP_M, h);
  splat<1, 1, 1, false, PACK>(tile, x, corner, 3u, c0.w * 1e-6f, zero, NATRAY_FP_P, h);
  return true;
}

// Local cell (16x + 4y + z) of a particle's key inside the block at (bx, by, bz), or 64 if
// the key is not in that block.
__device__ __forceinline__ uint32_t key_local(uint32_t key, int bx, int by, int bz) {
  const uint32_t ox = (key & KEY_MASK) - (uint32_t)(bx * 4), oy = ((key >> KEY_BITS) & KEY_MASK) - (uint32_t)(by * 4),
                 oz = (key >> (2 * KEY_BITS)) - (uint32_t)(bz * 4);
  return (ox > 3u || oy > 3u || oz > 3u) ? 64u : ox * 16u + oy * 4u + oz;
}

// Particles of a block that P2G orders in LDS, and the most particles one cell may hold for
// it; a block beyond either splats its particles unordered.
constexpr uint32_t P2G_MAXP = 2048;
constexpr uint32_t P2G_MAXR = 64;

// p2g (fluid.wgsl:254-303). The block's particles are counted per own cell, then laid out in
// LDS round-major: first the first particle of every cell that has one, then every cell's
// second, and so on. The 256 threads splat that list in order, so the lanes of one LDS atomic
// instruction hold particles of different cells (their stencils, not the same words), and no
// lane idles while a crowded cell still has particles. The per-cell counts are A_N.
template <class L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_p2g(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  __shared__ __align__(8) int tile[8 * P2G_TILE];
  __shared__ uint32_t cnt[64], drop[64];
  __shared__ uint32_t order[P2G_MAXP];
  __shared__ uint8_t ocell[P2G_MAXP];
  __shared__ unsigned long long rmask[P2G_MAXR];
  __shared__ uint32_t roff[P2G_MAXR];
  __shared__ uint32_t links[27];
  __shared__ uint32_t s_maxc;
  const uint32_t n = n_blocks(d), tid = threadIdx.x;
  const natray_params* U = d.U;
  const float h = U->origin[3];
  Dealer deal_s;
  __shared__ uint32_t deal_w[2];
  for (uint32_t b = deal_first(d, DEAL_P2G, n, v.wg, v.wgs, deal_s, deal_w); b < n; b = deal_next(deal_w)) {
    if (tid < 27) block_links(d, b, tid, links);
    if (tid < 64) {
      cnt[tid] = 0;
      drop[tid] = 0;
    }
    for (uint32_t t = tid; t < 8 * P2G_TILE; t += 256) tile[t] = 0;
    const uint32_t first = d.block_start[b], count = d.block_start[b + 1] - first;
    __syncthreads();
    deal_ahead(d, DEAL_P2G, n, v.wgs, deal_s, deal_w);
    if (count == 0u) {
      if (tid < 64) accr(d, NATRAY_A_N, b * 64u + tid) = 0;
      continue;
    }
    int bx, by, bz;
    slot_coord(U, d.block_list[b], bx, by, bz);
    const int corner[3] = {bx * 4 - 1, by * 4 - 1, bz * 4 - 1};
    if (count <= P2G_MAXP) {
      // One pass over the block's particles: each thread keeps its (at most eight) particles
      // and their cells in registers, and the count atomic gives each its rank in its cell.
      uint32_t pi[P2G_MAXP / 256], pl[P2G_MAXP / 256];
#pragma unroll
      for (uint32_t j = 0; j < P2G_MAXP / 256; ++j) {
        const uint32_t k = tid + 256u * j;
        pi[j] = k < count ? d.sorted[first + k] : NONE;
      }
#pragma unroll
      for (uint32_t j = 0; j < P2G_MAXP / 256; ++j) {
        pl[j] = NONE;
        if (pi[j] != NONE) {
          const uint32_t l = key_local(d.key[pi[j]], bx, by, bz);
          if (l < 64u) pl[j] = l | atomicAdd(&cnt[l], 1u) << 8;
        }
      }
      __syncthreads();
      if (tid < 64) {
        // Round r holds the cells with more than r particles: rmask[r] marks them, roff[r] is
        // where the round starts in the list.
        const uint32_t c = cnt[tid];
        uint32_t mx = c;
        for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
        if (mx <= P2G_MAXR) {
          uint32_t off = 0;
          for (uint32_t r = 0; r < mx; ++r) {
            const unsigned long long m = __ballot(c > r);
            if (tid == 0) {
              rmask[r] = m;
              roff[r] = off;
            }
            off += (uint32_t)__popcll(m);
          }
        }
        if (tid == 0) s_maxc = mx;
      }
      __syncthreads();
      const uint32_t maxc = s_maxc;
      if (maxc <= P2G_MAXR) {
#pragma unroll
        for (uint32_t j = 0; j < P2G_MAXP / 256; ++j) {
          if (pl[j] == NONE) continue;
          const uint32_t l = pl[j] & 255u, r = pl[j] >> 8;
          const uint32_t pos = roff[r] + (uint32_t)__popcll(rmask[r] & ((1ull << l) - 1ull));
          order[pos] = pi[j];
          ocell[pos] = (uint8_t)l;
        }
        __syncthreads();
        uint32_t placed = 0;
        for (uint32_t r = 0; r < maxc; ++r) placed += (uint32_t)__popcll(rmask[r]);
        if (count <= P2G_PACK_MAX) {
          for (uint32_t k = tid; k < placed; k += 256) {
            const uint32_t l = ocell[k];
            if (!splat_particle<true>(d, tile, order[k], l, bx, by, bz, corner, h)) atomicAdd(&drop[l], 1u);
          }
        } else {
          for (uint32_t k = tid; k < placed; k += 256) {
            const uint32_t l = ocell[k];
            if (!splat_particle<false>(d, tile, order[k], l, bx, by, bz, corner, h)) atomicAdd(&drop[l], 1u);
          }
        }
      } else {
        // A cell with more than P2G_MAXR particles: splat in stored order (counted above).
        for (uint32_t k = tid; k < count; k += 256) {
          const uint32_t i = d.sorted[first + k];
          const uint32_t l = key_local(d.key[i], bx, by, bz);
          if (l < 64u && !splat_particle<false>(d, tile, i, l, bx, by, bz, corner, h)) atomicAdd(&drop[l], 1u);
        }
      }
    } else {
      for (uint32_t k = tid; k < count; k += 256) {
        const uint32_t i = d.sorted[first + k];
        const uint32_t l = key_local(d.key[i], bx, by, bz);
        if (l < 64u && splat_particle<false>(d, tile, i, l, bx, by, bz, corner, h)) atomicAdd(&cnt[l], 1u);
      }
    }
    __syncthreads();
    for (uint32_t t = tid; t < TILE; t += 256) {
      uint32_t cell = tile_cell(links, t);
      if (cell == NONE) continue;
      const uint32_t pt = t / 49u * P2G_SX + t / 7u % 7u * P2G_SY + t % 7u;
#pragma unroll
      for (uint32_t lat = 0; lat < 4u; ++lat)
#pragma unroll
        for (uint32_t j = 0; j < 2u; ++j) {
          // j = 0: the weight (A_WT(lat), or A_WC); j = 1: the momentum (A_MOM(lat)), or A_PC.
          const uint32_t f = lat < 3u ? 2u * lat + 1u - j : NATRAY_A_WC + j;
          const int v = tile[2u * (lat * P2G_TILE + pt) + j];
          if (v != 0) atomicAdd(&accr(d, f, cell), v);
        }
    }
    if (tid < 64) accr(d, NATRAY_A_N, b * 64u + tid) = (int32_t)(cnt[tid] - drop[tid]);
  }
}

// ---- per-cell kernels: one wave per block --------------------------------------------------------

// Runs body(b, lane, lx, ly, lz, links) for the blocks b_lo, b_lo + stride, ... below b_hi,
// WAVES blocks at a time per workgroup (one per wave); `links` are the block's 27 neighbours
// in LDS.
template <uint32_t WAVES = 4, class Body>
__device__ __forceinline__ void for_blocks(const Dev& d, uint32_t b_lo, uint32_t b_hi, uint32_t stride, Body body) {
  __shared__ uint32_t lk[WAVES][27];
  const uint32_t w = threadIdx.x >> 6, lane = lane_id();
  const int lx = (int)(lane >> 4), ly = (int)((lane >> 2) & 3u), lz = (int)(lane & 3u);
  for (uint32_t base = b_lo; base < b_hi; base += stride) {
    const uint32_t b = base + w;
    const bool act = b < b_hi;
    __syncthreads();
    if (act && lane < 27) lk[w][lane] = d.block_nb[b * 27u + lane];
    __syncthreads();
    if (act) body(b, lane, lx, ly, lz, lk[w]);
  }
}

// grid_prep (fluid.wgsl:319-356)
template <class L>
__global__ __launch_bounds__(256) void k_grid_prep(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const float dt = dt_of(d);
  const natray_params* U = d.U;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u, [&](uint32_t b, uint32_t lane, int, int, int, const uint32_t*) {
    const uint32_t cell = b * 64u + lane;
    // Every sum the cell may use is read up front, so the loads go out together rather than
    // behind the tests that pick among them.
    static_assert(NATRAY_A_N < 11u && NATRAY_A_PC < 11u && NATRAY_A_WC < 11u && NATRAY_A_MF < 11u && NATRAY_A_MA < 11u &&
                      NATRAY_A_WT(2) < 11u && NATRAY_A_MOM(2) < 11u,
                  "grid_prep's sums are the first 11 fields and A_MW");
    int32_t a[11];
#pragma unroll
    for (uint32_t f = 0; f < 11u; ++f) a[f] = accr(d, f, cell);
    const int32_t a_mw = accr(d, NATRAY_A_MW, cell);
    int cx, cy, cz;
    cell_coord(d, cell, cx, cy, cz);
    float ty = GAS;
    if (is_solid_cell(d, cx, cy, cz)) {
      ty = SOLID;
    } else if (a[NATRAY_A_N] > 0) {
      ty = WATER;
    }
    cfr(d, NATRAY_F_TYPE, cell) = ty;
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
      float w = from_fp(a[NATRAY_A_WT(k)], NATRAY_FP_W);
      float u = 0.0f;
      if (w > 1e-6f) u = from_fp(a[NATRAY_A_MOM(k)], NATRAY_FP_M) / w + dt * U->gravity[k];
      cfr(d, NATRAY_F_U + k, cell) = u;
    }
    float wc = from_fp(a[NATRAY_A_WC], NATRAY_FP_W);
    float ma = from_fp(a[NATRAY_A_MA], NATRAY_FP_A);
    if (ty == GAS && a_mw > 0) {
      ty = RUBBER;
      cfr(d, NATRAY_F_TYPE, cell) = ty;
    }
    float p = 0.0f;
    if (ty == WATER && wc > 1e-6f) {
      p = from_fp(a[NATRAY_A_PC], NATRAY_FP_P) / wc * 1e6f;
    } else if (ty != WATER && ty != SOLID && ma > 0.0f) {
      ty = COVERED;
      cfr(d, NATRAY_F_TYPE, cell) = ty;
      p = from_fp(a[NATRAY_A_MF], NATRAY_FP_F) / ma * 1e6f;
    } else if (ty != WATER && ty != SOLID && (d.sc->flags & NATRAY_SCENE_MASK) && in_mask(d, cx, cy, cz)) {
      ty = COVERED;
      cfr(d, NATRAY_F_TYPE, cell) = ty;
      p = d.sc->mask_p;
    }
    cfr(d, NATRAY_F_P, cell) = p;
  });
}

// ---- pressure solve (fluid.wgsl:381-578) ------------------------------------------------------------
//
// One cooperative launch runs the whole CG: solve_init, then solve_q .. solve_beta for every
// iteration. The n blocks form m = ceil(n / K) chunks of K = ceil(n / MAX_CHUNKS) consecutive
// blocks (both fixed by n alone, so the sums do not depend on the launch); workgroup g owns
// the chunks [g m / G, (g + 1) m / G) and their blocks, wave w the blocks b0 + w + 16 r.
//
// Workgroups talk through 64-bit words that hold a 32-bit tag above a 32-bit payload, stored
// and loaded with agent-scope relaxed atomics: a reader waits until a word carries the tag
// it expects, and the payload beside it is then the value stored with that tag. Tags number
// the solve's reductions over the life of the buffers (Dev::bar[4] holds the last one used),
// and the host clears every tagged word long before the count could wrap.
//  - A reduction: each workgroup folds its chunks and publishes the chunk sums (one word per
//    f32, two per f64); every workgroup waits for all chunk words and sums them in total()'s
//    order itself, so no workgroup relays for the others. Consecutive reductions alternate
//    between two word buffers: a workgroup can only publish reduction t + 2 after it has seen
//    every chunk of t + 1, which every workgroup publishes after it has read all of t.
//  - Halo z: a wave that keeps its blocks in registers publishes z of its water cells with
//    the tag of the reduction that follows; neighbours read it in the next solve_q. The other
//    halo values a block needs (the old direction; p for the residual refresh) are rebuilt by
//    the reader from values it already holds, with the owner's arithmetic, so they are equal.
// When the blocks of a workgroup do not fit in registers (a launch with too few workgroups
// for the block count) every workgroup runs the same iteration on the cf fields in memory,
// with agent-scope release and acquire fences around each reduction. Both forms evaluate the
// same expressions in the same order and give the same bits.

// The pipelined solve replaces its residual only while |r| is above this many times the
// tolerance test's bound.
constexpr double PIPE_REPLACE = 1000.0;
// Polls of a solve wait before a wave gives up and raises the abort flag (Dev::bar[2]).
constexpr uint32_t kSpinLimit = 1u << 22;
// Solve iterations one launch may run, whatever limits.z asks: bounds the tags a launch uses.
constexpr uint32_t SOLVE_ITER_CAP = 1u << 24;

constexpr uint32_t SOLVE_THREADS = 1024;
constexpr uint32_t SOLVE_WAVES = SOLVE_THREADS / 64;
// Blocks a wave of the solve keeps in registers.
constexpr uint32_t SOLVE_RMAX = 2;
// The memory form's streaming passes: at most this many blocks per wave, the loads of this
// many blocks issued together.
constexpr uint32_t SOLVE_JMAX = 32;
constexpr uint32_t SOLVE_GROUP = 2;
// Words of one reduction buffer per chunk: 4 components x 2 halves. Word w is component w / H,
// half w % H (H = 1 for f32, 2 for f64); words 2j and 2j + 1 travel together as one 16-byte
// access, so the words of a chunk arrive in as few loads as they take.
constexpr uint32_t RED_SLOTS = 8;
constexpr uint32_t RED_PAIRS = RED_SLOTS / 2;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// Cache policy of the chunk records' 16-byte loads and stores: sc1, the bits of the agent-scope
// atomics the other tagged words use.
constexpr int kRedPolicy = 16;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t red_rsrc(const Dev& d) {
  return __builtin_amdgcn_make_buffer_rsrc(d.rwords, 0, (int)(RED_WORDS * 8u), 0x00020000);
}
// Byte offset of pair j of buffer `buf` in chunk c's record.
__device__ __forceinline__ uint32_t red_pair(uint32_t c, uint32_t buf, uint32_t j) {
  return ((c * RED_RECORD + buf * RED_SLOTS) + 2u * j) * 8u;
}

__device__ __forceinline__ unsigned long long tag_word(uint32_t tag, uint32_t payload) {
  return ((unsigned long long)tag << 32) | payload;
}
__device__ __forceinline__ unsigned long long load_word(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_word(unsigned long long* p, unsigned long long w) {
  __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Workgroup barrier of the register form: orders LDS only, so global stores to other
// workgroups' words and loads whose results are used later stay in flight across it.
// `all` makes it a full barrier (global memory too), as the memory form needs.
__device__ __forceinline__ void wg_sync(bool all) {
  if (all) {
    __syncthreads();
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
  }
}

// A reduction value as 32-bit payloads.
template <class R>
struct Halves;
template <>
struct Halves<float> {
  static constexpr int N = 1;
  static __device__ __forceinline__ uint32_t part(float v, int) { return __float_as_uint(v); }
  static __device__ __forceinline__ float join(uint32_t hi, uint32_t) { return __uint_as_float(hi); }
};
template <>
struct Halves<double> {
  static constexpr int N = 2;
  static __device__ __forceinline__ uint32_t part(double v, int i) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return i ? (uint32_t)b : (uint32_t)(b >> 32);
  }
  static __device__ __forceinline__ double join(uint32_t hi, uint32_t lo) {
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
  }
};

// block_sum's tree (fluid.wgsl:414-422) inside one wave; lane 0 ends with the sum.
template <class R>
__device__ __forceinline__ R wave_tree(R v) {
  return lane_tree<R>(v);
}

// Block-local symmetric red-black Gauss-Seidel (fluid.wgsl:384-410), five half sweeps from
// zero, inside one wave; or Jacobi.
__device__ __forceinline__ float precondition(float r, float diag, bool water, int lx, int ly, int lz, uint32_t mode) {
  if (mode == NATRAY_PRECOND_JACOBI) return water ? r / diag : 0.0f;
  const uint32_t lane = lane_id();
  const unsigned long long wm = __ballot(water);
  const bool wxm = lx > 0 && ((wm >> (lane - 16u)) & 1ull), wxp = lx < 3 && ((wm >> (lane + 16u)) & 1ull);
  const bool wym = ly > 0 && ((wm >> (lane - 4u)) & 1ull), wyp = ly < 3 && ((wm >> (lane + 4u)) & 1ull);
  const bool wzm = lz > 0 && ((wm >> (lane - 1u)) & 1ull), wzp = lz < 3 && ((wm >> (lane + 1u)) & 1ull);
  const uint32_t colour = (uint32_t)(lx + ly + lz) & 1u;
  // The first half sweep starts from z = +0, so each wet neighbour adds +0: s stays r, except
  // that r = -0 becomes +0 once any neighbour is added.
  float z = 0.0f;
  if (water && colour == 0u) z = ((wxm || wxp || wym || wyp || wzm || wzp) ? r + 0.0f : r) / diag;
#pragma unroll
  for (uint32_t k = 1; k < 5u; ++k) {
    const Faces f = block_faces(z);
    if (water && colour == (k & 1u)) {
      float s = r;
      if (wxm) s += f.xm;
      if (wxp) s += f.xp;
      if (wym) s += f.ym;
      if (wyp) s += f.yp;
      if (wzm) s += f.zm;
      if (wzp) s += f.zp;
      z = s / diag;
    }
  }
  return z;
}

template <class R>
__device__ __forceinline__ R prod(float a, float b) {
  return (R)a * (R)b;
}

__device__ __forceinline__ float sq(float x) { return __builtin_sqrtf(x); }
__device__ __forceinline__ double sq(double x) { return __builtin_sqrt(x); }

// converged (fluid.wgsl:441-443)
template <class R>
__device__ __forceinline__ bool converged(R rr, R tol, R b_norm) {
  return sq(rr) <= tol * (b_norm < (R)1 ? (R)1 : b_norm);
}

// The blocks and chunks of one solve workgroup.
struct SolveSpan {
  uint32_t n, K, m, G;
  uint32_t wg;      // this workgroup among the G of its scene
  uint32_t c0, c1;  // chunks of this workgroup
  uint32_t b0, b1;  // their blocks
  bool in_regs;     // every workgroup's blocks fit in registers (the same for all)
  bool stream;      // the memory form's streaming passes (the same for all)
  bool coherent;    // a reduction also publishes the plain stores before it to the other
                    // workgroups (agent-scope release and acquire)
};

// Work for the waves of a workgroup while a reduction completes: `wait(tag)` is called once
// per polling round and returns true once the wave has what it waits for. NoWait has nothing.
struct NoWait {
  __device__ __forceinline__ bool operator()(uint32_t) const { return true; }
};

// The register form's block partials of one workgroup in LDS, in the reduction's precision and
// component-major: v[k][b - b0]. A block's four components: the reduction's values, or (at the
// start) r.z, r.r, b.b and its water cells (a count, exact in either precision).
template <class R>
struct LPart {
  R v[4][SOLVE_WAVES * SOLVE_RMAX];
};

// The chunk fold's tree width: the next power of two >= K (at most MAX_CHUNKS).
__host__ __device__ __forceinline__ uint32_t chunk_width(uint32_t K) {
  uint32_t w = 1u;
  while (w < K) w <<= 1;
  return w;
}

// v[l] += v[l + off] for off = off0, off0 / 2, .., 1 (off0 <= 32, a power of two or 0) across
// the wave's lanes; lane 0 ends with the tree's sum of lanes [0, 2 off0).
template <class R>
__device__ __forceinline__ R lane_tree_from(R v, uint32_t off0) {
  if (off0 >= 32u) v += lane_xor32(v);
  if (off0 >= 16u) v += lane_xor16(v);
  if (off0 >= 8u) v += lane_up<8>(v);
  if (off0 >= 4u) v += lane_up<4>(v);
  if (off0 >= 2u) v += lane_up<2>(v);
  if (off0 >= 1u) v += lane_up<1>(v);
  return v;
}

// Component k of chunk [lo, hi)'s tree sum (the chunk fold: a binary tree over W slots, slot i
// = block lo + i's partial, zero past the chunk's last block), in lane 0 of the calling wave (all
// lanes active). The partials come from `lp` (register form) or Dev::part (memory form). Lane l
// holds slots l, l + 64, .. and combines them in the tree's order first.
template <class R>
__device__ __forceinline__ R chunk_tree(const Dev& d, const SolveSpan& sp, const LPart<R>* lp, uint32_t lo, uint32_t hi,
                                        uint32_t W, int k, uint32_t ln) {
  auto slot = [&](uint32_t i) -> R {
    const uint32_t b = min(lo + i, hi - 1u);
    const R x = sp.in_regs ? lp->v[k][b - sp.b0] : (R)d.part[b * 4u + (uint32_t)k];
    return lo + i < hi ? x : (R)0;
  };
  if (W == 256u) return lane_tree_from<R>((slot(ln) + slot(ln + 128u)) + (slot(ln + 64u) + slot(ln + 192u)), 32u);
  if (W == 128u) return lane_tree_from<R>(slot(ln) + slot(ln + 64u), 32u);
  return lane_tree_from<R>(slot(ln), W / 2u);
}

// Lane 0's value in every lane (all lanes active).
__device__ __forceinline__ float first_lane(float x) {
  return __uint_as_float((uint32_t)__builtin_amdgcn_readfirstlane((int)__float_as_uint(x)));
}
__device__ __forceinline__ double first_lane(double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A reduction's LDS: the chunk sums, and the abort flag per tag parity (a wave can leave once it
// has read it; it is rewritten two reductions later, after barriers every wave has passed).
template <class R>
struct RedLds {
  R red[4][256];
  uint32_t fail[2];
};

// Reduction `tag` over components [0, nc) of the block partials (nc = 0: a barrier only), in two
// halves. reduce_publish: once every wave has its block partials in `lp` (blocks b0.., register
// form) or Dev::part, the chunk folds are published. reduce_wait: every workgroup waits for
// every chunk and returns the same totals (false when the solve aborted); `wait` runs in every
// wave alongside the polling (see NoWait). Work that does not depend on the totals can run
// between the two halves, while the other workgroups' chunks arrive.
template <class R>
__device__ void reduce_publish(const Dev& d, const SolveSpan& sp, uint32_t tag, int nc, const LPart<R>* lp, RedLds<R>& L) {
  constexpr int H = Halves<R>::N;
  const uint32_t t = threadIdx.x;
  const int nw = nc > 0 ? nc : 1;
  const uint32_t buf = tag & 1u;
  if (t == 0) L.fail[buf] = 0u;
  wg_sync(!sp.in_regs);
  // The chunk fold, then publish: per chunk and pair of words, the components of both words
  // summed over the chunk's blocks, and both words (tag above payload each) stored as one
  // 16-byte record.
  const __amdgpu_buffer_rsrc_t rs = red_rsrc(d);
  const uint32_t words = (uint32_t)(nw * H), P = (words + 1u) / 2u;
  {
    // A wave per chunk and pair of words, its lanes over the chunk's blocks.
    const uint32_t wv = t >> 6, ln = t & 63u, tasks = (sp.c1 - sp.c0) * P, W = chunk_width(sp.K);
    for (uint32_t task = wv; task < tasks; task += SOLVE_WAVES) {
      const uint32_t j = task % P, c = sp.c0 + task / P, lo = c * sp.K, hi = min(sp.n, lo + sp.K);
      // The components of words 2j and 2j + 1 (the same one for f64, whose halves they are).
      const int k0 = (int)(2u * j) / H, k1 = (int)(2u * j + 1u) / H;
      R a0 = chunk_tree<R>(d, sp, lp, lo, hi, W, k0, ln);
      R a1 = k1 == k0 ? a0 : chunk_tree<R>(d, sp, lp, lo, hi, W, k1, ln);
      if (k0 >= nc) a0 = (R)0;
      if (k1 >= nc) a1 = (R)0;
      if (ln == 0) {
        if (sp.coherent) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const u32x4 rec = {Halves<R>::part(a0, (int)(2u * j) % H), tag, Halves<R>::part(a1, (int)(2u * j + 1u) % H), tag};
        __builtin_amdgcn_raw_buffer_store_b128(rec, rs, red_pair(c, buf, j), 0, kRedPolicy);
      }
    }
  }
}

template <class R, class Wait = NoWait>
__device__ bool reduce_wait(const Dev& d, const SolveSpan& sp, uint32_t tag, int nc, R* out, RedLds<R>& L,
                            Wait wait = Wait()) {
  constexpr int H = Halves<R>::N;
  const uint32_t t = threadIdx.x;
  const int nw = nc > 0 ? nc : 1;
  const uint32_t buf = tag & 1u;
  uint32_t* abort = d.bar + 2;
  const __amdgpu_buffer_rsrc_t rs = red_rsrc(d);
  const uint32_t P = ((uint32_t)(nw * H) + 1u) / 2u;
  // Every chunk's words, summed as total() does (fluid.wgsl:426-439): thread t holds chunk t,
  // then the tree over the 256 threads.
  R v[4] = {(R)0, (R)0, (R)0, (R)0};
  {
    const bool poller = t < 256u && t < sp.m;
    uint32_t pay[RED_SLOTS];
    const uint32_t full = (1u << P) - 1u;
    // A thread without a chunk has every pair.
    uint32_t have = poller ? 0u : full, spins = 0;
#pragma unroll
    for (uint32_t w = 0; w < RED_SLOTS; ++w) pay[w] = 0u;
    while (true) {
      // Issue every load of the round before looking at any, so one round trip serves both
      // the chunk words and the waves' own waits. The compiler fence keeps each round's loads
      // in the round (a record changes under the loop, from other workgroups).
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      u32x4 rec[RED_PAIRS];
#pragma unroll
      for (uint32_t j = 0; j < RED_PAIRS; ++j)
        rec[j] = (j < P && !((have >> j) & 1u)) ? __builtin_amdgcn_raw_buffer_load_b128(rs, red_pair(t, buf, j), 0, kRedPolicy)
                                                  : u32x4{0u, 0u, 0u, 0u};
      const bool waited = wait(tag);
#pragma unroll
      for (uint32_t j = 0; j < RED_PAIRS; ++j) {
        if (j >= P || ((have >> j) & 1u)) continue;
        // Each word's own tag decides; a record counts once both of its words carry the tag.
        if (rec[j].y == tag && rec[j].w == tag) {
          pay[2u * j] = rec[j].x;
          pay[2u * j + 1u] = rec[j].z;
          have |= 1u << j;
        }
      }
      if (have == full && waited) break;
      __builtin_amdgcn_s_sleep(1);
      if ((++spins & 255u) == 0u &&
          (__hip_atomic_load(abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kSpinLimit)) {
        __hip_atomic_store(abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        L.fail[buf] = 1u;
        break;
      }
    }
    if (poller) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k < nc) v[k] = (R)0 + Halves<R>::join(pay[k * H], pay[k * H + H - 1]);
    }
  }
  if (t < 256u) {
    if (sp.coherent) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (k < nc) L.red[k][t] = v[k];
  }
  wg_sync(!sp.in_regs);
  if (nc > 0) {
    // total()'s tree, in every wave alike: lane l sums (c_l + c_l+128) + (c_l+64 + c_l+192), the
    // tree's first two levels, then the 64 lanes, and lane 0's sum goes to the whole wave. `red`
    // is next written after the next reduction's first barrier, which each wave reaches only
    // after these reads.
    const uint32_t ln = t & 63u;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= nc) continue;
      const R x = lane_tree<R>((L.red[k][ln] + L.red[k][ln + 128u]) + (L.red[k][ln + 64u] + L.red[k][ln + 192u]));
      out[k] = first_lane(x);
    }
  }
  const bool ok = L.fail[buf] == 0u;
  if (!sp.in_regs) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  return ok;
}

template <class R, class Wait = NoWait>
__device__ bool reduce(const Dev& d, const SolveSpan& sp, uint32_t tag, int nc, const LPart<R>* lp, R* out,
                       Wait wait = Wait()) {
  __shared__ RedLds<R> L;
  reduce_publish<R>(d, sp, tag, nc, lp, L);
  return reduce_wait<R, Wait>(d, sp, tag, nc, out, L, wait);
}

// The solve's row of one cell (solve_init, fluid.wgsl:450-489): whether it is water, the
// diagonal, the right-hand side b (from the particle-carried pressure, always), the mask of
// water neighbours, and the residual b - A p0 for the starting pressure p0 (the carried
// pressure, or the warm start's guess).
struct Row {
  bool water;
  float diag, bv, r;
  uint32_t mask;
};
__device__ __forceinline__ Row cell_row(const Dev& d, uint32_t cell, int lx, int ly, int lz, const uint32_t* lk,
                                        const float* p0) {
  const natray_params* U = d.U;
  const float h = U->origin[3], dt = dt_of(d);
  const float* cP = d.cf + (size_t)NATRAY_F_P * d.capc;
  const float tc = cfr(d, NATRAY_F_TYPE, cell);
  Row w = {tc == WATER, 1.0f, 0.0f, 0.0f, 0u};
  if (!w.water) return w;
  // Every value the row may use is loaded before any test picks among them, so the loads go
  // out together; the arithmetic below is the reference's, in its order.
  uint32_t nb[6];  // k = 2 dim + (0: below, 1: above)
  float tn[6], pn[6], qn[6], uu[3], uc[3];
#pragma unroll
  for (uint32_t k = 0; k < 6u; ++k) {
    const int sg = (k & 1u) ? 1 : -1, dim = (int)(k >> 1);
    nb[k] = nbr(lk, lx, ly, lz, dim == 0 ? sg : 0, dim == 1 ? sg : 0, dim == 2 ? sg : 0);
    tn[k] = cell_type(d, nb[k]);
    pn[k] = nb[k] != NONE ? cP[nb[k]] : 0.0f;
    qn[k] = nb[k] != NONE ? p0[nb[k]] : 0.0f;
  }
#pragma unroll
  for (uint32_t dim = 0; dim < 3u; ++dim) {
    uc[dim] = cfr(d, NATRAY_F_U + dim, cell);
    uu[dim] = nb[2 * dim + 1] != NONE ? cfr(d, NATRAY_F_U + dim, nb[2 * dim + 1]) : 0.0f;
  }
  const int32_t src = accr(d, NATRAY_A_SRC, cell);
  const float pc = cP[cell], p0c = p0[cell];
  float cs = U->gravity[3] * dt;
  float kc = h * h / (cs * cs);
  float div = 0.0f, off = 0.0f;
  w.diag = kc;
  w.bv = kc * pc;
#pragma unroll
  for (uint32_t dim = 0; dim < 3; ++dim) {
    const uint32_t up = nb[2 * dim + 1], dn = nb[2 * dim];
    // face_u(up, cell) - face_u(cell, dn) (fluid.wgsl:368-371)
    const float fu = up == NONE ? 0.0f : bounded(d, uu[dim], tn[2 * dim + 1], tc, up, cell, dim);
    div += fu - bounded(d, uc[dim], tc, tn[2 * dim], cell, dn, dim);
#pragma unroll
    for (uint32_t e = 0; e < 2u; ++e) {
      const uint32_t k = 2 * dim + e;
      const float nt = tn[k];
      if (nt == SOLID) continue;
      w.diag += 1.0f;
      if (nt == WATER) {
        off -= qn[k];
        w.mask |= 1u << k;
      } else {
        // pressure_of (fluid.wgsl:375-379) of a cell that is not water, 0 past the grid
        w.bv += nb[k] == NONE ? 0.0f : pn[k];
      }
    }
  }
  div -= from_fp(src, NATRAY_FP_Q) / (h * h);
  w.bv -= U->fluid[0] * h / dt * div;
  w.r = w.bv - w.diag * p0c - off;
  w.mask |= 64u;
  return w;
}

// solve_init (fluid.wgsl:450-491) for one cell from p0: the operator row, b, r = b - A p0 and
// z. Writes diag, r, b, the neighbour mask and z; in register form also z as a tagged word
// (tag `ztag`) into zw. Stores the block partials (r.z, r.r, b.b, water cells) through `prow`
// from lane 0.
template <class R>
__device__ __forceinline__ void init_cell(const Dev& d, uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk,
                                          unsigned long long* zw, uint32_t ztag, R* prow, const float* p0) {
  const uint32_t capc = d.capc;
  const uint32_t cell = b * 64u + lane;
  const Row w = cell_row(d, cell, lx, ly, lz, lk, p0);
  if (w.water) {
    d.cf[(size_t)NATRAY_F_DIAG * capc + cell] = w.diag;
    d.cf[(size_t)NATRAY_F_R * capc + cell] = w.r;
    d.bvec[cell] = w.bv;
  }
  d.nbm[cell] = w.water ? w.mask : 0u;
  const float r = w.water ? w.r : 0.0f, bv = w.water ? w.bv : 0.0f;
  const float z = precondition(r, w.diag, w.water, lx, ly, lz, d.sc->precond);
  if (w.water) {
    d.cf[(size_t)NATRAY_F_Z * capc + cell] = z;
    if (zw) store_word(&zw[cell], tag_word(ztag, __float_as_uint(z)));
  }
  const R v0 = wave_tree<R>(prod<R>(r, z)), v1 = wave_tree<R>(prod<R>(r, r)), v2 = wave_tree<R>(prod<R>(bv, bv));
  const uint32_t nw = (uint32_t)__popcll(__ballot(w.water));
  if (lane == 0) {
    prow[0] = v0;
    prow[1] = v1;
    prow[2] = v2;
    prow[3] = (R)nw;
  }
}

// The starting pressure of this solve: the warm start's guess when it was accepted this step
// (Dev::bar[11]), else the particle-carried pressure.
__device__ __forceinline__ const float* start_pressure(const Dev& d) {
  return d.bar[11] ? d.guess : d.cf + (size_t)NATRAY_F_P * d.capc;
}

// The solver scalars of a solve, kept by every workgroup alike.
template <class R>
struct SolveScalars {
  R rz, alpha, beta, b_norm, r_norm;
  uint32_t iters, iters_max, unconverged, water;
};

// solve_start (fluid.wgsl:493-506) from the init totals; returns whether to iterate.
template <class R>
__device__ __forceinline__ bool solve_start(const Dev& d, const R* tt, bool ok, SolveScalars<R>& s) {
  const R tol = (R)d.U->solve[0];
  s.rz = tt[0];
  s.alpha = s.beta = (R)0;
  s.b_norm = sq(tt[2]);
  s.r_norm = sq(tt[1]);
  s.iters = 0;
  s.iters_max = d.st->iters_max;
  s.unconverged = d.st->unconverged;
  // The fourth component counts water cells (exact: at most 2^22 of them).
  s.water = (uint32_t)tt[3];
  return ok && s.rz > (R)0 && !converged<R>(tt[1], tol, s.b_norm) && min(d.U->limits[2], SOLVE_ITER_CAP) > 0u;
}

// solve_beta (fluid.wgsl:561-578) from the update totals; returns whether to iterate again.
template <class R>
__device__ __forceinline__ bool solve_beta(const Dev& d, const R* tt, SolveScalars<R>& s) {
  const R rz_new = tt[0], rr_new = tt[1];
  s.iters += 1u;
  s.r_norm = sq(rr_new);
  s.beta = s.rz > (R)0 ? rz_new / s.rz : (R)0;
  s.rz = rz_new;
  const bool done = converged<R>(rr_new, (R)d.U->solve[0], s.b_norm);
  if (done || s.iters >= min(d.U->limits[2], SOLVE_ITER_CAP)) {
    s.iters_max = s.iters > s.iters_max ? s.iters : s.iters_max;
    s.unconverged += done ? 0u : 1u;
    return false;
  }
  return true;
}

// End of a solve: the scene's workgroup 0 writes State, the solver scalars, the counters and
// the last tag.
template <class R>
__device__ __forceinline__ void solve_finish(const Dev& d, const SolveSpan& sp, uint32_t tag, const SolveScalars<R>& s) {
  if (sp.wg != 0 || threadIdx.x != 0) return;
  const uint32_t n = sp.n;
  __hip_atomic_store(d.bar + 4, tag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  natray_state* st = d.st;
  st->rz = (float)s.rz;
  st->alpha = (float)s.alpha;
  st->beta = (float)s.beta;
  st->b_norm = (float)s.b_norm;
  st->r_norm = (float)s.r_norm;
  st->solve_iters = s.iters;
  st->iters_max = s.iters_max;
  st->unconverged = s.unconverged;
  d.solver[0] = (double)s.rz;
  d.solver[1] = (double)s.alpha;
  d.solver[2] = (double)s.beta;
  d.solver[3] = (double)s.b_norm;
  d.solver[4] = (double)s.r_norm;
  d.stats[0] += 1ull;
  d.stats[1] += s.iters;
  d.stats[2] += (unsigned long long)s.water * s.iters;
  d.stats[3] += n;
  d.stats[4] += s.water;
  d.stats[5] += d.block_start[n];
  d.stats[6] += st->n_blocks_prev;
}

// Cell i of a field through a 32-bit byte offset over the field's base (a field holds fewer
// than 2^30 cells): one address register per cell instead of a 64-bit pointer.
__device__ __forceinline__ float ld_cell(const float* f, uint32_t i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(f) + (i << 2));
}
__device__ __forceinline__ void st_cell(float* f, uint32_t i, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(f) + (i << 2)) = v;
}

// The memory form of the solve: the same iteration on the cf fields in memory, for launches
// whose workgroups cannot keep their blocks in registers.
//
// Streaming passes (sp.stream: every workgroup holds at most SOLVE_WAVES x SOLVE_JMAX blocks):
// wave wv walks its own blocks b0 + wv + SOLVE_WAVES j with no workgroup barrier, the same
// blocks in every pass, so a cell's own fields are only ever read back by the wave that wrote
// them. Values a cell's neighbours in other workgroups need travel as tagged words (tag above
// the float), as in the register form: z in Dev::hz[cell] with the tag of the reduction after
// the pass that made it, the new direction in hz[capc (1 + k % 2) + cell] with the tag of
// iteration k's first reduction, and before a residual refresh p in the other direction buffer
// (whose last words every neighbour has read) with the tag of the refresh's reduction. A reader
// waits for the tag it expects, so the reductions need no agent-scope fences. Each wave keeps
// its blocks' neighbour masks and their six face-neighbour blocks (-x, +x, -y, +y, -z, +z) in
// LDS, walks its blocks in a rolled loop, issues the loads of SOLVE_GROUP blocks before it uses
// any, and takes a cell's neighbours inside the block from the wave's lanes. The arithmetic is
// that of the per-block passes, expression for expression.
//
// Per-block passes (larger workgroups): for_blocks, the workgroup's waves in lock step, with
// agent-scope release and acquire fences around each reduction (sp.coherent), since the
// neighbours' fields are read from memory.
template <bool MIXED>
__device__ __noinline__ void solve_memory(const Dev* dp, SolveSpan sp, uint32_t tag) {
  using R = typename std::conditional<MIXED, double, float>::type;
  const Dev& d = *dp;
  const uint32_t mode = d.sc->precond;
  const uint32_t b0 = sp.b0, b1 = sp.b1;
  const uint32_t capc = d.capc;
  const bool stream = sp.stream;
  float* const cP = d.cf + (size_t)NATRAY_F_P * capc;
  float* const cR = d.cf + (size_t)NATRAY_F_R * capc;
  float* const cZ = d.cf + (size_t)NATRAY_F_Z * capc;
  float* const cQ = d.cf + (size_t)NATRAY_F_Q * capc;
  float* const cDiag = d.cf + (size_t)NATRAY_F_DIAG * capc;
  unsigned long long* const zw = d.hz;
  const float* const p0 = start_pressure(d);
  for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    R v[4];
    init_cell<R>(d, b, lane, lx, ly, lz, lk, stream ? zw : nullptr, tag + 1u, v, p0);
    if (lane == 0)
      for (uint32_t k = 0; k < 4u; ++k) d.part[b * 4u + k] = (double)v[k];
  });
  R tt[4];
  SolveScalars<R> s;
  bool go = solve_start<R>(d, tt, reduce<R>(d, sp, ++tag, 4, nullptr, tt), s);
  if (p0 != cP) {
    // The iterate starts at the guess; every workgroup has read the carried pressure by now.
    for (uint32_t c = b0 * 64u + threadIdx.x; c < b1 * 64u; c += blockDim.x)
      if (d.nbm[c] & 64u) cP[c] = p0[c];
    __syncthreads();
  }
  const bool refresh_on = MIXED && d.sc->precision == NATRAY_CG_MIXED && d.sc->refresh_every > 0u;

  const uint32_t wv = threadIdx.x >> 6, ln = lane_id();
  const int lx = (int)(ln >> 4), ly = (int)((ln >> 2) & 3u), lz = (int)(ln & 3u);
  const uint32_t face_bit[3] = {lx == 0 ? 1u : (lx == 3 ? 2u : 0u), ly == 0 ? 4u : (ly == 3 ? 8u : 0u),
                                lz == 0 ? 16u : (lz == 3 ? 32u : 0u)};
  // The face neighbour this lane needs on each axis (its index among the six), and the cell of
  // that block the face touches (the coordinate on the axis 0 <-> 3).
  const uint32_t nb_dir[3] = {lx == 3 ? 1u : 0u, ly == 3 ? 3u : 2u, lz == 3 ? 5u : 4u};
  const uint32_t nb_cell[3] = {ln ^ 48u, ln ^ 12u, ln ^ 3u};
  __shared__ uint32_t nbl[SOLVE_WAVES][6][SOLVE_JMAX];
  __shared__ uint8_t mk[SOLVE_WAVES][SOLVE_JMAX][64];
  // Each pass reads its wave's first block from LDS, so that the blocks' addresses are formed
  // inside the iteration rather than held in registers across it.
  __shared__ uint32_t wave_b0[SOLVE_WAVES];
  if (stream) {
    for (uint32_t j = 0; j < SOLVE_JMAX; ++j) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * j;
      if (b >= b1) break;
      mk[wv][j][ln] = (uint8_t)(d.nbm[b * 64u + ln] & 127u);
      if (ln < 6u) {
        // block_nb's index of face neighbour ln: 13 -+ 9, 3 or 1.
        const uint32_t a = ln >> 1, step = a == 0 ? 9u : (a == 1 ? 3u : 1u);
        nbl[wv][ln][j] = d.block_nb[b * 27u + ((ln & 1u) ? 13u + step : 13u - step)];
      }
    }
    if (ln == 0u) wave_b0[wv] = b0 + wv;
    // Written and read by the same wave only.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
  }
  // The neighbour mask of this lane's cell in the wave's block j; w as in halo_of.
  auto mask_of = [&](uint32_t w, uint32_t j) -> uint32_t { return mk[w][j][ln]; };
  // The cell across face `a` of this lane in the wave's block j (used only where the mask
  // has the face's bit, so the neighbour block exists); w is the wave's index as a pass read
  // it from LDS.
  auto halo_of = [&](uint32_t w, uint32_t j, int a) -> uint32_t { return nbl[w][nb_dir[a]][j] * 64u + nb_cell[a]; };
  // Waits until word `w` (reloaded from p) carries `want`; false once the solve aborts.
  auto wait_word = [&](unsigned long long& w, const unsigned long long* p, uint32_t want) -> bool {
    uint32_t spins = 0;
    while ((uint32_t)(w >> 32) != want) {
      __builtin_amdgcn_s_sleep(1);
      w = load_word(p);
      if ((++spins & 255u) == 0u &&
          (__hip_atomic_load(d.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kSpinLimit)) {
        __hip_atomic_store(d.bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
    }
    return true;
  };
  // The tag of the last iteration's first reduction: its directions' words carry it.
  uint32_t tag_d = 0u;
  bool ok = true;

  while (go) {
    // solve_q (fluid.wgsl:508-534): the new direction and q = A d; neighbours' directions are
    // rebuilt from z and the old direction.
    const uint32_t old = s.iters % 2u;
    const float* cDo = d.cf + (size_t)(NATRAY_F_D + old) * capc;
    float* cDn = d.cf + (size_t)(NATRAY_F_D + 1u - old) * capc;
    const float fb = (float)s.beta;
    if (stream) {
      const uint32_t wb = wave_b0[wv], wl = wb - b0;
      // Directions of the neighbours: the last iteration's words, or before the first
      // iteration the direction field as the solve found it.
      const bool first = s.iters == 0u;
      const unsigned long long* const dw_old = zw + (size_t)capc * (1u + (s.iters + 1u) % 2u);
      unsigned long long* const dw_new = zw + (size_t)capc * (1u + s.iters % 2u);
      // This wave's blocks; a rolled loop, so that nothing per block lives across it.
      const uint32_t nbw = b1 > wb ? (b1 - wb + SOLVE_WAVES - 1u) / SOLVE_WAVES : 0u;
#pragma unroll 1
      for (uint32_t j0 = 0; j0 < nbw; j0 += SOLVE_GROUP) {
        float z[SOLVE_GROUP], dold[SOLVE_GROUP], dg[SOLVE_GROUP];
        unsigned long long hzw[SOLVE_GROUP][3], hdw[SOLVE_GROUP][3];
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          z[u] = dold[u] = 0.0f;
          dg[u] = 1.0f;
#pragma unroll
          for (int a = 0; a < 3; ++a) hzw[u][a] = hdw[u][a] = 0ull;
          const uint32_t mask = mask_of(wl, j);
          if (b >= b1 || !(mask & 64u)) continue;
          const uint32_t cell = b * 64u + ln;
          z[u] = ld_cell(cZ, cell);
          dold[u] = ld_cell(cDo, cell);
          dg[u] = ld_cell(cDiag, cell);
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            if (!(mask & face_bit[a])) continue;
            const uint32_t o = halo_of(wl, j, a);
            hzw[u][a] = load_word(&zw[o]);
            hdw[u][a] = first ? tag_word(tag_d, __float_as_uint(ld_cell(cDo, o))) : load_word(&dw_old[o]);
          }
        }
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          if (b >= b1) continue;
          const uint32_t cell = b * 64u + ln, mask = mask_of(wl, j);
          float hd[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            if (!(mask & 64u) || !(mask & face_bit[a])) continue;
            const uint32_t o = halo_of(wl, j, a);
            ok = ok && wait_word(hzw[u][a], &zw[o], tag) && wait_word(hdw[u][a], &dw_old[o], tag_d);
            hd[a] = __uint_as_float((uint32_t)hzw[u][a]) + fb * __uint_as_float((uint32_t)hdw[u][a]);
          }
          const float dn = z[u] + fb * dold[u];
          const Faces f = block_faces(dn);
          R v = (R)0;
          if (mask & 64u) {
            float q = dg[u] * dn;
            if (mask & 1u) q -= lx > 0 ? f.xm : hd[0];
            if (mask & 2u) q -= lx < 3 ? f.xp : hd[0];
            if (mask & 4u) q -= ly > 0 ? f.ym : hd[1];
            if (mask & 8u) q -= ly < 3 ? f.yp : hd[1];
            if (mask & 16u) q -= lz > 0 ? f.zm : hd[2];
            if (mask & 32u) q -= lz < 3 ? f.zp : hd[2];
            st_cell(cDn, cell, dn);
            st_cell(cQ, cell, q);
            store_word(&dw_new[cell], tag_word(tag + 1u, __float_as_uint(dn)));
            v = prod<R>(dn, q);
          }
          v = wave_tree<R>(v);
          if (ln == 0) d.part[b * 4u] = (double)v;
        }
      }
    } else {
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
        const uint32_t cell = b * 64u + lane;
        const uint32_t mask = d.nbm[cell];
        R v = (R)0;
        if (mask & 64u) {
          const float dn = cZ[cell] + fb * cDo[cell];
          float q = cDiag[cell] * dn;
          if (mask & 1u) { uint32_t o = nbr(lk, lx, ly, lz, -1, 0, 0); q -= cZ[o] + fb * cDo[o]; }
          if (mask & 2u) { uint32_t o = nbr(lk, lx, ly, lz, 1, 0, 0); q -= cZ[o] + fb * cDo[o]; }
          if (mask & 4u) { uint32_t o = nbr(lk, lx, ly, lz, 0, -1, 0); q -= cZ[o] + fb * cDo[o]; }
          if (mask & 8u) { uint32_t o = nbr(lk, lx, ly, lz, 0, 1, 0); q -= cZ[o] + fb * cDo[o]; }
          if (mask & 16u) { uint32_t o = nbr(lk, lx, ly, lz, 0, 0, -1); q -= cZ[o] + fb * cDo[o]; }
          if (mask & 32u) { uint32_t o = nbr(lk, lx, ly, lz, 0, 0, 1); q -= cZ[o] + fb * cDo[o]; }
          cDn[cell] = dn;
          cQ[cell] = q;
          v = prod<R>(dn, q);
        }
        v = wave_tree<R>(v);
        if (lane == 0) d.part[b * 4u] = (double)v;
      });
    }
    if (!reduce<R>(d, sp, ++tag, 1, nullptr, tt) || !ok) break;
    tag_d = tag;
    // solve_alpha (fluid.wgsl:536-541)
    s.alpha = tt[0] > (R)0 ? s.rz / tt[0] : (R)0;
    const float fa = (float)s.alpha;
    const bool refresh = refresh_on && (s.iters + 1u) % d.sc->refresh_every == 0u;
    // solve_update (fluid.wgsl:543-559): p += alpha d, r -= alpha q, z = M^-1 r.
    unsigned long long* const pw = zw + (size_t)capc * (1u + (s.iters + 1u) % 2u);
    if (stream) {
      const uint32_t wb = wave_b0[wv], wl = wb - b0;
      // This wave's blocks; a rolled loop, so that nothing per block lives across it.
      const uint32_t nbw = b1 > wb ? (b1 - wb + SOLVE_WAVES - 1u) / SOLVE_WAVES : 0u;
#pragma unroll 1
      for (uint32_t j0 = 0; j0 < nbw; j0 += SOLVE_GROUP) {
        float p[SOLVE_GROUP], dn[SOLVE_GROUP], rr[SOLVE_GROUP], q[SOLVE_GROUP], dg[SOLVE_GROUP];
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          p[u] = dn[u] = rr[u] = q[u] = 0.0f;
          dg[u] = 1.0f;
          if (b >= b1 || !(mask_of(wl, j) & 64u)) continue;
          const uint32_t cell = b * 64u + ln;
          p[u] = ld_cell(cP, cell);
          dn[u] = ld_cell(cDn, cell);
          rr[u] = ld_cell(cR, cell);
          q[u] = ld_cell(cQ, cell);
          dg[u] = ld_cell(cDiag, cell);
        }
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          if (b >= b1) continue;
          const uint32_t cell = b * 64u + ln;
          const bool water = (mask_of(wl, j) & 64u) != 0u;
          float r = 0.0f, diag = 1.0f;
          if (water) {
            const float pn = p[u] + fa * dn[u];
            st_cell(cP, cell, pn);
            // Before a refresh, p for the neighbours' A p, in the direction buffer the last
            // iteration used (every neighbour has read it by now).
            if (refresh) store_word(&pw[cell], tag_word(tag + 1u, __float_as_uint(pn)));
            r = rr[u] - fa * q[u];
            st_cell(cR, cell, r);
            diag = dg[u];
          }
          if (!refresh) {
            const float z = precondition(r, diag, water, lx, ly, lz, mode);
            if (water) {
              st_cell(cZ, cell, z);
              store_word(&zw[cell], tag_word(tag + 1u, __float_as_uint(z)));
            }
            const R v0 = wave_tree<R>(prod<R>(r, z)), v1 = wave_tree<R>(prod<R>(r, r));
            if (ln == 0) {
              d.part[b * 4u + 0] = (double)v0;
              d.part[b * 4u + 1] = (double)v1;
            }
          }
        }
      }
    } else {
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t*) {
        const uint32_t cell = b * 64u + lane;
        const bool water = (d.nbm[cell] & 64u) != 0u;
        float r = 0.0f, diag = 1.0f;
        if (water) {
          cP[cell] += fa * cDn[cell];
          r = cR[cell] - fa * cQ[cell];
          cR[cell] = r;
          diag = cDiag[cell];
        }
        if (!refresh) {
          const float z = precondition(r, diag, water, lx, ly, lz, mode);
          if (water) cZ[cell] = z;
          const R v0 = wave_tree<R>(prod<R>(r, z)), v1 = wave_tree<R>(prod<R>(r, r));
          if (lane == 0) {
            d.part[b * 4u + 0] = (double)v0;
            d.part[b * 4u + 1] = (double)v1;
          }
        }
      });
    }
    if (refresh && stream) {
      // True residual r = b - A p in double once every p is updated, then the preconditioner;
      // the neighbours' p across block faces from their words.
      if (!reduce<R>(d, sp, ++tag, 0, nullptr, tt)) break;
      const uint32_t wb = wave_b0[wv], wl = wb - b0;
      // This wave's blocks; a rolled loop, so that nothing per block lives across it.
      const uint32_t nbw = b1 > wb ? (b1 - wb + SOLVE_WAVES - 1u) / SOLVE_WAVES : 0u;
#pragma unroll 1
      for (uint32_t j0 = 0; j0 < nbw; j0 += SOLVE_GROUP) {
        float p[SOLVE_GROUP], dg[SOLVE_GROUP], bv[SOLVE_GROUP];
        unsigned long long hpw[SOLVE_GROUP][3];
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          p[u] = bv[u] = 0.0f;
          dg[u] = 1.0f;
#pragma unroll
          for (int a = 0; a < 3; ++a) hpw[u][a] = 0ull;
          const uint32_t mask = mask_of(wl, j);
          if (b >= b1 || !(mask & 64u)) continue;
          const uint32_t cell = b * 64u + ln;
          p[u] = ld_cell(cP, cell);
          dg[u] = ld_cell(cDiag, cell);
          bv[u] = ld_cell(d.bvec, cell);
#pragma unroll
          for (int a = 0; a < 3; ++a)
            if (mask & face_bit[a]) hpw[u][a] = load_word(&pw[halo_of(wl, j, a)]);
        }
#pragma unroll
        for (uint32_t u = 0; u < SOLVE_GROUP; ++u) {
          const uint32_t j = j0 + u, b = wb + SOLVE_WAVES * j;
          if (b >= b1) continue;
          const uint32_t cell = b * 64u + ln, mask = mask_of(wl, j);
          const bool water = (mask & 64u) != 0u;
          float hp[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            if (!water || !(mask & face_bit[a])) continue;
            ok = ok && wait_word(hpw[u][a], &pw[halo_of(wl, j, a)], tag);
            hp[a] = __uint_as_float((uint32_t)hpw[u][a]);
          }
          const Faces f = block_faces(p[u]);
          float r = 0.0f, diag = 1.0f;
          if (water) {
            diag = dg[u];
            double ap = (double)diag * (double)p[u];
            if (mask & 1u) ap -= (double)(lx > 0 ? f.xm : hp[0]);
            if (mask & 2u) ap -= (double)(lx < 3 ? f.xp : hp[0]);
            if (mask & 4u) ap -= (double)(ly > 0 ? f.ym : hp[1]);
            if (mask & 8u) ap -= (double)(ly < 3 ? f.yp : hp[1]);
            if (mask & 16u) ap -= (double)(lz > 0 ? f.zm : hp[2]);
            if (mask & 32u) ap -= (double)(lz < 3 ? f.zp : hp[2]);
            r = (float)((double)bv[u] - ap);
            st_cell(cR, cell, r);
          }
          const float z = precondition(r, diag, water, lx, ly, lz, mode);
          if (water) {
            st_cell(cZ, cell, z);
            store_word(&zw[cell], tag_word(tag + 1u, __float_as_uint(z)));
          }
          const R v0 = wave_tree<R>(prod<R>(r, z)), v1 = wave_tree<R>(prod<R>(r, r));
          if (ln == 0) {
            d.part[b * 4u + 0] = (double)v0;
            d.part[b * 4u + 1] = (double)v1;
          }
        }
      }
    } else if (refresh) {
      // True residual r = b - A p in double once every p is updated, then the preconditioner.
      if (!reduce<R>(d, sp, ++tag, 0, nullptr, tt)) break;
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
        const uint32_t cell = b * 64u + lane;
        const uint32_t mask = d.nbm[cell];
        const bool water = (mask & 64u) != 0u;
        float r = 0.0f, diag = 1.0f;
        if (water) {
          diag = cDiag[cell];
          double ap = (double)diag * (double)cP[cell];
          if (mask & 1u) ap -= (double)cP[nbr(lk, lx, ly, lz, -1, 0, 0)];
          if (mask & 2u) ap -= (double)cP[nbr(lk, lx, ly, lz, 1, 0, 0)];
          if (mask & 4u) ap -= (double)cP[nbr(lk, lx, ly, lz, 0, -1, 0)];
          if (mask & 8u) ap -= (double)cP[nbr(lk, lx, ly, lz, 0, 1, 0)];
          if (mask & 16u) ap -= (double)cP[nbr(lk, lx, ly, lz, 0, 0, -1)];
          if (mask & 32u) ap -= (double)cP[nbr(lk, lx, ly, lz, 0, 0, 1)];
          r = (float)((double)d.bvec[cell] - ap);
          cR[cell] = r;
        }
        const float z = precondition(r, diag, water, lx, ly, lz, mode);
        if (water) cZ[cell] = z;
        const R v0 = wave_tree<R>(prod<R>(r, z)), v1 = wave_tree<R>(prod<R>(r, r));
        if (lane == 0) {
          d.part[b * 4u + 0] = (double)v0;
          d.part[b * 4u + 1] = (double)v1;
        }
      });
    }
    if (!reduce<R>(d, sp, ++tag, 2, nullptr, tt)) break;
    go = solve_beta<R>(d, tt, s);
  }
  solve_finish<R>(d, sp, tag, s);
}

// This workgroup's chunks and blocks in its scene's solve, and whether every workgroup keeps
// its blocks in registers.
__device__ __forceinline__ SolveSpan solve_span(const View& v, bool streams) {
  SolveSpan sp;
  sp.n = n_blocks(v.d);
  sp.K = sp.n ? (sp.n + MAX_CHUNKS - 1u) / MAX_CHUNKS : 1u;
  sp.m = sp.n ? (sp.n + sp.K - 1u) / sp.K : 0u;
  sp.G = v.wgs;
  sp.wg = v.wg;
  sp.c0 = (uint32_t)((uint64_t)v.wg * sp.m / sp.G);
  sp.c1 = (uint32_t)((uint64_t)(v.wg + 1) * sp.m / sp.G);
  sp.b0 = sp.c0 * sp.K;
  sp.b1 = min(sp.n, sp.c1 * sp.K);
  // The most blocks a workgroup holds: its chunks are at most ceil(m / G).
  const uint32_t most = (sp.m + sp.G - 1u) / sp.G * sp.K;
  sp.in_regs = most <= SOLVE_WAVES * SOLVE_RMAX;
  sp.stream = streams && !sp.in_regs && most <= SOLVE_WAVES * SOLVE_JMAX;
  sp.coherent = !sp.in_regs && !sp.stream;
  return sp;
}

template <bool MIXED, class L>
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  using R = typename std::conditional<MIXED, double, float>::type;
  constexpr uint32_t RM = SOLVE_RMAX;
  if (d.st->dt_ns == 0u) return;
  // A test of the plain launch's failure path: the last workgroup stays out, as one that is not
  // resident would, and the others' bounded waits abort the solve.
  if constexpr (std::is_same<L, One>::value) {
    if (d.solve_plain == 2u && v.wg + 1u == v.wgs) return;
  }
  // An earlier solve of this batch lost its barrier: the host reports it, nothing more runs.
  if (__hip_atomic_load(d.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
  const SolveSpan sp = solve_span(v, true);
  uint32_t tag = __hip_atomic_load(d.bar + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!sp.in_regs) {
    // Called with the device copy of Dev, which the callee reads with scalar loads.
    solve_memory<MIXED>(d.self, sp, tag);
    return;
  }

  // Register form. Wave wv holds blocks b0 + wv + SOLVE_WAVES r (r < RM), the blocks
  // for_blocks gives it. Per block and lane: p, r, z, the direction d, q, diag, the neighbour
  // mask, and for each axis the halo cell across the block face (when it is water) with its
  // direction and p as this wave rebuilds them, and its tagged z as last loaded.
  const uint32_t mode = d.sc->precond;
  const uint32_t b0 = sp.b0, b1 = sp.b1;
  const uint32_t capc = d.capc;
  float* const cP = d.cf + (size_t)NATRAY_F_P * capc;
  const float* const p0 = start_pressure(d);
  unsigned long long* const hz = d.hz;
  __shared__ LPart<R> lp;
  for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    R v[4];
    init_cell<R>(d, b, lane, lx, ly, lz, lk, d.hz, tag + 1u, v, p0);
    if (lane == 0)
      for (uint32_t k = 0; k < 4u; ++k) lp.v[k][b - b0] = v[k];
  });
  R tt[4];
  SolveScalars<R> s;
  bool go = solve_start<R>(d, tt, reduce<R>(d, sp, ++tag, 4, &lp, tt), s);
  const bool refresh_on = MIXED && d.sc->precision == NATRAY_CG_MIXED && d.sc->refresh_every > 0u;

  __shared__ uint32_t lkr[SOLVE_WAVES][SOLVE_RMAX][27];
  const uint32_t wv = threadIdx.x >> 6, ln = lane_id();
  const int lx = (int)(ln >> 4), ly = (int)((ln >> 2) & 3u), lz = (int)(ln & 3u);
  // Mask bit of the neighbour across the block face on each axis, if this lane is on one.
  const uint32_t face_bit[3] = {lx == 0 ? 1u : (lx == 3 ? 2u : 0u), ly == 0 ? 4u : (ly == 3 ? 8u : 0u),
                                lz == 0 ? 16u : (lz == 3 ? 32u : 0u)};
  float rP[RM], rR[RM], rZ[RM], rD[RM], rQ[RM], rDg[RM];
  uint32_t rM[RM];
  float hD[RM][3], hPr[RM][3];
  unsigned long long hW[RM][3];
  auto halo_cell = [&](uint32_t r, int a) -> uint32_t {
    const int sg = (face_bit[a] & 0x2Au) ? 1 : -1;
    return nbr(lkr[wv][r], lx, ly, lz, a == 0 ? sg : 0, a == 1 ? sg : 0, a == 2 ? sg : 0);
  };
  auto has_halo = [&](uint32_t r, int a) -> bool { return (rM[r] & face_bit[a]) != 0u; };
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b < b1 && ln < 27) lkr[wv][r][ln] = d.block_nb[b * 27u + ln];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    rP[r] = rR[r] = rZ[r] = rD[r] = rQ[r] = 0.0f;
    rDg[r] = 1.0f;
    rM[r] = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      hD[r][a] = hPr[r][a] = 0.0f;
      hW[r][a] = 0ull;
    }
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b < b1) {
      const uint32_t cell = b * 64u + ln;
      rM[r] = d.nbm[cell];
      rP[r] = p0[cell];
      rR[r] = d.cf[(size_t)NATRAY_F_R * capc + cell];
      rZ[r] = d.cf[(size_t)NATRAY_F_Z * capc + cell];
      rDg[r] = d.cf[(size_t)NATRAY_F_DIAG * capc + cell];
      rD[r] = d.cf[(size_t)NATRAY_F_D * capc + cell];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!has_halo(r, a)) continue;
        const uint32_t o = halo_cell(r, a);
        hD[r][a] = d.cf[(size_t)NATRAY_F_D * capc + o];
        if (refresh_on) hPr[r][a] = p0[o];
        hW[r][a] = load_word(&hz[o]);
      }
    }
  }

  bool ok = true;
  while (go) {
    // solve_q (fluid.wgsl:508-534): the new direction and q = A d.
    float* cDn = d.cf + (size_t)(NATRAY_F_D + 1u - s.iters % 2u) * capc;
    const float fb = (float)s.beta;
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      const uint32_t cell = b * 64u + ln, mask = rM[r];
      // Halo directions: z across the face (tagged with the last reduction's tag), rebuilt
      // as the owner does, z + beta d_old.
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!has_halo(r, a)) continue;
        uint32_t spins = 0;
        while ((uint32_t)(hW[r][a] >> 32) != tag) {
          __builtin_amdgcn_s_sleep(1);
          hW[r][a] = load_word(&hz[halo_cell(r, a)]);
          if ((++spins & 255u) == 0u &&
              (__hip_atomic_load(d.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kSpinLimit)) {
            __hip_atomic_store(d.bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = false;
            break;
          }
        }
        hD[r][a] = __uint_as_float((uint32_t)hW[r][a]) + fb * hD[r][a];
      }
      const float dn = rZ[r] + fb * rD[r];
      const Faces f = block_faces(dn);
      R v = (R)0;
      if (mask & 64u) {
        float q = rDg[r] * dn;
        if (mask & 1u) q -= lx > 0 ? f.xm : hD[r][0];
        if (mask & 2u) q -= lx < 3 ? f.xp : hD[r][0];
        if (mask & 4u) q -= ly > 0 ? f.ym : hD[r][1];
        if (mask & 8u) q -= ly < 3 ? f.yp : hD[r][1];
        if (mask & 16u) q -= lz > 0 ? f.zm : hD[r][2];
        if (mask & 32u) q -= lz < 3 ? f.zp : hD[r][2];
        cDn[cell] = dn;
        rD[r] = dn;
        rQ[r] = q;
        v = prod<R>(dn, q);
      }
      v = wave_tree<R>(v);
      if (ln == 0) lp.v[0][b - b0] = v;
    }
    if (!reduce<R>(d, sp, ++tag, 1, &lp, tt) || !ok) break;
    // solve_alpha (fluid.wgsl:536-541)
    s.alpha = tt[0] > (R)0 ? s.rz / tt[0] : (R)0;
    const float fa = (float)s.alpha;
    const bool refresh = refresh_on && (s.iters + 1u) % d.sc->refresh_every == 0u;
    // solve_update (fluid.wgsl:543-559): p += alpha d, r -= alpha q, z = M^-1 r; with the
    // refresh, r = b - A p in double instead, p of the halo rebuilt as its owner updates it.
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      const uint32_t cell = b * 64u + ln, mask = rM[r];
      const bool water = (mask & 64u) != 0u;
      float rr = 0.0f, diag = 1.0f;
      if (water) {
        rP[r] += fa * rD[r];
        rR[r] = rR[r] - fa * rQ[r];
        rr = rR[r];
        diag = rDg[r];
      }
      if (refresh_on) {
#pragma unroll
        for (int a = 0; a < 3; ++a)
          if (has_halo(r, a)) hPr[r][a] += fa * hD[r][a];
      }
      if (refresh) {
        const float p = rP[r];
        const Faces f = block_faces(p);
        if (water) {
          double ap = (double)diag * (double)p;
          if (mask & 1u) ap -= (double)(lx > 0 ? f.xm : hPr[r][0]);
          if (mask & 2u) ap -= (double)(lx < 3 ? f.xp : hPr[r][0]);
          if (mask & 4u) ap -= (double)(ly > 0 ? f.ym : hPr[r][1]);
          if (mask & 8u) ap -= (double)(ly < 3 ? f.yp : hPr[r][1]);
          if (mask & 16u) ap -= (double)(lz > 0 ? f.zm : hPr[r][2]);
          if (mask & 32u) ap -= (double)(lz < 3 ? f.zp : hPr[r][2]);
          rr = (float)((double)d.bvec[cell] - ap);
          rR[r] = rr;
        }
      }
      const float z = precondition(rr, diag, water, lx, ly, lz, mode);
      if (water) {
        rZ[r] = z;
        store_word(&hz[cell], tag_word(tag + 1u, __float_as_uint(z)));
      }
      const R v0 = wave_tree<R>(prod<R>(rr, z)), v1 = wave_tree<R>(prod<R>(rr, rr));
      if (ln == 0) {
        lp.v[0][b - b0] = v0;
        lp.v[1][b - b0] = v1;
      }
    }
    // Halo z for the next solve_q: neighbours publish it before their part of the reduction
    // below, so every wave reloads its stale halo words while the reduction completes.
    auto halo_poll = [&](uint32_t want) -> bool {
      bool fresh = true;
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        if (b0 + wv + SOLVE_WAVES * r >= b1) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (!has_halo(r, a) || (uint32_t)(hW[r][a] >> 32) == want) continue;
          hW[r][a] = load_word(&hz[halo_cell(r, a)]);
          fresh = fresh && (uint32_t)(hW[r][a] >> 32) == want;
        }
      }
      return fresh;
    };
    if (!reduce<R>(d, sp, ++tag, 2, &lp, tt, halo_poll)) break;
    go = solve_beta<R>(d, tt, s);
  }
  // The fields the other kernels and the host read: p, r, z and q of this wave's water cells
  // (solve_q stored the directions).
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b < b1 && (rM[r] & 64u)) {
      const uint32_t cell = b * 64u + ln;
      cP[cell] = rP[r];
      d.cf[(size_t)NATRAY_F_R * capc + cell] = rR[r];
      d.cf[(size_t)NATRAY_F_Z * capc + cell] = rZ[r];
      if (s.iters > 0u) d.cf[(size_t)NATRAY_F_Q * capc + cell] = rQ[r];
    }
  }
  solve_finish<R>(d, sp, tag, s);
}

// ---- pipelined CG (NATRAY_CG_PIPELINED; not in the reference) ---------------------------------------
//
// Ghysels and Vanroose's pipelined preconditioned CG (Parallel Computing 40, 2014, Algorithm 3):
// the Krylov iterates of the solve above in exact arithmetic, with ONE reduction per iteration,
// of (r, u), (w, u) and (r, r), where u = M^-1 r and w = A u are carried by recurrences. Per
// iteration: m = M^-1 w, n = A m, then
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
//   z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p,
//   x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z
// with gamma = (r, u), delta = (w, u). Every refresh_every iterations the residual is replaced
// (Cools et al., SIAM J. Sci. Comput. 40, 2018): r = b - A x in double, u = M^-1 r, w = A u,
// s = A p, q = M^-1 s, z = A q. Dots, norms and the tolerance test are in double, vectors in
// f32. The iteration stops when (r, r) meets the tolerance (tested from the second reduction
// on; the first one's (r, r) is the start's) or after limits.z updates. When the solve ends,
// x is in P, r in R, u in Z, s in Q, p in D and w in D + 1.
//
// A replacement brings the recurrences' r, u, w, s, q and z back to what x and p give, which
// upsets the recurrences when r has already fallen to the rounding gap between the recurred
// and the true residual: so a replacement is made only while |r| > PIPE_REPLACE tol max(|b|, 1)
// (the tolerance test's right-hand side times PIPE_REPLACE), where the gap is still small
// against r; the gap that builds up afterwards starts from residuals that small.
//
// Register form: m of the water cells on block faces is published as tagged words and read by
// the neighbours while the iteration's reduction completes, so the halo exchange costs no wait
// of its own. A replacement exchanges x, u, p and q in turn. Halo words are numbered 8 t + k
// (t the iteration's reduction tag; k = 0 for m, 1..4 for x, u, p, q of a replacement; u0 at
// the start has 8 t + 5 with t the start's reduction) and alternate between two word buffers:
// m of reduction t and u and q of its replacement in buffer t & 1, x and p in the other, u0
// in the start's. A wave rewrites a buffer only after a reduction or an exchange that every
// neighbour joins after it has read the buffer's last words.

// The pipelined iteration's scalars from the reduction's (r, u), (w, u), (r, r); returns
// whether to update (false once (r, r) meets the tolerance or limits.z updates ran).
__device__ __forceinline__ bool pipe_scalars(const Dev& d, const double* tt, SolveScalars<double>& s) {
  const double g = tt[0], dl = tt[1], rr = tt[2];
  s.r_norm = sq(rr);
  if (s.iters > 0u) {
    const bool done = converged<double>(rr, (double)d.U->solve[0], s.b_norm);
    if (done || s.iters >= min(d.U->limits[2], SOLVE_ITER_CAP)) {
      s.iters_max = s.iters > s.iters_max ? s.iters : s.iters_max;
      s.unconverged += done ? 0u : 1u;
      s.rz = g;
      return false;
    }
  }
  const double beta = (s.iters > 0u && s.rz > 0.0) ? g / s.rz : 0.0;
  const double den = (s.iters > 0u && s.alpha != 0.0) ? dl - beta * g / s.alpha : dl;
  s.alpha = den > 0.0 ? g / den : 0.0;
  s.beta = beta;
  s.rz = g;
  return true;
}

// (A v) at this lane's cell in register form: diag v minus the water neighbours' v in the order
// -x, +x, -y, +y, -z, +z, those across a block face from h (one per axis). Every lane of the
// wave calls it (block_faces is a wave operation); the result is meaningful on water cells.
__device__ __forceinline__ float apply_row(float v, float diag, uint32_t mask, const float h[3], int lx, int ly, int lz) {
  const Faces f = block_faces(v);
  float q = diag * v;
  if (mask & 1u) q -= lx > 0 ? f.xm : h[0];
  if (mask & 2u) q -= lx < 3 ? f.xp : h[0];
  if (mask & 4u) q -= ly > 0 ? f.ym : h[1];
  if (mask & 8u) q -= ly < 3 ? f.yp : h[1];
  if (mask & 16u) q -= lz > 0 ? f.zm : h[2];
  if (mask & 32u) q -= lz < 3 ? f.zp : h[2];
  return q;
}

// The same row on a field in memory.
__device__ __forceinline__ float mem_row(const float* v, float diag, uint32_t mask, uint32_t cell, const uint32_t* lk, int lx,
                                         int ly, int lz) {
  float q = diag * v[cell];
  if (mask & 1u) q -= v[nbr(lk, lx, ly, lz, -1, 0, 0)];
  if (mask & 2u) q -= v[nbr(lk, lx, ly, lz, 1, 0, 0)];
  if (mask & 4u) q -= v[nbr(lk, lx, ly, lz, 0, -1, 0)];
  if (mask & 8u) q -= v[nbr(lk, lx, ly, lz, 0, 1, 0)];
  if (mask & 16u) q -= v[nbr(lk, lx, ly, lz, 0, 0, -1)];
  if (mask & 32u) q -= v[nbr(lk, lx, ly, lz, 0, 0, 1)];
  return q;
}

// Waits until the word at p carries `want` (w holds the last load). Gives up like the
// reductions' waits: returns false, with the abort flag raised, after kSpinLimit polls or
// once another wave has raised it.
__device__ __forceinline__ bool wait_word(const Dev& d, unsigned long long& w, const unsigned long long* p, uint32_t want) {
  uint32_t spins = 0;
  while ((uint32_t)(w >> 32) != want) {
    __builtin_amdgcn_s_sleep(1);
    w = load_word(p);
    if ((++spins & 255u) == 0u &&
        (__hip_atomic_load(d.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || spins > kSpinLimit)) {
      __hip_atomic_store(d.bar + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
  }
  return true;
}

// The block tail of the memory form: m = M^-1 w into mo and the block partials (r, u), (w, u),
// (r, r) of the values this lane holds (0 off water).
__device__ __forceinline__ void pipe_tail_memory(const Dev& d, uint32_t b, uint32_t lane, int lx, int ly, int lz, float r,
                                                 float u, float w, float diag, bool water, float* mo) {
  const float m = precondition(w, diag, water, lx, ly, lz, d.sc->precond);
  if (water) mo[b * 64u + lane] = m;
  const double v0 = wave_tree<double>(prod<double>(r, u)), v1 = wave_tree<double>(prod<double>(w, u)),
               v2 = wave_tree<double>(prod<double>(r, r));
  if (lane == 0) {
    d.part[b * 4u + 0] = v0;
    d.part[b * 4u + 1] = v1;
    d.part[b * 4u + 2] = v2;
  }
}

// The memory form of the pipelined solve: every field in memory (z, q and two m buffers in
// Dev::pcg), one reduction per iteration and three barriers more per replacement.
__device__ __noinline__ void solve_pipe_memory(const Dev* dp, SolveSpan sp, uint32_t tag) {
  const Dev& d = *dp;
  const uint32_t mode = d.sc->precond;
  const uint32_t b0 = sp.b0, b1 = sp.b1;
  const uint32_t capc = d.capc;
  float* const cX = d.cf + (size_t)NATRAY_F_P * capc;
  float* const cR = d.cf + (size_t)NATRAY_F_R * capc;
  float* const cU = d.cf + (size_t)NATRAY_F_Z * capc;
  float* const cS = d.cf + (size_t)NATRAY_F_Q * capc;
  float* const cP = d.cf + (size_t)NATRAY_F_D * capc;
  float* const cW = d.cf + (size_t)(NATRAY_F_D + 1u) * capc;
  const float* const cDiag = d.cf + (size_t)NATRAY_F_DIAG * capc;
  float* const cZ = d.pcg;
  float* const cQ = d.pcg + capc;
  float* const cM[2] = {d.pcg + 2u * capc, d.pcg + 3u * capc};
  const float* const p0 = start_pressure(d);
  for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    init_cell<double>(d, b, lane, lx, ly, lz, lk, nullptr, 0u, d.part + b * 4u, p0);
  });
  double tt[4];
  SolveScalars<double> s;
  bool go = solve_start<double>(d, tt, reduce<double>(d, sp, ++tag, 4, nullptr, tt), s);
  if (p0 != cX) {
    // The iterate starts at the guess; every workgroup has read the carried pressure by now.
    for (uint32_t c = b0 * 64u + threadIdx.x; c < b1 * 64u; c += blockDim.x)
      if (d.nbm[c] & 64u) cX[c] = p0[c];
    __syncthreads();
  }
  const bool refresh_on = d.sc->refresh_every > 0u;
  const double replace_above = PIPE_REPLACE * (double)d.U->solve[0] * (s.b_norm < 1.0 ? 1.0 : s.b_norm);
  if (go) {
    // w = A u; z, q, s and p start at zero; m = M^-1 w and the first partials.
    for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
      const uint32_t cell = b * 64u + lane, mask = d.nbm[cell];
      const bool water = (mask & 64u) != 0u;
      float r = 0.0f, u = 0.0f, w = 0.0f, diag = 1.0f;
      if (water) {
        diag = cDiag[cell];
        r = cR[cell];
        u = cU[cell];
        w = mem_row(cU, diag, mask, cell, lk, lx, ly, lz);
        cW[cell] = w;
        cZ[cell] = cQ[cell] = cS[cell] = cP[cell] = 0.0f;
      }
      pipe_tail_memory(d, b, lane, lx, ly, lz, r, u, w, diag, water, cM[0]);
    });
  }
  while (go) {
    if (!reduce<double>(d, sp, ++tag, 3, nullptr, tt)) break;
    if (!pipe_scalars(d, tt, s)) break;
    const float fa = (float)s.alpha, fb = (float)s.beta;
    const float* const mi = cM[s.iters & 1u];
    float* const mo = cM[(s.iters + 1u) & 1u];
    const bool refresh = refresh_on && (s.iters + 1u) % d.sc->refresh_every == 0u && s.r_norm > replace_above;
    for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
      const uint32_t cell = b * 64u + lane, mask = d.nbm[cell];
      const bool water = (mask & 64u) != 0u;
      float r = 0.0f, u = 0.0f, w = 0.0f, diag = 1.0f;
      if (water) {
        diag = cDiag[cell];
        const float n = mem_row(mi, diag, mask, cell, lk, lx, ly, lz);
        const float z = n + fb * cZ[cell];
        const float q = mi[cell] + fb * cQ[cell];
        const float sv = cW[cell] + fb * cS[cell];
        const float p = cU[cell] + fb * cP[cell];
        cX[cell] = cX[cell] + fa * p;
        r = cR[cell] - fa * sv;
        u = cU[cell] - fa * q;
        w = cW[cell] - fa * z;
        cZ[cell] = z;
        cQ[cell] = q;
        cS[cell] = sv;
        cP[cell] = p;
        cR[cell] = r;
        cU[cell] = u;
        cW[cell] = w;
      }
      if (!refresh) pipe_tail_memory(d, b, lane, lx, ly, lz, r, u, w, diag, water, mo);
    });
    s.iters += 1u;
    if (refresh) {
      // r = b - A x in double, u = M^-1 r.
      if (!reduce<double>(d, sp, ++tag, 0, nullptr, tt)) break;
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
        const uint32_t cell = b * 64u + lane, mask = d.nbm[cell];
        const bool water = (mask & 64u) != 0u;
        float r = 0.0f, diag = 1.0f;
        if (water) {
          diag = cDiag[cell];
          double ap = (double)diag * (double)cX[cell];
          if (mask & 1u) ap -= (double)cX[nbr(lk, lx, ly, lz, -1, 0, 0)];
          if (mask & 2u) ap -= (double)cX[nbr(lk, lx, ly, lz, 1, 0, 0)];
          if (mask & 4u) ap -= (double)cX[nbr(lk, lx, ly, lz, 0, -1, 0)];
          if (mask & 8u) ap -= (double)cX[nbr(lk, lx, ly, lz, 0, 1, 0)];
          if (mask & 16u) ap -= (double)cX[nbr(lk, lx, ly, lz, 0, 0, -1)];
          if (mask & 32u) ap -= (double)cX[nbr(lk, lx, ly, lz, 0, 0, 1)];
          r = (float)((double)d.bvec[cell] - ap);
          cR[cell] = r;
        }
        const float u = precondition(r, diag, water, lx, ly, lz, mode);
        if (water) cU[cell] = u;
      });
      // w = A u, s = A p, q = M^-1 s.
      if (!reduce<double>(d, sp, ++tag, 0, nullptr, tt)) break;
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
        const uint32_t cell = b * 64u + lane, mask = d.nbm[cell];
        const bool water = (mask & 64u) != 0u;
        float sv = 0.0f, diag = 1.0f;
        if (water) {
          diag = cDiag[cell];
          cW[cell] = mem_row(cU, diag, mask, cell, lk, lx, ly, lz);
          sv = mem_row(cP, diag, mask, cell, lk, lx, ly, lz);
          cS[cell] = sv;
        }
        const float q = precondition(sv, diag, water, lx, ly, lz, mode);
        if (water) cQ[cell] = q;
      });
      // z = A q; m = M^-1 w and the partials.
      if (!reduce<double>(d, sp, ++tag, 0, nullptr, tt)) break;
      for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
        const uint32_t cell = b * 64u + lane, mask = d.nbm[cell];
        const bool water = (mask & 64u) != 0u;
        float r = 0.0f, u = 0.0f, w = 0.0f, diag = 1.0f;
        if (water) {
          diag = cDiag[cell];
          cZ[cell] = mem_row(cQ, diag, mask, cell, lk, lx, ly, lz);
          r = cR[cell];
          u = cU[cell];
          w = cW[cell];
        }
        pipe_tail_memory(d, b, lane, lx, ly, lz, r, u, w, diag, water, mo);
      });
    }
  }
  solve_finish<double>(d, sp, tag, s);
}

template <class L>
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve_pipe(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  constexpr uint32_t RM = SOLVE_RMAX;
  if (d.st->dt_ns == 0u) return;
  // A test of the plain launch's failure path: the last workgroup stays out, as one that is not
  // resident would, and the others' bounded waits abort the solve.
  if constexpr (std::is_same<L, One>::value) {
    if (d.solve_plain == 2u && v.wg + 1u == v.wgs) return;
  }
  if (__hip_atomic_load(d.bar + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return;
  const SolveSpan sp = solve_span(v, false);
  uint32_t tag = __hip_atomic_load(d.bar + 4, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (!sp.in_regs) {
    solve_pipe_memory(d.self, sp, tag);
    return;
  }

  // Register form. Wave wv holds blocks b0 + wv + SOLVE_WAVES r (r < RM). Per block and lane:
  // x, r, u, w, z, q, s, p, m, diag and the neighbour mask; per axis the tagged word of the
  // halo cell across the block face (when it is water) as last loaded.
  const uint32_t mode = d.sc->precond;
  const uint32_t b0 = sp.b0, b1 = sp.b1;
  const uint32_t capc = d.capc;
  const float* const p0 = start_pressure(d);
  unsigned long long* const hz = d.hz;
  __shared__ LPart<double> lp;
  const uint32_t t0 = tag + 1u;  // the start's reduction
  unsigned long long* const hz0 = hz + (size_t)(t0 & 1u) * capc;
  for_blocks<SOLVE_WAVES>(d, b0, b1, SOLVE_WAVES, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    double v[4];
    init_cell<double>(d, b, lane, lx, ly, lz, lk, hz0, 8u * t0 + 5u, v, p0);
    if (lane == 0)
      for (uint32_t k = 0; k < 4u; ++k) lp.v[k][b - b0] = v[k];
  });
  double tt[4];
  SolveScalars<double> s;
  bool go = solve_start<double>(d, tt, reduce<double>(d, sp, ++tag, 4, &lp, tt), s);
  const bool refresh_on = d.sc->refresh_every > 0u;
  const double replace_above = PIPE_REPLACE * (double)d.U->solve[0] * (s.b_norm < 1.0 ? 1.0 : s.b_norm);

  __shared__ uint32_t lkr[SOLVE_WAVES][SOLVE_RMAX][27];
  const uint32_t wv = threadIdx.x >> 6, ln = lane_id();
  const int lx = (int)(ln >> 4), ly = (int)((ln >> 2) & 3u), lz = (int)(ln & 3u);
  const uint32_t face_bit[3] = {lx == 0 ? 1u : (lx == 3 ? 2u : 0u), ly == 0 ? 4u : (ly == 3 ? 8u : 0u),
                                lz == 0 ? 16u : (lz == 3 ? 32u : 0u)};
  float rX[RM], rR[RM], rU[RM], rW[RM], rZ[RM], rQ[RM], rS[RM], rP[RM], rMv[RM], rDg[RM];
  uint32_t rM[RM];
  unsigned long long hW[RM][3];
  auto halo_cell = [&](uint32_t r, int a) -> uint32_t {
    const int sg = (face_bit[a] & 0x2Au) ? 1 : -1;
    return nbr(lkr[wv][r], lx, ly, lz, a == 0 ? sg : 0, a == 1 ? sg : 0, a == 2 ? sg : 0);
  };
  auto has_halo = [&](uint32_t r, int a) -> bool { return (rM[r] & face_bit[a]) != 0u; };
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b < b1 && ln < 27) lkr[wv][r][ln] = d.block_nb[b * 27u + ln];
  }
  __syncthreads();
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    rX[r] = rR[r] = rU[r] = rW[r] = rZ[r] = rQ[r] = rS[r] = rP[r] = rMv[r] = 0.0f;
    rDg[r] = 1.0f;
    rM[r] = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) hW[r][a] = 0ull;
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b >= b1) continue;
    const uint32_t cell = b * 64u + ln;
    rM[r] = d.nbm[cell];
    if (rM[r] & 64u) {
      rX[r] = p0[cell];
      rR[r] = d.cf[(size_t)NATRAY_F_R * capc + cell];
      rU[r] = d.cf[(size_t)NATRAY_F_Z * capc + cell];
      rDg[r] = d.cf[(size_t)NATRAY_F_DIAG * capc + cell];
    }
  }
  bool ok = true;
  // The halo of word buffer `buf` with tag `want` for block r: waits for each halo word and
  // returns the payloads (0 where there is no halo cell).
  auto halo = [&](uint32_t r, const unsigned long long* buf, uint32_t want, float h[3]) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      h[a] = 0.0f;
      if (!has_halo(r, a)) continue;
      ok = wait_word(d, hW[r][a], &buf[halo_cell(r, a)], want) && ok;
      h[a] = __uint_as_float((uint32_t)hW[r][a]);
    }
  };
  // Publishes vals of this wave's water cells into buffer `buf` with tag `tg`.
  auto publish = [&](const float* vals, unsigned long long* buf, uint32_t tg) {
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b < b1 && (rM[r] & 64u)) store_word(&buf[b * 64u + ln], tag_word(tg, __float_as_uint(vals[r])));
    }
  };
  if (go) {
    // w = A u from u0 and its halo (published by init_cell with the start's reduction).
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      float h[3];
      halo(r, hz0, 8u * t0 + 5u, h);
      const float w = apply_row(rU[r], rDg[r], rM[r], h, lx, ly, lz);
      if (rM[r] & 64u) rW[r] = w;
    }
  }
  __shared__ RedLds<double> rl;
  while (go) {
    const uint32_t T = tag + 1u;  // this iteration's reduction
    unsigned long long* const hzT = hz + (size_t)(T & 1u) * capc;
    // The block partials of (r, u), (w, u) and (r, r), published at once.
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      const double v0 = wave_tree<double>(prod<double>(rR[r], rU[r])), v1 = wave_tree<double>(prod<double>(rW[r], rU[r])),
                   v2 = wave_tree<double>(prod<double>(rR[r], rR[r]));
      if (ln == 0) {
        lp.v[0][b - b0] = v0;
        lp.v[1][b - b0] = v1;
        lp.v[2][b - b0] = v2;
      }
    }
    reduce_publish<double>(d, sp, ++tag, 3, &lp, rl);
    // While the other workgroups' chunks arrive: m = M^-1 w, published for the neighbours.
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      const bool water = (rM[r] & 64u) != 0u;
      const float m = precondition(rW[r], rDg[r], water, lx, ly, lz, mode);
      rMv[r] = water ? m : 0.0f;
    }
    publish(rMv, hzT, 8u * T);
    // The halo m words, polled while the reduction completes.
    auto halo_poll = [&](uint32_t) -> bool {
      bool fresh = true;
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        if (b0 + wv + SOLVE_WAVES * r >= b1) continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if (!has_halo(r, a) || (uint32_t)(hW[r][a] >> 32) == 8u * T) continue;
          hW[r][a] = load_word(&hzT[halo_cell(r, a)]);
          fresh = fresh && (uint32_t)(hW[r][a] >> 32) == 8u * T;
        }
      }
      return fresh;
    };
    if (!reduce_wait<double>(d, sp, tag, 3, tt, rl, halo_poll) || !ok) break;
    if (!pipe_scalars(d, tt, s)) break;
    const float fa = (float)s.alpha, fb = (float)s.beta;
    const bool refresh = refresh_on && (s.iters + 1u) % d.sc->refresh_every == 0u && s.r_norm > replace_above;
#pragma unroll
    for (uint32_t r = 0; r < RM; ++r) {
      const uint32_t b = b0 + wv + SOLVE_WAVES * r;
      if (b >= b1) continue;
      float hm[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) hm[a] = has_halo(r, a) ? __uint_as_float((uint32_t)hW[r][a]) : 0.0f;
      const float n = apply_row(rMv[r], rDg[r], rM[r], hm, lx, ly, lz);
      if (rM[r] & 64u) {
        rZ[r] = n + fb * rZ[r];
        rQ[r] = rMv[r] + fb * rQ[r];
        rS[r] = rW[r] + fb * rS[r];
        rP[r] = rU[r] + fb * rP[r];
        rX[r] = rX[r] + fa * rP[r];
        rR[r] = rR[r] - fa * rS[r];
        rU[r] = rU[r] - fa * rQ[r];
        rW[r] = rW[r] - fa * rZ[r];
      }
    }
    s.iters += 1u;
    if (refresh) {
      // Residual replacement through four halo exchanges: x (r = b - A x in double, u = M^-1 r),
      // u (w = A u), p (s = A p, q = M^-1 s), q (z = A q); x and p in the buffer m of this
      // iteration is not in, u and q in the other.
      unsigned long long* const hzA = hz + (size_t)((T + 1u) & 1u) * capc;
      publish(rX, hzA, 8u * T + 1u);
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        const uint32_t b = b0 + wv + SOLVE_WAVES * r;
        if (b >= b1) continue;
        const uint32_t cell = b * 64u + ln, mask = rM[r];
        const bool water = (mask & 64u) != 0u;
        float h[3];
        halo(r, hzA, 8u * T + 1u, h);
        const Faces f = block_faces(rX[r]);
        if (water) {
          double ap = (double)rDg[r] * (double)rX[r];
          if (mask & 1u) ap -= (double)(lx > 0 ? f.xm : h[0]);
          if (mask & 2u) ap -= (double)(lx < 3 ? f.xp : h[0]);
          if (mask & 4u) ap -= (double)(ly > 0 ? f.ym : h[1]);
          if (mask & 8u) ap -= (double)(ly < 3 ? f.yp : h[1]);
          if (mask & 16u) ap -= (double)(lz > 0 ? f.zm : h[2]);
          if (mask & 32u) ap -= (double)(lz < 3 ? f.zp : h[2]);
          rR[r] = (float)((double)d.bvec[cell] - ap);
        }
        const float u = precondition(rR[r], rDg[r], water, lx, ly, lz, mode);
        if (water) rU[r] = u;
      }
      publish(rU, hzT, 8u * T + 2u);
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        if (b0 + wv + SOLVE_WAVES * r >= b1) continue;
        float h[3];
        halo(r, hzT, 8u * T + 2u, h);
        const float w = apply_row(rU[r], rDg[r], rM[r], h, lx, ly, lz);
        if (rM[r] & 64u) rW[r] = w;
      }
      publish(rP, hzA, 8u * T + 3u);
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        if (b0 + wv + SOLVE_WAVES * r >= b1) continue;
        const bool water = (rM[r] & 64u) != 0u;
        float h[3];
        halo(r, hzA, 8u * T + 3u, h);
        const float sv = apply_row(rP[r], rDg[r], rM[r], h, lx, ly, lz);
        if (water) rS[r] = sv;
        const float q = precondition(rS[r], rDg[r], water, lx, ly, lz, mode);
        if (water) rQ[r] = q;
      }
      publish(rQ, hzT, 8u * T + 4u);
#pragma unroll
      for (uint32_t r = 0; r < RM; ++r) {
        if (b0 + wv + SOLVE_WAVES * r >= b1) continue;
        float h[3];
        halo(r, hzT, 8u * T + 4u, h);
        const float z = apply_row(rQ[r], rDg[r], rM[r], h, lx, ly, lz);
        if (rM[r] & 64u) rZ[r] = z;
      }
    }
  }
  // The fields the other kernels and the host read.
#pragma unroll
  for (uint32_t r = 0; r < RM; ++r) {
    const uint32_t b = b0 + wv + SOLVE_WAVES * r;
    if (b < b1 && (rM[r] & 64u)) {
      const uint32_t cell = b * 64u + ln;
      d.cf[(size_t)NATRAY_F_P * capc + cell] = rX[r];
      d.cf[(size_t)NATRAY_F_R * capc + cell] = rR[r];
      d.cf[(size_t)NATRAY_F_Z * capc + cell] = rU[r];
      if (s.iters > 0u) {
        d.cf[(size_t)NATRAY_F_Q * capc + cell] = rS[r];
        d.cf[(size_t)NATRAY_F_D * capc + cell] = rP[r];
        d.cf[(size_t)(NATRAY_F_D + 1u) * capc + cell] = rW[r];
      }
    }
  }
  solve_finish<double>(d, sp, tag, s);
}

// ---- learned warm start of the solve ---------------------------------------------------------------
//
// Not in the reference, whose solve always starts from the particle-carried pressure
// (solve_init, fluid.wgsl:450-489). Per step: r0 = b - A p_carried on every water cell; the 16^3
// tiles of cells that hold water are listed (at most a fixed number); each listed tile's inputs
// (r0 / s with s the tile's largest |r0|, and the cell kind 1 water / -1 solid / 0 otherwise)
// go to the network, whose output o gives guess = p_carried + s o on the tile's water cells.
// The solve starts from the guess when |b - A guess|^2 < |r0|^2 over all water cells (both sums
// in FP64, fixed order), else from the carried pressure as without the warm start. b is the
// same either way, so the solve meets the same tolerance on the same system.

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t o = (uint32_t)__shfl_xor((int)v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// The cell of voxel v (256 x + 16 y + z) of tile t, or NONE where no block holds it.
__device__ __forceinline__ uint32_t tile_voxel_cell(const Dev& d, uint32_t t, uint32_t v, uint32_t nb) {
  const uint32_t tz = t % d.tdz, ty = (t / d.tdz) % d.tdy, tx = t / (d.tdz * d.tdy);
  const uint32_t cx = tx * 16u + (v >> 8), cy = ty * 16u + ((v >> 4) & 15u), cz = tz * 16u + (v & 15u);
  const uint32_t bx = cx >> 2, by = cy >> 2, bz = cz >> 2;
  const natray_params* U = d.U;
  if (bx >= U->bdims[0] || by >= U->bdims[1] || bz >= U->bdims[2]) return NONE;
  const uint32_t b = d.block_ptr[block_slot(U, (int)bx, (int)by, (int)bz)];
  if (b >= nb) return NONE;
  return b * 64u + (cx & 3u) * 16u + (cy & 3u) * 4u + (cz & 3u);
}

// r0 from the carried pressure, the carried pressure kept, the guess preset to it, and the
// tiles with water marked with their largest |r0|.
template <class L>
__global__ __launch_bounds__(256) void k_ws_prepare(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const float* cP = d.cf + (size_t)NATRAY_F_P * d.capc;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    const uint32_t cell = b * 64u + lane;
    const Row w = cell_row(d, cell, lx, ly, lz, lk, cP);
    const float p = cP[cell];
    d.ws_r0[cell] = w.water ? w.r : 0.0f;
    d.ws_pcar[cell] = p;
    d.guess[cell] = p;
    const uint32_t m = wave_max(w.water ? __float_as_uint(__builtin_fabsf(w.r)) + 1u : 0u);
    if (lane == 0 && m != 0u) {
      int bx, by, bz;
      slot_coord(d.U, d.block_list[b], bx, by, bz);
      const uint32_t t = ((uint32_t)(bx >> 2) * d.tdy + (uint32_t)(by >> 2)) * d.tdz + (uint32_t)(bz >> 2);
      atomicMax(&d.tile_map[t], m);
    }
  });
}

// One workgroup: lists the marked tiles in tile order (the first `cap`) with their scale,
// clears the marks, and stores the listed count in bar[10].
template <class L>
__global__ __launch_bounds__(1024) void k_ws_tiles(L l, uint32_t cap) {
  const View v = view(l);
  const Dev& d = v.d;
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t base;
  const uint32_t T = d.tdx * d.tdy * d.tdz;
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (threadIdx.x == 0) base = 0u;
  __syncthreads();
  for (uint32_t t0 = 0; t0 < T; t0 += 1024u) {
    const uint32_t t = t0 + threadIdx.x;
    const uint32_t m = t < T ? d.tile_map[t] : 0u;
    const unsigned long long bal = __ballot(m != 0u);
    if (lane == 0) wsum[wave] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t slot = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
    for (uint32_t w = 0; w < wave; ++w) slot += wsum[w];
    if (t < T) {
      if (m != 0u && slot < cap) {
        d.tile_list[slot] = t;
        d.tile_scale[slot] = __uint_as_float(m - 1u);
      }
      d.tile_map[t] = 0u;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t add = 0;
      for (uint32_t w = 0; w < 16u; ++w) add += wsum[w];
      base += add;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    d.bar[10] = base < cap ? base : cap;
    d.ws_stat[4] = (double)base;
  }
}

// The network's inputs of the listed tiles: FP32 into ws_in ([slot][voxel][2]) and BF16 into
// the network's input buffer ([slot][voxel][8], channels 2..7 zero).
template <class L>
__global__ __launch_bounds__(256) void k_ws_gather(L l, bf16x8* __restrict__ net_in) {
  const View v = view(l);
  const Dev& d = v.d;
  const uint32_t nb = n_blocks(d);
  const size_t total = (size_t)d.bar[10] * 4096u;
  for (size_t i = (size_t)v.wg * blockDim.x + threadIdx.x; i < total; i += (size_t)v.wgs * blockDim.x) {
    const uint32_t slot = (uint32_t)(i >> 12), v = (uint32_t)(i & 4095u);
    const uint32_t cell = tile_voxel_cell(d, d.tile_list[slot], v, nb);
    float x = 0.0f, kind = 0.0f;
    if (cell != NONE) {
      const float ty = cfr(d, NATRAY_F_TYPE, cell);
      const float s = d.tile_scale[slot];
      if (ty == WATER) {
        kind = 1.0f;
        x = s > 0.0f ? d.ws_r0[cell] / s : 0.0f;
      } else if (ty == SOLID) {
        kind = -1.0f;
      }
    }
    d.ws_in[2 * i] = x;
    d.ws_in[2 * i + 1] = kind;
    if (net_in) {
      bf16x8 o;
#pragma unroll
      for (int j = 0; j < 8; ++j) o[j] = (__bf16)0.0f;
      o[0] = (__bf16)x;
      o[1] = (__bf16)kind;
      net_in[i] = o;
    }
  }
}

// guess = p_carried + s o on the water cells of the listed tiles (o: channel 0 of `pred`,
// [slot][voxel][16] FP32).
template <class L>
__global__ __launch_bounds__(256) void k_ws_scatter(L l, const float* __restrict__ pred) {
  const View v = view(l);
  const Dev& d = v.d;
  const uint32_t nb = n_blocks(d);
  const size_t total = (size_t)d.bar[10] * 4096u;
  for (size_t i = (size_t)v.wg * blockDim.x + threadIdx.x; i < total; i += (size_t)v.wgs * blockDim.x) {
    const uint32_t slot = (uint32_t)(i >> 12), v = (uint32_t)(i & 4095u);
    const uint32_t cell = tile_voxel_cell(d, d.tile_list[slot], v, nb);
    if (cell == NONE || cfr(d, NATRAY_F_TYPE, cell) != WATER) continue;
    d.guess[cell] = d.ws_pcar[cell] + d.tile_scale[slot] * pred[i * 16u];
  }
}

// Per block: the squared residuals of the guess and of the carried pressure (lane tree, FP64).
template <class L>
__global__ __launch_bounds__(256) void k_ws_check(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    const uint32_t cell = b * 64u + lane;
    const Row w = cell_row(d, cell, lx, ly, lz, lk, d.guess);
    const double rg = w.water ? (double)w.r : 0.0, r0 = w.water ? (double)d.ws_r0[cell] : 0.0;
    const double sg = lane_tree<double>(rg * rg), s0 = lane_tree<double>(r0 * r0);
    if (lane == 0) {
      d.ws_part[2 * b] = sg;
      d.ws_part[2 * b + 1] = s0;
    }
  });
}

// One workgroup: sums the block residuals in a fixed order and accepts the guess (bar[11] = 1)
// when its residual is the smaller; keeps the sums and counts in ws_stat.
template <class L>
__global__ __launch_bounds__(1024) void k_ws_decide(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  __shared__ double sh[2][1024];
  if (d.st->dt_ns == 0u) return;
  const uint32_t n = n_blocks(d);
  double sg = 0.0, s0 = 0.0;
  for (uint32_t b = threadIdx.x; b < n; b += 1024u) {
    sg += d.ws_part[2 * b];
    s0 += d.ws_part[2 * b + 1];
  }
  sh[0][threadIdx.x] = sg;
  sh[1][threadIdx.x] = s0;
  __syncthreads();
  for (uint32_t h = 512u; h > 0u; h >>= 1) {
    if (threadIdx.x < h) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + h];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool accept = sh[0][0] < sh[1][0];
    d.bar[11] = accept ? 1u : 0u;
    d.ws_stat[0] += 1.0;
    d.ws_stat[1] += accept ? 1.0 : 0.0;
    d.ws_stat[2] = sh[0][0];
    d.ws_stat[3] = sh[1][0];
  }
}

// Training samples of the last solve on the listed tiles, [slot][voxel]: target
// (p - p_carried) / s, the water mask, and the operator's diagonal and water-neighbour bits.
// The zero guess (warm-start mode 3): this step's solve starts from the guess buffer, which the
// host holds at zero.
template <class L>
__global__ __launch_bounds__(64) void k_ws_zero_start(L l) {
  const View v = view(l);
  if (v.wg == 0u && threadIdx.x == 0u) v.d.bar[11] = 1u;
}

template <class L>
__global__ __launch_bounds__(256) void k_ws_samples(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  const uint32_t nb = n_blocks(d);
  const float* cP = d.cf + (size_t)NATRAY_F_P * d.capc;
  const size_t total = (size_t)d.bar[10] * 4096u;
  for (size_t i = (size_t)v.wg * blockDim.x + threadIdx.x; i < total; i += (size_t)v.wgs * blockDim.x) {
    const uint32_t slot = (uint32_t)(i >> 12), v = (uint32_t)(i & 4095u);
    const uint32_t cell = tile_voxel_cell(d, d.tile_list[slot], v, nb);
    const float s = d.tile_scale[slot];
    const bool on = cell != NONE && s > 0.0f && cfr(d, NATRAY_F_TYPE, cell) == WATER;
    d.ws_tgt[i] = on ? (cP[cell] - d.ws_pcar[cell]) / s : 0.0f;
    d.ws_mask[i] = on ? 1.0f : 0.0f;
    d.ws_diag[i] = on ? cfr(d, NATRAY_F_DIAG, cell) : 0.0f;
    d.ws_nbm[i] = on ? (d.nbm[cell] & 63u) : 0u;
  }
}

// ---- project, extend_1, extend_2 (fluid.wgsl:580-674) ------------------------------------------------

template <class L>
__global__ __launch_bounds__(256) void k_project(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const float k = dt_of(d) / (d.U->fluid[0] * d.U->origin[3]);
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u,
             [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
               const uint32_t cell = b * 64u + lane;
               const float ty = cfr(d, NATRAY_F_TYPE, cell);
               // Both pressures are loaded before the face test that decides whether they are
               // used (pressure_of, fluid.wgsl:375-379, on the loaded values).
               const float pc = cfr(d, NATRAY_F_P, cell), qc = ty == WATER ? fmax_(pc, NATRAY_P_FLOOR) : pc;
               uint32_t nb[3];
               float nt[3], pn[3], un[3];
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim) {
                 nb[dim] = nbr(lk, lx, ly, lz, -(dim == 0), -(dim == 1), -(dim == 2));
                 nt[dim] = cell_type(d, nb[dim]);
                 pn[dim] = nb[dim] != NONE ? cfr(d, NATRAY_F_P, nb[dim]) : 0.0f;
                 un[dim] = cfr(d, NATRAY_F_U + dim, cell);
               }
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim) {
                 float u = bounded(d, un[dim], ty, nt[dim], cell, nb[dim], dim);
                 if (ty != SOLID && nt[dim] != SOLID && (ty == WATER || nt[dim] == WATER)) {
                   const float qn = nb[dim] == NONE ? 0.0f : (nt[dim] == WATER ? fmax_(pn[dim], NATRAY_P_FLOOR) : pn[dim]);
                   u -= k * (qc - qn);
                 }
                 cfr(d, NATRAY_F_U + dim, cell) = u;
               }
             });
}

// The six face neighbours, k = 2 axis + (0: -1, 1: +1).
__device__ __forceinline__ void face_offsets(int k, int& ox, int& oy, int& oz) {
  const int sgn = (k & 1) ? 1 : -1;
  ox = (k >> 1) == 0 ? sgn : 0;
  oy = (k >> 1) == 1 ? sgn : 0;
  oz = (k >> 1) == 2 ? sgn : 0;
}

// extend_1 (fluid.wgsl:607-640). A wave whose cells are all water has every face set and keeps
// its velocities. Any other wave loads, before it tests anything, every cell type and velocity
// the tests may need, so their loads are in flight together rather than one dependent round
// trip per test; the faces and sums are those of the reference, in its order.
template <class L>
__global__ __launch_bounds__(256) void k_extend1(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u,
             [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
               const uint32_t cell = b * 64u + lane;
               const float tc = cfr(d, NATRAY_F_TYPE, cell);
               float u[3];
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim) u[dim] = cfr(d, NATRAY_F_U + dim, cell);
               if (__ballot(tc != WATER) == 0ull) {
#pragma unroll
                 for (uint32_t dim = 0; dim < 3; ++dim) cfr(d, NATRAY_F_EXT + dim, cell) = u[dim];
                 cfr(d, NATRAY_F_EXT_SET, cell) = 7.0f;
                 return;
               }
               uint32_t nb[6];
               float tn[6];
#pragma unroll
               for (int k = 0; k < 6; ++k) {
                 int ox, oy, oz;
                 face_offsets(k, ox, oy, oz);
                 nb[k] = nbr(lk, lx, ly, lz, ox, oy, oz);
                 tn[k] = cell_type(d, nb[k]);
               }
               // face_set (fluid.wgsl:607-612) of neighbour k for component dim: the neighbour is
               // water, or the cell below it along dim is.
               float tb[3][6], un[3][6];
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim)
#pragma unroll
                 for (int k = 0; k < 6; ++k) {
                   int ox, oy, oz;
                   face_offsets(k, ox, oy, oz);
                   tb[dim][k] = nb[k] == NONE ? GAS
                                              : cell_type(d, nbr(lk, lx, ly, lz, ox - (dim == 0), oy - (dim == 1), oz - (dim == 2)));
                   un[dim][k] = nb[k] == NONE ? 0.0f : cfr(d, NATRAY_F_U + dim, nb[k]);
                 }
               uint32_t done = 0;
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim) {
                 float ud = u[dim];
                 if (tc == WATER || tn[2 * dim] == WATER) {
                   done |= 1u << dim;
                 } else {
                   float sum = 0.0f, cnt = 0.0f;
#pragma unroll
                   for (int k = 0; k < 6; ++k) {
                     if (nb[k] != NONE && (tn[k] == WATER || tb[dim][k] == WATER)) {
                       sum += un[dim][k];
                       cnt += 1.0f;
                     }
                   }
                   if (cnt > 0.0f) {
                     ud = sum / cnt;
                     done |= 1u << dim;
                   }
                 }
                 cfr(d, NATRAY_F_EXT + dim, cell) = ud;
               }
               cfr(d, NATRAY_F_EXT_SET, cell) = (float)done;
             });
}

// extend_2 (fluid.wgsl:642-674), then bounded: as extend_1, a wave whose faces are all set
// skips the neighbours, and any other loads every neighbour's mark and value up front.
template <class L>
__global__ __launch_bounds__(256) void k_extend2(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u,
             [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
               const uint32_t cell = b * 64u + lane;
               const float ty = cfr(d, NATRAY_F_TYPE, cell);
               const uint32_t set = (uint32_t)cfr(d, NATRAY_F_EXT_SET, cell);
               uint32_t nb[6];
#pragma unroll
               for (int k = 0; k < 6; ++k) {
                 int ox, oy, oz;
                 face_offsets(k, ox, oy, oz);
                 nb[k] = nbr(lk, lx, ly, lz, ox, oy, oz);
               }
               float u[3];
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim)
                 u[dim] = (set >> dim & 1u) ? cfr(d, NATRAY_F_EXT + dim, cell) : cfr(d, NATRAY_F_U + dim, cell);
               if (__ballot(set != 7u) != 0ull) {
                 uint32_t sn[6];
                 float en[3][6];
#pragma unroll
                 for (int k = 0; k < 6; ++k) sn[k] = nb[k] == NONE ? 0u : (uint32_t)cfr(d, NATRAY_F_EXT_SET, nb[k]);
#pragma unroll
                 for (uint32_t dim = 0; dim < 3; ++dim)
#pragma unroll
                   for (int k = 0; k < 6; ++k) en[dim][k] = nb[k] == NONE ? 0.0f : cfr(d, NATRAY_F_EXT + dim, nb[k]);
#pragma unroll
                 for (uint32_t dim = 0; dim < 3; ++dim) {
                   if (set >> dim & 1u) continue;
                   float sum = 0.0f, cnt = 0.0f;
#pragma unroll
                   for (int k = 0; k < 6; ++k) {
                     if (nb[k] != NONE && (sn[k] >> dim & 1u)) {
                       sum += en[dim][k];
                       cnt += 1.0f;
                     }
                   }
                   if (cnt > 0.0f) u[dim] = sum / cnt;
                 }
               }
#pragma unroll
               for (uint32_t dim = 0; dim < 3; ++dim) {
                 const uint32_t lo = nb[2 * dim];
                 cfr(d, NATRAY_F_U + dim, cell) = bounded(d, u[dim], ty, cell_type(d, lo), cell, lo, dim);
               }
             });
}

// ---- g2p (fluid.wgsl:767-901) ------------------------------------------------------------------------

struct Gathered {
  float s, g0, g1, g2, w;
};

// gather (fluid.wgsl:780-802)
template <int OX, int OY, int OZ>
__device__ __forceinline__ Gathered gather(const float* tg, const float X[3], const int corner[3], uint32_t fa, uint32_t fw,
                                           float h) {
  const float o[3] = {0.5f * OX, 0.5f * OY, 0.5f * OZ};
  int base[3];
  float fx[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    base[k] = (int)floorf(X[k] - o[k] - 0.5f);
    fx[k] = X[k] - o[k] - (float)base[k];
  }
  float wx[3], wy[3], wz[3];
  axis_weights(fx[0], wx);
  axis_weights(fx[1], wy);
  axis_weights(fx[2], wz);
  const int t0 = (base[0] - corner[0]) * 49 + (base[1] - corner[1]) * 7 + (base[2] - corner[2]);
  const float* ta = tg + fa * TILE + t0;
  const float* tw = tg + fw * TILE + t0;
  Gathered r = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float weight = wx[a] * wy[b] * wz[c];
        const int t = a * 49 + b * 7 + c;
        const float u = weight * ta[t];
        r.s += u;
        r.g0 += u * ((float)a - fx[0]);
        r.g1 += u * ((float)b - fx[1]);
        r.g2 += u * ((float)c - fx[2]);
        r.w += weight * tw[t];
      }
  r.g0 *= h;
  r.g1 *= h;
  r.g2 *= h;
  return r;
}

// carve for the moving capsule (the role of bullet.wgsl:395-424).
__device__ __forceinline__ void carve(const Dev& d, float x[3], float v[3]) {
  const float r = d.sc->capsule_a[3];
  float on[3], ab[3];
  float d2 = capsule_d2(d, x[0], x[1], x[2], now_of(d), on, ab);
  if (d2 >= r * r) return;
  float dd[3] = {x[0] - on[0], x[1] - on[1], x[2] - on[2]};
  if (d2 < 1e-14f) {
    if (ab[0] * ab[0] + ab[2] * ab[2] > 1e-20f) {
      dd[0] = -ab[2]; dd[1] = 0.0f; dd[2] = ab[0];
    } else {
      dd[0] = 0.0f; dd[1] = ab[2]; dd[2] = -ab[1];
    }
    d2 = (dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2];
  }
  const float len = __builtin_sqrtf(d2);
  const float dir[3] = {dd[0] / len, dd[1] / len, dd[2] / len};
  x[0] = on[0] + dir[0] * r;
  x[1] = on[1] + dir[1] * r;
  x[2] = on[2] + dir[2] * r;
  const float* cv = d.sc->capsule_v;
  const float wall = (cv[0] * dir[0] + cv[1] * dir[1]) + cv[2] * dir[2];
  const float vn = (v[0] * dir[0] + v[1] * dir[1]) + v[2] * dir[2];
  const float push = fmax_(wall - vn, 0.0f);
  v[0] += dir[0] * push;
  v[1] += dir[1] * push;
  v[2] += dir[2] * push;
}

// ---- spray and wet ground (fluid.wgsl:676-765, sim_defs.wgsl:233-240; NATRAY_SCENE_SPRAY) --------
//
// A particle that leaves the resolved flow becomes spray (break_up) or, on the ground, wets it
// (ground_touch); the drops fly through still air until they land (spray_update, before G2P).
// exp, log, cos and pow come from natray_math.h, which the CPU oracle uses too, so the records
// are the oracle's bit for bit; their slots come from an atomic add, so their order is not.

constexpr float RHO_AIR = 1.2f;          // sim_defs.wgsl:101
constexpr float MU_AIR = 1.8e-5f;        // sim_defs.wgsl:102
constexpr float EVAPORATION = 1.3e-10f;  // fluid.wgsl:684
constexpr float WE_CRIT = 12.0f;         // fluid.wgsl:718
constexpr float PI_F = 3.14159265358979f;  // common.wgsl:1
constexpr uint32_t BREAKUP_DROPS = 4;    // fluid.wgsl:719

// pcg, hash3, rnd, rnd_gauss (common.wgsl:3-22)
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
  const uint32_t t = v * 747796405u + 2891336453u;
  const uint32_t w = ((t >> ((t >> 28u) + 4u)) ^ t) * 277803737u;
  return (w >> 22u) ^ w;
}
__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c) { return pcg(a ^ pcg(b ^ pcg(c))); }
__device__ __forceinline__ float rnd(uint32_t& seed) {
  seed = pcg(seed);
  return (float)(seed >> 8u) * (1.0f / 16777216.0f);
}
__device__ __forceinline__ float rnd_gauss(uint32_t& seed) {
  const float u1 = fmax_(rnd(seed), 1e-7f);
  const float u2 = rnd(seed);
  return __builtin_sqrtf(-2.0f * nm_logf(u1)) * nm_cosf(2.0f * PI_F * u2);
}

// ground_touch (sim_defs.wgsl:233-240); the cell index is clamped to the map, which matters
// only for map sides that are not powers of two.
__device__ __forceinline__ void ground_touch(const Dev& d, float px, float pz) {
  const float hw = d.U->wet[0];
  const uint32_t res = d.wet_res;
  const float ux = (px + hw) / (2.0f * hw), uz = (pz + hw) / (2.0f * hw);
  if (ux < 0.0f || uz < 0.0f || ux >= 1.0f || uz >= 1.0f || res == 0u) return;
  const uint32_t cx = min((uint32_t)(ux * (float)res), res - 1u), cz = min((uint32_t)(uz * (float)res), res - 1u);
  atomicMin(&d.touch[(size_t)cz * res + cx], __float_as_uint(now_of(d)));
}

// break_up (fluid.wgsl:720-765): four drops and a parcel of finer drops, in five consecutive
// spray slots. G2P lists the particles that break up and k_break_up runs this for them, so the
// G2P kernel keeps no call frame of its own.
__device__ __forceinline__ void break_up(const Dev& d, float x0, float x1, float x2, float v0, float v1, float v2, uint32_t id) {
  const uint32_t cap = d.spray_cap;
  const uint32_t first = atomicAdd(&d.st->spray_count, BREAKUP_DROPS + 1u);
  if (first >= cap) return;
  uint32_t seed = hash3(d.st->step, id, 0xb4eau);
  const float hs = d.U->fluid[3];
  const float vol = hs * hs * hs;
  const float speed = __builtin_sqrtf((v0 * v0 + v1 * v1) + v2 * v2);
  // breakup_diameter (fluid.wgsl:713-716)
  const float bd = (WE_CRIT * NATRAY_SIGMA_W) / (RHO_AIR * fmax_(speed * speed, 1e-4f));
  const float d_med = fmin_(0.55f * hs, 0.5f * bd);
  float dd[BREAKUP_DROPS];
  float used = 0.0f;
#pragma unroll
  for (uint32_t k = 0; k < BREAKUP_DROPS; ++k) {
    dd[k] = fmin_(fmax_(d_med * nm_expf(0.5f * rnd_gauss(seed)), 2e-5f), 1.24f * hs);
    used += 0.5236f * dd[k] * dd[k] * dd[k];
  }
  const float df = 0.5f * d_med;
  const bool fine = df < 1e-4f;
  const float fit = nm_powf(vol / used, 1.0f / 3.0f);
  const float scale = fine ? fmin_(1.0f, fit) : fit;
  const float tnow = now_of(d);
  const float x[3] = {x0, x1, x2}, v[3] = {v0, v1, v2};
#pragma unroll
  for (uint32_t k = 0; k < BREAKUP_DROPS; ++k) {
    const uint32_t slot = first + k;
    if (slot >= cap) return;
    const float dk = dd[k] * scale;
    const float spread = __builtin_sqrtf(NATRAY_SIGMA_W / (d.U->fluid[0] * dk)) + 0.05f * speed;
    const float g0 = rnd_gauss(seed);
    const float g1 = rnd_gauss(seed);
    const float g2 = rnd_gauss(seed);
    const float w0 = g0 + 1e-6f, w1 = g1 + 1e-6f, w2 = g2 + 1e-6f;
    const float len = __builtin_sqrtf((w0 * w0 + w1 * w1) + w2 * w2);
    const float u[3] = {w0 / len, w1 / len, w2 / len};
    const float o0 = rnd(seed);
    const float o1 = rnd(seed);
    const float o2 = rnd(seed);
    const float rv = rnd(seed);
    float4* rec = reinterpret_cast<float4*>(d.spray) + (size_t)slot * 3;
    rec[0] = make_float4(x[0] + (o0 - 0.5f) * hs, x[1] + (o1 - 0.5f) * hs, x[2] + (o2 - 0.5f) * hs, 0.5f * dk);
    rec[1] = make_float4(v[0] + u[0] * spread * rv, v[1] + u[1] * spread * rv, v[2] + u[2] * spread * rv, 1.0f);
    rec[2] = make_float4(0.0f, 0.0f, 0.0f, tnow);
  }
  const uint32_t slot = first + BREAKUP_DROPS;
  if (slot >= cap) return;
  const float n = (vol - used * scale * scale * scale) / (0.5236f * df * df * df);
  float4* rec = reinterpret_cast<float4*>(d.spray) + (size_t)slot * 3;
  rec[0] = make_float4(x[0], x[1], x[2], (fine && n >= 2.0f) ? 0.5f * df : 0.0f);
  rec[1] = make_float4(v[0], v[1], v[2], fmax_(n, 2.0f));
  rec[2] = make_float4(0.0f, 0.0f, 0.0f, tnow);
}

template <class L>
__global__ __launch_bounds__(256) void k_break_up(L l) {
  const View vw = view(l);
  const Dev& d = vw.d;
  if (d.st->dt_ns == 0u || !d.spray) return;
  const uint32_t n = d.bar[12];
  for (uint32_t k = vw.wg * 256u + threadIdx.x; k < n; k += vw.wgs * 256u) {
    const float4* e = reinterpret_cast<const float4*>(d.breakups) + (size_t)k * 2;
    const float4 a = e[0], b = e[1];
    break_up(d, a.y, a.z, a.w, b.x, b.y, b.z, __float_as_uint(a.x));
  }
}

// drag_cd (fluid.wgsl:676-679)
__device__ __forceinline__ float drag_cd(float re) {
  if (re < 1e-4f) return 24.0f / 1e-4f;
  return 24.0f / re * (1.0f + 0.15f * nm_powf(re, 0.687f)) + 0.42f / (1.0f + 42500.0f * nm_powf(re, -1.16f));
}

// spray_update (fluid.wgsl:686-711), grid-stride over the records made so far.
template <class L>
__global__ __launch_bounds__(256) void k_spray_update(L l) {
  const View vw = view(l);
  const Dev& d = vw.d;
  if (d.st->dt_ns == 0u || !d.spray) return;
  const uint32_t n = min(d.st->spray_count, d.spray_cap);
  const float step = dt_of(d), tnow = now_of(d);
  const float g[3] = {d.U->gravity[0], d.U->gravity[1], d.U->gravity[2]};
  for (uint32_t i = vw.wg * 256u + threadIdx.x; i < n; i += vw.wgs * 256u) {
    float4* rec = reinterpret_cast<float4*>(d.spray) + (size_t)i * 3;
    const float4 pos = rec[0], vel = rec[1];
    float r = pos.w;
    if (r <= 0.0f) continue;
    const float sp = __builtin_sqrtf((vel.x * vel.x + vel.y * vel.y) + vel.z * vel.z);
    const float re = 2.0f * r * sp * RHO_AIR / MU_AIR;
    const float k = 3.0f * RHO_AIR * drag_cd(re) * sp / (8.0f * d.U->fluid[0] * r);
    const float den = 1.0f + k * step;
    const float vn[3] = {(vel.x + step * g[0]) / den, (vel.y + step * g[1]) / den, (vel.z + step * g[2]) / den};
    r = __builtin_sqrtf(fmax_(r * r - EVAPORATION * step, 0.0f));
    if (r < 2.5e-6f) r = 0.0f;
    const float p[3] = {pos.x + step * vn[0], pos.y + step * vn[1], pos.z + step * vn[2]};
    if (p[1] <= r) {
      reinterpret_cast<float4*>(d.landed)[i] = make_float4(p[0], p[2], r * nm_powf(vel.w, 1.0f / 3.0f), tnow);
      r = 0.0f;
    }
    rec[0] = make_float4(p[0], p[1], p[2], r);
    rec[1] = make_float4(vn[0], vn[1], vn[2], vel.w);
  }
}

// ---- bullets: vectors, cavity sections and carve (bullet.wgsl:85-177, 388-424) --------------------
//
// Every vector expression is written out in the order the WGSL source evaluates it (left to
// right, no contraction), so that the CPU restatement gives the same bits.

struct b3 {
  float x, y, z;
};
__device__ __forceinline__ b3 operator+(b3 a, b3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ b3 operator-(b3 a, b3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ b3 operator*(b3 a, float k) { return {a.x * k, a.y * k, a.z * k}; }
__device__ __forceinline__ b3 operator/(b3 a, float k) { return {a.x / k, a.y / k, a.z / k}; }
__device__ __forceinline__ b3 adds(b3 a, float k) { return {a.x + k, a.y + k, a.z + k}; }
__device__ __forceinline__ float dot3(b3 a, b3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ b3 cross3(b3 a, b3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len3(b3 a) { return __builtin_sqrtf(dot3(a, a)); }
__device__ __forceinline__ b3 norm3(b3 a) { return a / len3(a); }
__device__ __forceinline__ b3 xyz(const float* v) { return {v[0], v[1], v[2]}; }
__device__ __forceinline__ b3 xyz(float4 v) { return {v.x, v.y, v.z}; }
// qrot and qmul (bullet.wgsl:91-98); quaternions are (x, y, z, w).
__device__ __forceinline__ b3 qrot(float4 q, b3 v) {
  const b3 qv = xyz(q);
  const b3 t = cross3(qv, v) * 2.0f;
  return (v + t * q.w) + cross3(qv, t);
}
__device__ __forceinline__ float4 qmul(float4 a, float4 b) {
  const b3 r = (xyz(b) * a.w + xyz(a) * b.w) + cross3(xyz(a), xyz(b));
  return make_float4(r.x, r.y, r.z, a.w * b.w - dot3(xyz(a), xyz(b)));
}

constexpr float R_HAND = 1.5f;  // a section this wide is the grid's (bullet.wgsl:85)

// Section k of bullet bi: a.xyz and when made, b.xyz and d(R^2)/dt, R^2 at the start
// (bullet.wgsl:42-50).
__device__ __forceinline__ const float4* section_of(const Dev& d, uint32_t bi, uint32_t k) {
  return reinterpret_cast<const float4*>(d.sections) + ((size_t)bi * NATRAY_BULLET_SECTIONS + k) * 3;
}
// section_r2 (bullet.wgsl:152-154)
__device__ __forceinline__ float section_r2(const float4* s, float t) { return s[2].x + s[1].w * fmax_(t - s[0].w, 0.0f); }
__device__ __forceinline__ float hand2_of(float hh) { return R_HAND * R_HAND * hh * hh; }

// carve (bullet.wgsl:388-424): a particle inside a cavity section still thinner than the grid
// goes onto its wall and takes at least the wall's speed.
__device__ __forceinline__ void carve_cavity(const Dev& d, float x[3], float v[3], uint32_t id) {
  const float t = now_of(d);
  const float hand2 = hand2_of(d.U->origin[3]);
  b3 X = xyz(x), V = xyz(v);
  for (uint32_t bi = 0; bi < d.n_bullets; ++bi) {
    const float* reach = d.bullets[bi].reach;
    if (reach[3] <= 0.0f || len3(X - xyz(reach)) > reach[3]) continue;
    for (uint32_t k = 0; k < NATRAY_BULLET_SECTIONS; ++k) {
      const float4* s = section_of(d, bi, k);
      const float r2 = section_r2(s, t);
      if (s[1].w <= 0.0f || r2 >= hand2) continue;
      const b3 a = xyz(s[0]), ab = xyz(s[1]) - a;
      const float f = dot3(X - a, ab) / fmax_(dot3(ab, ab), 1e-20f);
      if (f < 0.0f || f > 1.0f) continue;
      const b3 on = a + ab * f;
      b3 dd = X - on;
      if (dot3(dd, dd) >= r2) continue;
      if (dot3(dd, dd) < 1e-14f) {
        uint32_t seed = hash3(id, d.st->step, k);
        const float g0 = rnd_gauss(seed);
        const float g1 = rnd_gauss(seed);
        const float g2 = rnd_gauss(seed);
        dd = cross3(adds(ab, 1e-9f), b3{g0, g1, g2});
      }
      const b3 dir = norm3(dd);
      const float r = __builtin_sqrtf(r2);
      X = on + dir * r;
      const float wall = s[1].w / (2.0f * r);
      V = V + dir * fmax_(wall - dot3(V, dir), 0.0f);
    }
  }
  x[0] = X.x, x[1] = X.y, x[2] = X.z;
  v[0] = V.x, v[1] = V.y, v[2] = V.z;
}

// g2p_one (fluid.wgsl:806-861); returns the particle's speed, 0 once it is gone.
template <bool CAV>
__device__ __forceinline__ float g2p_one(const Dev& d, const float* tg, const uint32_t* links, uint32_t i, const int corner[3]) {
  const natray_params* U = d.U;
  const float hh = U->origin[3];
  float4* P = reinterpret_cast<float4*>(d.P) + (size_t)i * 5;
  const float4 pos = P[0];
  const float c1w = P[3].w;
  const float X[3] = {(pos.x - U->origin[0]) / hh, (pos.y - U->origin[1]) / hh, (pos.z - U->origin[2]) / hh};
  const int own[3] = {(int)floorf(X[0]), (int)floorf(X[1]), (int)floorf(X[2])};
  if ((unsigned)(own[0] - corner[0] - 1) > 3u || (unsigned)(own[1] - corner[1] - 1) > 3u ||
      (unsigned)(own[2] - corner[2] - 1) > 3u)
    return 0.0f;  // not in this block: only possible if the particle was moved behind the library's back
  const Gathered gu = gather<0, 1, 1>(tg, X, corner, NATRAY_A_MOM(0), NATRAY_A_WT(0), hh);
  const Gathered gv = gather<1, 0, 1>(tg, X, corner, NATRAY_A_MOM(1), NATRAY_A_WT(1), hh);
  const Gathered gw = gather<1, 1, 0>(tg, X, corner, NATRAY_A_MOM(2), NATRAY_A_WT(2), hh);
  const Gathered gp = gather<1, 1, 1>(tg, X, corner, NATRAY_A_PC, NATRAY_A_WC, hh);
  float v[3] = {gu.s, gv.s, gw.s};
  const float m_near = gu.w + gv.w + gw.w;
  const float step = dt_of(d);
  float x[3] = {pos.x + step * v[0], pos.y + step * v[1], pos.z + step * v[2]};
  if (CAV) carve_cavity(d, x, v, i);
  if (d.sc->flags & NATRAY_SCENE_CAPSULE) carve(d, x, v);
  const float spacing = U->fluid[3];
  float lo[3], hi[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    lo[k] = U->origin[k] + 3.5f * hh;
    hi[k] = U->origin[k] + ((float)U->dims[k] - 4.5f) * hh;
  }
  const bool escaped = x[0] < lo[0] || x[2] < lo[2] || x[0] > hi[0] || x[2] > hi[2] || x[1] > hi[1];
  const float fill = m_near / 24.0f;
  const bool on_ground = x[1] < 1.5f * spacing;
  const float r_thin = 0.75f * spacing;
  const float t_thin = 3.0f * __builtin_sqrtf(U->fluid[0] * r_thin * r_thin * r_thin / NATRAY_SIGMA_W);
  // `covered` only matters below fill 0.45 (the thin test and, below 0.2, the removal), so the
  // cell's mark is read only there.
  bool covered = false;
  if (fill < 0.45f) {
    const uint32_t own_cell =
        tile_cell(links, (uint32_t)((own[0] - corner[0]) * 49 + (own[1] - corner[1]) * 7 + (own[2] - corner[2])));
    covered = own_cell != NONE && accr(d, NATRAY_A_MA, own_cell) > 0;
    if ((d.sc->flags & NATRAY_SCENE_MASK) && in_mask(d, own[0], own[1], own[2])) covered = true;
  }
  const float thin_age = (fill < 0.45f && !on_ground && !covered) ? c1w + step : 0.0f;
  if (escaped || (!covered && fill < 0.2f) || thin_age > t_thin) {
    if (d.spray) {
      // Too thin for the grid: on the ground it soaks in, elsewhere it breaks up into spray.
      if (on_ground) {
        ground_touch(d, x[0], x[2]);
      } else {
        const uint32_t k = atomicAdd(&d.bar[12], 1u);  // below N: a particle dies once a step
        float4* e = reinterpret_cast<float4*>(d.breakups) + (size_t)k * 2;
        e[0] = make_float4(__uint_as_float(i), x[0], x[1], x[2]);
        e[1] = make_float4(v[0], v[1], v[2], 0.0f);
      }
    }
    P[1] = make_float4(v[0], v[1], v[2], 0.0f);
    d.key[i] = NONE;
    return 0.0f;
  }
  float xx[3] = {x[0], x[1], x[2]}, vv[3] = {v[0], v[1], v[2]};
  if (xx[1] < 0.3f * spacing) {
    xx[1] = 0.3f * spacing;
    vv[1] = fmax_(vv[1], 0.0f);
  }
  if (d.spray && xx[1] < 0.7f * spacing) ground_touch(d, xx[0], xx[2]);
  const float cs = 4.0f / (hh * hh);
  P[0] = make_float4(xx[0], xx[1], xx[2], pos.w);
  P[1] = make_float4(vv[0], vv[1], vv[2], 1.0f);
  P[2] = make_float4(gu.g0 * cs, gu.g1 * cs, gu.g2 * cs, gp.w > 1e-6f ? gp.s / gp.w : 0.0f);
  P[3] = make_float4(gv.g0 * cs, gv.g1 * cs, gv.g2 * cs, thin_age);
  P[4] = make_float4(gw.g0 * cs, gw.g1 * cs, gw.g2 * cs, 0.0f);
  d.key[i] = pack_key(d, xx[0], xx[1], xx[2], true);
  return __builtin_sqrtf((vv[0] * vv[0] + vv[1] * vv[1]) + vv[2] * vv[2]);
}

template <class L, bool CAV = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_g2p(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  __shared__ float tg[8 * TILE];
  __shared__ uint32_t links[27];
  const uint32_t n = n_blocks(d), tid = threadIdx.x;
  float fastest = 0.0f;
  Dealer deal_s;
  __shared__ uint32_t deal_w[2];
  for (uint32_t b = deal_first(d, DEAL_G2P, n, v.wg, v.wgs, deal_s, deal_w); b < n; b = deal_next(deal_w)) {
    if (tid < 27) links[tid] = d.block_nb[b * 27u + tid];
    const uint32_t first = d.block_start[b], count = d.block_start[b + 1] - first;
    __syncthreads();
    deal_ahead(d, DEAL_G2P, n, v.wgs, deal_s, deal_w);
    if (count == 0u) continue;
    for (uint32_t t = tid; t < TILE; t += 256) {
      const uint32_t cell = tile_cell(links, t);
      const bool water = cell != NONE && cfr(d, NATRAY_F_TYPE, cell) == WATER;
#pragma unroll
      for (uint32_t k = 0; k < 3; ++k) {
        float u = 0.0f, w = 0.0f;
        if (cell != NONE) {
          u = cfr(d, NATRAY_F_U + k, cell);
          w = from_fp(accr(d, NATRAY_A_WT(k), cell), NATRAY_FP_W);
        }
        tg[NATRAY_A_MOM(k) * TILE + t] = u;
        tg[NATRAY_A_WT(k) * TILE + t] = w;
      }
      tg[NATRAY_A_PC * TILE + t] = water ? pressure_of(d, cell) : 0.0f;
      tg[NATRAY_A_WC * TILE + t] = water ? 1.0f : 0.0f;
    }
    __syncthreads();
    int bx, by, bz;
    slot_coord(d.U, d.block_list[b], bx, by, bz);
    const int corner[3] = {bx * 4 - 1, by * 4 - 1, bz * 4 - 1};
    for (uint32_t k = tid; k < count; k += 256) fastest = fmax_(fastest, g2p_one<CAV>(d, tg, links, d.sorted[first + k], corner));
  }
  // The workgroup's fastest particle goes to its own slot; k_vmax takes the maximum. One
  // atomic per wave on State.max_speed would serialise thousands of device-scope atomics on
  // one word.
  __shared__ float s_fast[4];
  for (int off = 32; off > 0; off >>= 1) fastest = fmax_(fastest, __shfl_xor(fastest, off));
  if (lane_id() == 0) s_fast[tid >> 6] = fastest;
  __syncthreads();
  if (tid == 0) d.vmax_part[v.wg] = fmax_(fmax_(s_fast[0], s_fast[1]), fmax_(s_fast[2], s_fast[3]));
}

// The step's fastest particle from k_g2p's per-workgroup maxima (the atomicMax of
// fluid.wgsl:898-900 as a reduction).
template <class L>
__global__ __launch_bounds__(256) void k_vmax(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const uint32_t parts = d.g2p_wgs;
  __shared__ float s[4];
  float m = 0.0f;
  for (uint32_t k = threadIdx.x; k < parts; k += 256) m = fmax_(m, d.vmax_part[k]);
  for (int off = 32; off > 0; off >>= 1) m = fmax_(m, __shfl_xor(m, off));
  if (lane_id() == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) d.st->max_speed = fmax_(fmax_(s[0], s[1]), fmax_(s[2], s[3]));
}

// ---- membrane coupling (membrane.wgsl:151-256, with grid_prep's RUBBER/COVERED cells) -----------
//
// The membrane's own step (its forces and substeps) is not part of this library: its vertices
// come from the caller each step (natray_sim_set_membrane). These kernels are what the water
// substep runs with them: claiming the blocks around rubber off the water (mark_rubber), the
// rubber's pull and motion splatted to the grid (mem_splat, read by grid_prep), and the water
// sampled at each vertex after the extension (mem_sample).

constexpr float CONTACT = 0.05f;  // rubber this deep in water is on it (membrane.wgsl:38)

__device__ __forceinline__ uint32_t mem_count(const Dev& d) {
  const uint32_t n = d.U->mem[0];
  return n < d.mem_cap ? n : d.mem_cap;
}
__device__ __forceinline__ bool on_water(const Dev& d, uint32_t i) {
  return d.mem_mass[i] > 0.0f && d.mem_water[i].w >= CONTACT;
}
// The cell of a point, floor((p - origin) / h) per axis (own_cell, fluid.wgsl:127-129).
__device__ __forceinline__ void point_cell(const Dev& d, float4 p, int& cx, int& cy, int& cz) {
  const natray_params* U = d.U;
  const float h = U->origin[3];
  cx = (int)floorf((p.x - U->origin[0]) / h);
  cy = (int)floorf((p.y - U->origin[1]) / h);
  cz = (int)floorf((p.z - U->origin[2]) / h);
}
// The storage index of the cell at (cx, cy, cz), or NONE outside the domain or where no block
// of this step holds it (cell_at, membrane.wgsl:156-162).
__device__ __forceinline__ uint32_t cell_at(const Dev& d, int cx, int cy, int cz) {
  const natray_params* U = d.U;
  if (cx < 0 || cy < 0 || cz < 0 || cx >= (int)U->dims[0] || cy >= (int)U->dims[1] || cz >= (int)U->dims[2]) return NONE;
  const uint32_t b = d.block_ptr[block_slot(U, cx >> 2, cy >> 2, cz >> 2)];
  if (b >= n_blocks(d)) return NONE;
  return b * 64u + (uint32_t)(cx & 3) * 16u + (uint32_t)(cy & 3) * 4u + (uint32_t)(cz & 3);
}

// mark_rubber (membrane.wgsl:164-180): the blocks of cells c-2..c+3 around rubber off the
// water are claimed, so the air there is simulated too.
template <class L>
__global__ __launch_bounds__(256) void k_mark_rubber(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const uint32_t i = v.wg * blockDim.x + threadIdx.x;
  if (i >= mem_count(d) || d.mem_push[i].y <= 0.0f || on_water(d, i)) return;
  int cx, cy, cz;
  point_cell(d, d.mem_x[i], cx, cy, cz);
  const natray_params* U = d.U;
  const int bmx = (int)U->bdims[0] - 1, bmy = (int)U->bdims[1] - 1, bmz = (int)U->bdims[2] - 1;
  const int lx = max(0, min((cx - 2) >> 2, bmx)), ly = max(0, min((cy - 2) >> 2, bmy)), lz = max(0, min((cz - 2) >> 2, bmz));
  const int hx = max(0, min((cx + 3) >> 2, bmx)), hy = max(0, min((cy + 3) >> 2, bmy)), hz = max(0, min((cz + 3) >> 2, bmz));
  for (int x = lx; x <= hx; ++x)
    for (int y = ly; y <= hy; ++y)
      for (int z = lz; z <= hz; ++z) {
        const uint32_t s = block_slot(U, x, y, z);
        if (d.block_ptr[s] != PENDING) d.block_ptr[s] = PENDING;
      }
}

// mem_splat (membrane.wgsl:182-216): every vertex with area adds its area (mm^2) and area
// times velocity to its own cell (A_MW, A_MV); a vertex on the water also spreads its push
// (A_MF) and area (A_MA) over the 27 cells around it with the particles' B-spline weights.
template <class L>
__global__ __launch_bounds__(256) void k_mem_splat(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const uint32_t i = v.wg * blockDim.x + threadIdx.x;
  if (i >= mem_count(d) || d.mem_mass[i] == 0.0f) return;
  const float2 pa = d.mem_push[i];
  if (pa.y <= 0.0f) return;
  const float4 p = d.mem_x[i];
  int cx, cy, cz;
  point_cell(d, p, cx, cy, cz);
  const uint32_t own = cell_at(d, cx, cy, cz);
  if (own != NONE) {
    const float a = pa.y * 1e6f;
    const float4 mv = d.mem_v[i];
    atomicAdd(&accr(d, NATRAY_A_MW, own), to_fp(a, NATRAY_FP_A));
    atomicAdd(&accr(d, NATRAY_A_MV + 0u, own), to_fp(a * mv.x, NATRAY_FP_A));
    atomicAdd(&accr(d, NATRAY_A_MV + 1u, own), to_fp(a * mv.y, NATRAY_FP_A));
    atomicAdd(&accr(d, NATRAY_A_MV + 2u, own), to_fp(a * mv.z, NATRAY_FP_A));
  }
  if (!on_water(d, i)) return;
  const natray_params* U = d.U;
  const float h = U->origin[3];
  const float x[3] = {(p.x - U->origin[0]) / h - 0.5f, (p.y - U->origin[1]) / h - 0.5f, (p.z - U->origin[2]) / h - 0.5f};
  int base[3];
  float w[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float bf = floorf(x[k] - 0.5f);
    base[k] = (int)bf;
    axis_weights(x[k] - bf, w[k]);
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        const uint32_t cell = cell_at(d, base[0] + a, base[1] + b, base[2] + c);
        if (cell == NONE) continue;
        const float wt = w[0][a] * w[1][b] * w[2][c];
        atomicAdd(&accr(d, NATRAY_A_MF, cell), to_fp(wt * pa.x, NATRAY_FP_F));
        atomicAdd(&accr(d, NATRAY_A_MA, cell), to_fp(wt * pa.y * 1e6f, NATRAY_FP_A));
      }
}

// sample_grid (membrane.wgsl:218-237): the B-spline sums around x (in cells, on the lattice)
// of cell field fu weighted by transfer field fw, and of the weights.
__device__ __forceinline__ float2 sample_grid(const Dev& d, const float x[3], uint32_t fu, uint32_t fw) {
  int base[3];
  float w[3][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float bf = floorf(x[k] - 0.5f);
    base[k] = (int)bf;
    axis_weights(x[k] - bf, w[k]);
  }
  float sx = 0.0f, sy = 0.0f;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        const uint32_t cell = cell_at(d, base[0] + a, base[1] + b, base[2] + c);
        if (cell == NONE) continue;
        const float wt = w[0][a] * w[1][b] * w[2][c] * from_fp(accr(d, fw, cell), NATRAY_FP_W);
        sx = sx + wt * cfr(d, fu, cell);
        sy = sy + wt;
      }
  return make_float2(sx, sy);
}

// mem_sample (membrane.wgsl:239-254): the velocity of the faces around each vertex, weighted
// by the water they carry, and how full of water the cells around it are (a bulk cell holds 8
// particles whose weights sum to 1 each).
template <class L>
__global__ __launch_bounds__(256) void k_mem_sample(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u) return;
  const uint32_t i = v.wg * blockDim.x + threadIdx.x;
  if (i >= mem_count(d) || d.mem_mass[i] == 0.0f) return;
  const natray_params* U = d.U;
  const float h = U->origin[3];
  const float4 p = d.mem_x[i];
  const float x[3] = {(p.x - U->origin[0]) / h, (p.y - U->origin[1]) / h, (p.z - U->origin[2]) / h};
  float vel[3];
#pragma unroll
  for (uint32_t dim = 0; dim < 3; ++dim) {
    const float xs[3] = {dim == 0 ? x[0] : x[0] - 0.5f, dim == 1 ? x[1] : x[1] - 0.5f, dim == 2 ? x[2] : x[2] - 0.5f};
    const float2 s = sample_grid(d, xs, NATRAY_F_U + dim, NATRAY_A_WT(dim));
    vel[dim] = s.y > 1e-6f ? s.x / s.y : 0.0f;
  }
  const float xc[3] = {x[0] - 0.5f, x[1] - 0.5f, x[2] - 0.5f};
  const float fill = sample_grid(d, xc, NATRAY_F_TYPE, NATRAY_A_WC).y / 8.0f;
  d.mem_water[i] = make_float4(vel[0], vel[1], vel[2], fill);
}

// ---- bullets: their step and the water's push (bullet.wgsl:100-386) ------------------------------
//
// One workgroup of 256 per bullet (the reference's own shape): each thread takes patches k,
// k + 256, ..., the workgroup sums their forces in the reference's tree, and thread 0 steps the
// body and records the cavity section the nose opened. The rubber's impulses and held rubber
// (bullet_hit, written by rubber_contact in the membrane's substeps) are zero here.

constexpr float LOGVINOVICH_K = 2.0f, CP_WAKE = -0.4f, CF_SKIN = 0.003f;

// cp_patch (bullet.wgsl:100-111)
__device__ __forceinline__ float cp_patch(float s) {
  if (s < 0.0f) return CP_WAKE * s * s;
  const float beta = nm_asinf(fmin_(s, 1.0f));
  if (beta < PI_F / 12.0f) return 2.13f * s * s;
  const float x = beta / PI_F - 0.25f;
  return 0.5f + 1.81f * x - 2.0f * x * x;
}

// bullet_step (bullet.wgsl:179-377) in two kernels with the reference's bits. The patches are
// latency-bound (each samples the grid at about 250 stencil cells, two dependent loads deep), so
// k_bullet_patches gives every patch a wave: its 64 lanes load the 4^3 cells that hold all the
// patch's stencils at once, and each stencil's sums are taken from the lanes in the sequential
// loop's order (readlane), so they round as the loop does. Each patch leaves a record;
// k_bullet_body sums the records the way the reference's workgroup of 256 does (lane li takes
// patches li, li + 256, ... in order, then the tree) and steps the body.

// The 27 cells of a stencil around x (in cells, on a lattice), lane l < 27 taking cell
// (l / 9, l / 3 % 3, l % 3): its B-spline weights and its index (NONE where there is none).
struct Stencil {
  int base[3];
  float w[3][3];
};
__device__ __forceinline__ Stencil stencil_of(const float x[3]) {
  Stencil st;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float bf = floorf(x[k] - 0.5f);
    st.base[k] = (int)bf;
    axis_weights(x[k] - bf, st.w[k]);
  }
  return st;
}
__device__ __forceinline__ uint32_t stencil_cell(const Dev& d, const Stencil& st, uint32_t lane, int& a, int& b, int& c) {
  a = (int)(lane / 9u), b = (int)((lane / 3u) % 3u), c = (int)(lane % 3u);
  if (lane >= 27u) return NONE;
  return cell_at(d, st.base[0] + a, st.base[1] + b, st.base[2] + c);
}
// sum over lanes 0..26 in lane order of (v, w) where `present`: the bits of the sequential
// sx = sx + v, sy = sy + w over the stencil. Wave-uniform result.
__device__ __forceinline__ float2 ordered_sum27(float v, float w, bool present) {
  const unsigned long long m = __ballot(present);
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i)
    if ((m >> i) & 1ull) {
      sx = sx + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
      sy = sy + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), i));
    }
  return make_float2(sx, sy);
}
// sample_pressure (bullet.wgsl:113-134) by one wave.
__device__ __forceinline__ float2 wave_sample_pressure(const Dev& d, const float x[3], bool gas_only, uint32_t lane) {
  const Stencil st = stencil_of(x);
  int a, b, c;
  const uint32_t cell = stencil_cell(d, st, lane, a, b, c);
  bool present = false;
  float v = 0.0f, w = 0.0f;
  if (cell != NONE) {
    const float ty = cfr(d, NATRAY_F_TYPE, cell);
    present = !(ty == SOLID || (gas_only && ty == WATER));
    if (present) {
      w = st.w[0][a] * st.w[1][b] * st.w[2][c];
      v = w * cfr(d, NATRAY_F_P, cell);
    }
  }
  return ordered_sum27(v, w, present);
}
// in_cavity (bullet.wgsl:156-170) by one wave: lane l < 32 tests section l.
__device__ __forceinline__ bool wave_in_cavity(const Dev& d, uint32_t bi, b3 p, float t, uint32_t lane) {
  bool hit = false;
  if (lane < NATRAY_BULLET_SECTIONS) {
    const float4* s = section_of(d, bi, lane);
    if (!(s[1].w <= 0.0f || section_r2(s, t) >= hand2_of(d.U->origin[3]))) {
      const b3 a = xyz(s[0]), ab = xyz(s[1]) - a;
      const float f = dot3(p - a, ab) / fmax_(dot3(ab, ab), 1e-20f);
      if (!(f < 0.0f || f > 1.0f)) {
        const b3 dd = p - (a + ab * f);
        hit = dot3(dd, dd) < section_r2(s, t);
      }
    }
  }
  return __ballot(hit) != 0ull;
}

// The 4^3 cells around a patch hold every stencil it samples: the three face lattices, the
// cell centres and the pushes. Per axis those stencils' bases are floor(X - 1) or one more, so a
// wave loads the neighbourhood once, lane l taking cell b0 + (l >> 4, (l >> 2) & 3, l & 3), and
// each stencil is a 3^3 sub-block of it at offsets 0 or 1 per axis.
struct Sub {
  int off[3];
  float w[3][3];
};
__device__ __forceinline__ Sub sub_of(const float xs[3], const int b0[3]) {
  Sub st;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float bf = floorf(xs[k] - 0.5f);
    st.off[k] = __builtin_amdgcn_readfirstlane((int)bf - b0[k]);
    axis_weights(xs[k] - bf, st.w[k]);
  }
  return st;
}
__device__ __forceinline__ float pick3(const float w[3], int i) { return i == 0 ? w[0] : (i == 1 ? w[1] : w[2]); }
// This lane's (a, b, c) in the sub-block, if it is in it.
__device__ __forceinline__ bool in_sub(const Sub& st, uint32_t lane, int& a, int& b, int& c) {
  a = (int)(lane >> 4) - st.off[0];
  b = (int)((lane >> 2) & 3u) - st.off[1];
  c = (int)(lane & 3u) - st.off[2];
  return (unsigned)a < 3u && (unsigned)b < 3u && (unsigned)c < 3u;
}
// The sums over the sub-block's cells where `present`, in the sequential loop's (a, b, c) order.
__device__ __forceinline__ float2 sub_sum(const Sub& st, float v, float w, bool present) {
  const unsigned long long m = __ballot(present);
  float sx = 0.0f, sy = 0.0f;
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const int ln = ((i / 9 + st.off[0]) << 4) | (((i / 3) % 3 + st.off[1]) << 2) | (i % 3 + st.off[2]);
    if ((m >> ln) & 1ull) {
      sx = sx + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), ln));
      sy = sy + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(w), ln));
    }
  }
  return make_float2(sx, sy);
}

// Whether bullet b takes part in this step (it joins at the first step that starts once it has
// arrived), as bullet_step's first lines decide.
__device__ __forceinline__ uint32_t bullet_state(const Dev& d, const natray_bullet& b) {
  uint32_t state = b.info[2];
  if (state == 0u && d.st->t_ns - d.st->dt_ns >= b.info[0]) state = 1u;
  return state;
}

// Words of a patch record: fp.xyz, 1 when the patch is on the wetted face; r x fp, its drag;
// its face area across the path, then p times it (xyz); r . vhat.
constexpr uint32_t PATCH_REC = 12;

// The patches of bullet_step (bullet.wgsl:226-284): one wave per patch, NATRAY_MAX_BULLETS x
// 144 workgroups of 4 waves.
template <class L>
__global__ __launch_bounds__(256) void k_bullet_patches(L l) {
  const View vw = view(l);
  const Dev& d = vw.d;
  if (d.st->dt_ns == 0u) return;
  const natray_params* U = d.U;
  const uint32_t lane = lane_id();
  const uint32_t per = NATRAY_BULLET_PATCHES / 4u;
  const uint32_t bi = vw.wg / per, k = (vw.wg % per) * 4u + (threadIdx.x >> 6);
  if (bi >= d.n_bullets) return;
  const natray_bullet& bref = d.bullets[bi];
  if (k >= bref.info[1] || bullet_state(d, bref) != 1u) return;
  const float step = dt_of(d), t_end = now_of(d), hh = U->origin[3];
  const b3 origin = xyz(U->origin), bx = xyz(bref.x), bv = xyz(bref.v);
  // The pressure ahead of it (bullet.wgsl:194-196).
  float pf;
  {
    const b3 ahead = bx + norm3(adds(bv, 1e-9f)) * (bref.shape[2] + 3.0f * hh);
    const b3 xa = (ahead - origin) / hh;
    const float xs[3] = {xa.x - 0.5f, xa.y - 0.5f, xa.z - 0.5f};
    const float2 sp = wave_sample_pressure(d, xs, false, lane);
    pf = sp.y > 1e-6f ? sp.x / sp.y : 0.0f;
  }
  const float4 q = make_float4(bref.q[0], bref.q[1], bref.q[2], bref.q[3]);
  const b3 lb = qrot(make_float4(-q.x, -q.y, -q.z, q.w), xyz(bref.l));
  const b3 w = qrot(q, b3{lb.x / bref.shape[1], lb.y / bref.shape[1], lb.z / bref.shape[0]});
  const float rho = U->fluid[0];
  const b3 vhat = norm3(adds(bv, 1e-9f));
  const float4* pts = reinterpret_cast<const float4*>(d.patches) + (size_t)bi * NATRAY_BULLET_PATCHES * 2;
  const float4 pp = pts[2 * k], pn = pts[2 * k + 1];
  const b3 r = qrot(q, xyz(pp));
  const b3 p = bx + r;
  const b3 n = qrot(q, xyz(pn));
  const float area = pp.w;
  const b3 Xv = (p - origin) / hh;
  const float X[3] = {Xv.x, Xv.y, Xv.z};
  const float xc[3] = {X[0] - 0.5f, X[1] - 0.5f, X[2] - 0.5f};
  // The neighbourhood: one cell per lane, and the fields every stencil reads there.
  int b0[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) b0[k] = __builtin_amdgcn_readfirstlane((int)floorf(xc[k] - 0.5f) < (int)floorf(X[k] - 0.5f)
                                                                      ? (int)floorf(xc[k] - 0.5f)
                                                                      : (int)floorf(X[k] - 0.5f));
  const uint32_t cell = cell_at(d, b0[0] + (int)(lane >> 4), b0[1] + (int)((lane >> 2) & 3u), b0[2] + (int)(lane & 3u));
  const bool here = cell != NONE;
  int32_t wt_i[3] = {0, 0, 0}, wc_i = 0;
  float u_c[3] = {0.0f, 0.0f, 0.0f}, ty = 0.0f, p_c = 0.0f;
  if (here) {
#pragma unroll
    for (uint32_t dim = 0; dim < 3; ++dim) {
      wt_i[dim] = accr(d, NATRAY_A_WT(dim), cell);
      u_c[dim] = cfr(d, NATRAY_F_U + dim, cell);
    }
    wc_i = accr(d, NATRAY_A_WC, cell);
    ty = cfr(d, NATRAY_F_TYPE, cell);
    p_c = cfr(d, NATRAY_F_P, cell);
  }
  // sample_grid (membrane.wgsl:218-237) on the three face lattices.
  float uw[3], sw[3];
  Sub face[3];
#pragma unroll
  for (uint32_t dim = 0; dim < 3; ++dim) {
    const float xs[3] = {X[0] - (dim == 0 ? 0.0f : 0.5f), X[1] - (dim == 1 ? 0.0f : 0.5f), X[2] - (dim == 2 ? 0.0f : 0.5f)};
    face[dim] = sub_of(xs, b0);
    int a, b, c;
    float v = 0.0f, wgt = 0.0f;
    if (here && in_sub(face[dim], lane, a, b, c)) {
      wgt = pick3(face[dim].w[0], a) * pick3(face[dim].w[1], b) * pick3(face[dim].w[2], c) * from_fp(wt_i[dim], NATRAY_FP_W);
      v = wgt * u_c[dim];
    }
    const float2 sg = sub_sum(face[dim], v, wgt, here);
    uw[dim] = sg.y > 1e-6f ? sg.x / sg.y : 0.0f;
    sw[dim] = sg.y;
  }
  // The water around it (sample_grid of F_TYPE weighted by A_WC) and the gas's pressure
  // (sample_pressure over GAS cells), both on the cell centres.
  const Sub cen_s = sub_of(xc, b0);
  float fill, p_gas;
  {
    int a, b, c;
    float wf = 0.0f, vf = 0.0f, wg = 0.0f, vg = 0.0f;
    const bool in = in_sub(cen_s, lane, a, b, c);
    const bool gas = here && !(ty == SOLID || ty == WATER);
    if (here && in) {
      const float w3 = pick3(cen_s.w[0], a) * pick3(cen_s.w[1], b) * pick3(cen_s.w[2], c);
      wf = w3 * from_fp(wc_i, NATRAY_FP_W);
      vf = wf * ty;
      if (gas) {
        wg = w3;
        vg = w3 * p_c;
      }
    }
    fill = fmin_(fmax_(sub_sum(cen_s, vf, wf, here).y / 8.0f, 0.0f), 1.0f);
    const float2 gs = sub_sum(cen_s, vg, wg, gas);
    p_gas = gs.y > 1e-6f ? gs.x / gs.y : 0.0f;
  }
  if (wave_in_cavity(d, bi, p, t_end - step, lane)) fill = 0.0f;
  const b3 vp = bv + cross3(w, r);
  const b3 rel = vp - b3{uw[0], uw[1], uw[2]};
  const float sp = len3(rel);
  const float wn = dot3(rel, n);
  const float qd = 0.5f * rho * sp * sp;
  const float p_water = fmax_(NATRAY_P_FLOOR, pf + qd * cp_patch(wn / fmax_(sp, 1e-9f)));
  // The wet share of the patch: pressure, and skin friction along the surface.
  b3 fw = n * (-fill * p_water * area);
  const b3 rt = rel - n * wn;
  const float rtl = len3(rt);
  if (rtl > 1e-9f) fw = fw - (rt * (fill * CF_SKIN * 0.5f * rho * rtl)) * area;
  const b3 fp = fw - n * ((1.0f - fill) * fmax_(p_gas, NATRAY_P_FLOOR) * area);
  // push_water (bullet.wgsl:136-150) on the three face lattices: integer atomics, any order.
  {
    const float hs = U->fluid[3];
    const float fwd[3] = {fw.x, fw.y, fw.z};
#pragma unroll
    for (uint32_t dim = 0; dim < 3; ++dim) {
      const float j = -fwd[dim] * step;
      if (sw[dim] <= 1e-6f || j == 0.0f) continue;
      const float kk = j / (U->fluid[0] * hs * hs * hs * sw[dim]);
      int a, b, c;
      if (here && wt_i[dim] > 0 && in_sub(face[dim], lane, a, b, c))
        atomicAdd(&accr(d, NATRAY_A_BU + dim, cell),
                  to_fp(kk * pick3(face[dim].w[0], a) * pick3(face[dim].w[1], b) * pick3(face[dim].w[2], c), NATRAY_FP_M));
    }
  }
  if (lane == 0) {
    const b3 tq = cross3(r, fp);
    const bool wet_face = wn > 0.0f && fill > 0.0f;
    float face = 0.0f, drag = 0.0f, rv = 0.0f;
    b3 pface = {0.0f, 0.0f, 0.0f};
    if (wet_face) {
      face = fill * area * fmax_(dot3(n, vhat), 0.0f);
      drag = -dot3(fw, vhat);
      pface = p * face;
      rv = dot3(r, vhat);
    }
    float4* rec = reinterpret_cast<float4*>(d.bullet_rec) + ((size_t)bi * NATRAY_BULLET_PATCHES + k) * (PATCH_REC / 4);
    rec[0] = make_float4(fp.x, fp.y, fp.z, wet_face ? 1.0f : 0.0f);
    rec[1] = make_float4(tq.x, tq.y, tq.z, drag);
    rec[2] = make_float4(face, pface.x, pface.y, pface.z);
    // r . vhat rides in the spare word of the next record's slot group: a 13th word.
    d.bullet_rec[(size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_PATCHES * PATCH_REC + (size_t)bi * NATRAY_BULLET_PATCHES + k] = rv;
  }
}

// The body of bullet_step (bullet.wgsl:286-377): the records summed as the reference's
// workgroup sums its patches, the body stepped, the cavity section recorded and the thin
// sections' volume source splatted. One workgroup of 256 per bullet.
template <class L>
__global__ __launch_bounds__(256) void k_bullet_body(L l) {
  const View vw = view(l);
  const Dev& d = vw.d;
  if (d.st->dt_ns == 0u) return;
  const natray_params* U = d.U;
  const uint32_t bi = vw.wg, li = threadIdx.x;
  natray_bullet* B = d.bullets + bi;
  const natray_bullet b = *B;
  const uint32_t state = bullet_state(d, b);
  __syncthreads();  // every thread has read the bullet before thread 0 writes it
  float4* pose = reinterpret_cast<float4*>(d.pose_log) + ((size_t)bi * NATRAY_BULLET_LOG + d.st->step % NATRAY_BULLET_LOG) * 3;
  if (state != 1u) {
    if (li == 0) {
      B->info[2] = state;
      pose[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    return;
  }
  const float step = dt_of(d), t_end = now_of(d), hh = U->origin[3];
  const b3 origin = xyz(U->origin), bx = xyz(b.x), bv = xyz(b.v);
  const float rho = U->fluid[0];
  const b3 vhat = norm3(adds(bv, 1e-9f));
  __shared__ float4 red_f[256], red_t[256], red_c[256];
  b3 f = {0.0f, 0.0f, 0.0f}, tq = {0.0f, 0.0f, 0.0f}, cen = {0.0f, 0.0f, 0.0f};
  float cav_x = 0.0f, cav_y = 0.0f, cen_w = 3.0e38f;
  const float4* recs = reinterpret_cast<const float4*>(d.bullet_rec) + (size_t)bi * NATRAY_BULLET_PATCHES * (PATCH_REC / 4);
  const float* rvs = d.bullet_rec + (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_PATCHES * PATCH_REC + (size_t)bi * NATRAY_BULLET_PATCHES;
  for (uint32_t k = li; k < b.info[1]; k += 256u) {
    const float4 r0 = recs[3 * k], r1 = recs[3 * k + 1], r2 = recs[3 * k + 2];
    f = f + xyz(r0);
    tq = tq + xyz(r1);
    if (r0.w != 0.0f) {
      cav_x = cav_x + r1.w;
      cav_y = cav_y + r2.x;
      cen = cen + b3{r2.y, r2.z, r2.w};
      cen_w = fmin_(cen_w, rvs[k]);
    }
  }
  red_f[li] = make_float4(f.x, f.y, f.z, cav_x);
  red_t[li] = make_float4(tq.x, tq.y, tq.z, cav_y);
  red_c[li] = make_float4(cen.x, cen.y, cen.z, cen_w);
  __syncthreads();
  for (uint32_t off = 128u; off > 0u; off >>= 1u) {
    if (li < off) {
      const float4 a = red_f[li], c = red_f[li + off];
      red_f[li] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w + c.w);
      const float4 e = red_t[li], g = red_t[li + off];
      red_t[li] = make_float4(e.x + g.x, e.y + g.y, e.z + g.z, e.w + g.w);
      const float4 m = red_c[li], o = red_c[li + off];
      red_c[li] = make_float4(m.x + o.x, m.y + o.y, m.z + o.z, fmin_(m.w, o.w));
    }
    __syncthreads();
  }
  const float hand2 = hand2_of(hh);
  // Thread 0 steps the body; the neighbour-block test, the reach over the sections and the
  // sections' source run on many lanes (an or, a max of non-negative floats and integer
  // atomics: the same values as the sequential loops).
  __shared__ float s_x[3];
  __shared__ uint32_t s_reach;
  const float4 rf = red_f[0], rt4 = red_t[0], rc = red_c[0];
  const float mass = b.x[3];
  // The rubber's impulses (bullet_hit) are zero: rubber_contact is not part of this library.
  const b3 hit = {0.0f, 0.0f, 0.0f}, hit_l = {0.0f, 0.0f, 0.0f};
  const b3 force = xyz(rf) + xyz(U->gravity) * mass;
  const float drag = rf.w, face = rt4.w;
  const b3 v = bv + (force * step + hit) / mass;
  const b3 x = bx + v * step;
  if (li == 0) {
    s_x[0] = x.x, s_x[1] = x.y, s_x[2] = x.z;
    s_reach = 0u;
  }
  // Out of reach of any water it can go a block per step: it cannot miss one on the way,
  // since the blocks around it are checked every step.
  bool nb_in_use = false;
  if (li < 27u) {
    const b3 xc = (x - origin) / hh;
    const int k = (int)li;
    const int nb[3] = {((int)floorf(xc.x) >> 2) + k / 9 - 1, ((int)floorf(xc.y) >> 2) + (k / 3) % 3 - 1,
                       ((int)floorf(xc.z) >> 2) + k % 3 - 1};
    nb_in_use = nb[0] >= 0 && nb[1] >= 0 && nb[2] >= 0 && nb[0] < (int)U->bdims[0] && nb[1] < (int)U->bdims[1] &&
                nb[2] < (int)U->bdims[2] && d.block_ptr[block_slot(U, nb[0], nb[1], nb[2])] < n_blocks(d);
  }
  const bool near = __syncthreads_or(nb_in_use) != 0;
  float4* secs = reinterpret_cast<float4*>(d.sections) + (size_t)bi * NATRAY_BULLET_SECTIONS * 3;
  uint32_t info[4] = {b.info[0], b.info[1], 1u, b.info[3]};
  b3 lnew = {0.0f, 0.0f, 0.0f};
  float4 qn = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  float roll = 0.0f, pace = 0.0f;
  if (li == 0) {
    lnew = (xyz(b.l) + xyz(rt4) * step) + hit_l;
    // A symmetric top turns about its angular momentum at |L| / Iy and about its own axis at
    // the rest of its spin; taking the two exactly keeps the spin and the cone whatever the step.
    const float4 q = make_float4(b.q[0], b.q[1], b.q[2], b.q[3]);
    const b3 axis0 = qrot(q, b3{0.0f, 0.0f, 1.0f});
    const float spin_extra = dot3(lnew, axis0) * (1.0f / b.shape[0] - 1.0f / b.shape[1]) * step;
    const float cone = len3(lnew) / b.shape[1] * step;
    qn = qmul(q, make_float4(0.0f, 0.0f, nm_sinf(0.5f * spin_extra), nm_cosf(0.5f * spin_extra)));
    if (cone > 1e-12f) {
      const b3 ax = norm3(lnew) * nm_sinf(0.5f * cone);
      qn = qmul(make_float4(ax.x, ax.y, ax.z, nm_cosf(0.5f * cone)), qn);
    }
    const float ql = __builtin_sqrtf(((qn.x * qn.x + qn.y * qn.y) + qn.z * qn.z) + qn.w * qn.w);
    qn = make_float4(qn.x / ql, qn.y / ql, qn.z / ql, qn.w / ql);
    roll = b.v[3] + dot3(lnew, axis0) / b.shape[0] * step;
    const float xa[3] = {x.x, x.y, x.z};
    for (int k = 0; k < 3; ++k)
      if (xa[k] < U->origin[k] || xa[k] > U->origin[k] + (float)U->dims[k] * hh) info[2] = 2u;
    pace = near ? len3(v) : len3(v) * U->fluid[1] / 4.0f;
    if (drag > 0.0f && face > 0.0f) {
      // The face's centre, moved back along the path to where the flow leaves it.
      const b3 mid = xyz(rc) / face;
      const b3 a = mid + vhat * (rc.w - dot3(mid - bx, vhat));
      const b3 e = a + v * step;
      float4* sn = secs + (size_t)(info[3] % NATRAY_BULLET_SECTIONS) * 3;
      sn[0] = make_float4(a.x, a.y, a.z, t_end - 0.5f * step);
      sn[1] = make_float4(e.x, e.y, e.z, 2.0f * __builtin_sqrtf(drag / (PI_F * rho * LOGVINOVICH_K)));
      sn[2] = make_float4(face / PI_F, 0.0f, 0.0f, 0.0f);
      info[3] += 1u;
    }
  }
  __syncthreads();  // the new section is in place
  if (li < NATRAY_BULLET_SECTIONS) {
    const float4* sk = secs + (size_t)li * 3;
    if (!(sk[1].w <= 0.0f || section_r2(sk, t_end) >= hand2)) {
      const float r = fmax_(len3(xyz(sk[0]) - x), len3(xyz(sk[1]) - x));
      atomicMax(&s_reach, __float_as_uint(r));  // r >= 0: its bits order as the floats do
    }
  }
  __syncthreads();
  if (li == 0) {
    const float reach = __uint_as_float(s_reach);
    B->x[0] = x.x, B->x[1] = x.y, B->x[2] = x.z;
    B->v[0] = v.x, B->v[1] = v.y, B->v[2] = v.z, B->v[3] = roll;
    B->q[0] = qn.x, B->q[1] = qn.y, B->q[2] = qn.z, B->q[3] = qn.w;
    B->l[0] = lnew.x, B->l[1] = lnew.y, B->l[2] = lnew.z, B->l[3] = pace;
    B->reach[0] = x.x, B->reach[1] = x.y, B->reach[2] = x.z, B->reach[3] = reach > 0.0f ? reach + R_HAND * hh : 0.0f;
    for (int k = 0; k < 4; ++k) B->info[k] = info[k];
    pose[0] = make_float4(x.x, x.y, x.z, __uint_as_float(d.st->t_ns));
    pose[1] = qn;
    pose[2] = make_float4(v.x, v.y, v.z, roll);
  }
  // Sections thinner than the grid push the water out of the way as they open: one (section,
  // stencil cell) pair per thread.
  for (uint32_t pair = li; pair < NATRAY_BULLET_SECTIONS * 27u; pair += 256u) {
    const uint32_t k = pair / 27u, cidx = pair % 27u;
    const float4* sk = secs + (size_t)k * 3;
    const float r2 = section_r2(sk, t_end);
    if (!(sk[1].w > 0.0f && r2 < hand2)) continue;
    const b3 a = xyz(sk[0]), e = xyz(sk[1]);
    const float qv = PI_F * sk[1].w * len3(e - a);
    const b3 m = ((a + e) * 0.5f - origin) / hh;
    const float xs[3] = {m.x - 0.5f, m.y - 0.5f, m.z - 0.5f};
    const Stencil st = stencil_of(xs);
    const int a3 = (int)(cidx / 9u), b3i = (int)((cidx / 3u) % 3u), c3 = (int)(cidx % 3u);
    const uint32_t cell = cell_at(d, st.base[0] + a3, st.base[1] + b3i, st.base[2] + c3);
    if (cell == NONE) continue;
    atomicAdd(&accr(d, NATRAY_A_SRC, cell), to_fp(qv * st.w[0][a3] * st.w[1][b3i] * st.w[2][c3], NATRAY_FP_Q));
  }
}

// bullet_apply (bullet.wgsl:379-386): the bullets' pushes onto the water's faces.
template <class L>
__global__ __launch_bounds__(256) void k_bullet_apply(L l) {
  const View vw = view(l);
  const Dev& d = vw.d;
  if (d.st->dt_ns == 0u) return;
  const uint32_t cells = n_blocks(d) * 64u;
  for (uint32_t cell = vw.wg * blockDim.x + threadIdx.x; cell < cells; cell += vw.wgs * blockDim.x)
#pragma unroll
    for (uint32_t dim = 0; dim < 3; ++dim) {
      const int32_t du = accr(d, NATRAY_A_BU + dim, cell);
      if (du != 0) cfr(d, NATRAY_F_U + dim, cell) += from_fp(du, NATRAY_FP_M);
    }
}

// ---- the air boundary (air.wgsl:85-148, 271-277; fluid.wgsl:91-107 for the block maps) -----------
//
// The air's own substeps (HLL fluxes) are not part of this library; what the water substep does
// with the gas is: the gas of the last step's blocks carried into this step's (air_carry), cells
// that stop being water given the gas next to them or vapour (air_fill), and the gas's
// pressure handed to the solve for GAS and RUBBER cells.

constexpr float GAMMA = 1.4f, P_ATM = 101325.0f, P_VAPOUR = 2339.0f, RHO_ATM = 1.204f, RHO_VAPOUR = 0.0278f;

struct Gas {
  float rho, mx, my, mz, e;
};
__device__ __forceinline__ Gas gas_at_rest(float rho, float p) { return {rho, 0.0f, 0.0f, 0.0f, p / (GAMMA - 1.0f)}; }
__device__ __forceinline__ float gas_p(const Gas& g) {
  return (GAMMA - 1.0f) * (g.e - 0.5f * ((g.mx * g.mx + g.my * g.my) + g.mz * g.mz) / g.rho);
}
__device__ __forceinline__ Gas read_gas(const float* air, uint32_t capc, uint32_t cell) {
  return {air[cell], air[(size_t)capc + cell], air[2 * (size_t)capc + cell], air[3 * (size_t)capc + cell],
          air[4 * (size_t)capc + cell]};
}
__device__ __forceinline__ void write_gas(float* air, uint32_t capc, uint32_t cell, const Gas& g) {
  air[cell] = g.rho;
  air[(size_t)capc + cell] = g.mx;
  air[2 * (size_t)capc + cell] = g.my;
  air[3 * (size_t)capc + cell] = g.mz;
  air[4 * (size_t)capc + cell] = g.e;
}

// keep_blocks (air.wgsl:271-277): the blocks whose air was doing anything last step are
// claimed again.
template <class L>
__global__ __launch_bounds__(256) void k_keep_blocks(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u || !d.air_x) return;
  for (uint32_t i = v.wg * blockDim.x + threadIdx.x; i < d.st->n_blocks_prev; i += v.wgs * blockDim.x)
    if (d.block_active[i] != 0u) d.block_ptr[d.block_list_prev[i]] = PENDING;
}

// air_carry (air.wgsl:85-98): each cell of this step's blocks gets the gas it had last step, or
// the ambient air where its block was not allocated then.
template <class L>
__global__ __launch_bounds__(256) void k_air_carry(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u || !d.air_x) return;
  const uint32_t n_prev = d.st->n_blocks_prev;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u, [&](uint32_t b, uint32_t lane, int, int, int, const uint32_t*) {
    const uint32_t s = d.block_list[b], o = d.prev_ptr[s];
    Gas g = gas_at_rest(RHO_ATM, P_ATM);
    if (o < n_prev && d.block_list_prev[o] == s) g = read_gas(d.air_x, d.capc, o * 64u + lane);
    write_gas(d.air_tmp, d.capc, b * 64u + lane, g);
    if (lane == 0) d.block_active[b] = 0u;
  });
}

// air_fill (air.wgsl:100-148), after grid_prep: water and solid cells hold no gas; a cell
// without gas takes the mean of its neighbours' or, with none, vapour; a GAS cell's pressure is
// its gas's, a RUBBER cell's the lowest of its GAS neighbours' (gauge).
template <class L>
__global__ __launch_bounds__(256) void k_air_fill(L l) {
  const View v = view(l);
  const Dev& d = v.d;
  if (d.st->dt_ns == 0u || !d.air_x) return;
  const uint32_t capc = d.capc;
  for_blocks(d, v.wg * 4u, n_blocks(d), v.wgs * 4u, [&](uint32_t b, uint32_t lane, int lx, int ly, int lz, const uint32_t* lk) {
    const uint32_t cell = b * 64u + lane;
    const float ty = cfr(d, NATRAY_F_TYPE, cell);
    if (ty == WATER || ty == SOLID) {
      write_gas(d.air_x, capc, cell, Gas{0.0f, 0.0f, 0.0f, 0.0f, 0.0f});
      return;
    }
    uint32_t nbs[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int sg = (k % 2 == 1) ? 1 : -1;
      nbs[k] = nbr(lk, lx, ly, lz, k / 2 == 0 ? sg : 0, k / 2 == 1 ? sg : 0, k / 2 == 2 ? sg : 0);
    }
    Gas g = read_gas(d.air_tmp, capc, cell);
    if (g.rho <= 0.0f) {
      Gas sum{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
      float n = 0.0f;
      for (int k = 0; k < 6; ++k) {
        if (nbs[k] == NONE) continue;
        const Gas q = read_gas(d.air_tmp, capc, nbs[k]);
        if (q.rho > 0.0f) {
          sum = Gas{sum.rho + q.rho, sum.mx + q.mx, sum.my + q.my, sum.mz + q.mz, sum.e + q.e};
          n += 1.0f;
        }
      }
      g = n > 0.0f ? Gas{sum.rho / n, sum.mx / n, sum.my / n, sum.mz / n, sum.e / n} : gas_at_rest(RHO_VAPOUR, P_VAPOUR);
    }
    write_gas(d.air_x, capc, cell, g);
    if (ty == GAS) cfr(d, NATRAY_F_P, cell) = gas_p(g) - P_ATM;
    if (ty == RUBBER) {
      float p = 3.0e38f;
      for (int k = 0; k < 6; ++k) {
        if (nbs[k] == NONE || cfr(d, NATRAY_F_TYPE, nbs[k]) != GAS) continue;
        const Gas q = read_gas(d.air_tmp, capc, nbs[k]);
        if (q.rho > 0.0f) p = fmin_(p, gas_p(q));
      }
      cfr(d, NATRAY_F_P, cell) = (p < 3.0e38f ? p : gas_p(g)) - P_ATM;
    }
  });
}

// ---- keyframe pack of the water (keys.wgsl:9-37, pack.wgsl:33-39) ------------------------------

// pack_pos (keys.wgsl:21-24): 21 bits per axis over an 8 m box from (-4, -0.5, -4), rounded to
// the nearest step; x and the low 11 bits of y in the first word, the high 10 bits of y, z and
// a flag (bit 31) in the second.
__device__ __forceinline__ uint2 pack_pos(float px, float py, float pz, bool flag) {
  const float lo[3] = {-4.0f, -0.5f, -4.0f}, p[3] = {px, py, pz};
  uint32_t q[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float t = fmin_(fmax_((p[k] - lo[k]) / 8.0f, 0.0f), 1.0f);
    q[k] = (uint32_t)(t * 2097151.0f + 0.5f);
  }
  return make_uint2(q[0] | (q[1] << 21), (q[1] >> 11) | (q[2] << 10) | (flag ? 1u << 31 : 0u));
}

// pack_water (pack.wgsl:33-39): two words per particle, the flag for a live particle.
template <class L>
__global__ __launch_bounds__(256) void k_pack_water(L l, uint2* __restrict__ out) {
  const View v = view(l);
  const Dev& d = v.d;
  const uint32_t i = v.wg * blockDim.x + threadIdx.x;
  if (i >= d.N) return;
  const float4* p = reinterpret_cast<const float4*>(d.P) + (size_t)i * 5;
  const float4 pos = p[0], vel = p[1];
  out[i] = pack_pos(pos.x, pos.y, pos.z, vel.w > 0.5f);
}

// ---- the rest of the keyframe (pack.wgsl:41-124; layout keys.wgsl:1-54) ----------------------------
//
// After the 4 header words and the water's 2 words per particle: the membrane's vertices (2
// words each), one alive bit per triangle, then the spray: a live bit per slot, the live count
// before each word of bits, and 3 words per live slot (position, then the radius's bits) in slot
// order. The spray index (live bits and their exclusive prefix) is built first, into `index`:
// [0] slots, [1] live slots, [2..3] unused, [4..4+cw) bits, [4+cw..4+2cw) prefix, where cw is the
// bit words of the spray capacity.

__host__ __device__ __forceinline__ uint32_t bit_words(uint32_t n) { return (n + 31u) / 32u; }

// pack_mem (pack.wgsl:41-46)
__global__ __launch_bounds__(256) void k_pack_mem(const float4* __restrict__ x, uint32_t n, uint2* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = x[i];
  out[i] = pack_pos(p.x, p.y, p.z, false);
}

// pack_alive (pack.wgsl:48-58)
__global__ __launch_bounds__(256) void k_pack_alive(const uint32_t* __restrict__ alive, uint32_t n_tris,
                                                    uint32_t* __restrict__ out) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w * 32u >= n_tris) return;
  uint32_t bits = 0u;
  for (uint32_t k = 0; k < 32u; ++k) {
    const uint32_t t = w * 32u + k;
    if (t < n_tris && alive[t] != 0u) bits |= 1u << k;
  }
  out[w] = bits;
}

// index_bits (pack.wgsl:60-74): the live bits of the first n slots.
__global__ __launch_bounds__(256) void k_spray_index_bits(const float4* __restrict__ spray, uint32_t n,
                                                          uint32_t* __restrict__ index) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w == 0u) index[0] = n;
  if (w >= bit_words(n)) return;
  uint32_t bits = 0u;
  for (uint32_t k = 0; k < 32u; ++k) {
    const uint32_t i = w * 32u + k;
    if (i < n && spray[(size_t)i * 3].w > 0.0f) bits |= 1u << k;
  }
  index[4u + w] = bits;
}

// index_scan (pack.wgsl:76-101): the exclusive prefix of the live counts over the bit words, in
// one workgroup (integers: any order gives the same values), and the live total in index[1].
__global__ __launch_bounds__(256) void k_spray_index_scan(uint32_t n, uint32_t cw, uint32_t* __restrict__ index) {
  __shared__ uint32_t part[256];
  const uint32_t li = threadIdx.x, nw = bit_words(n);
  const uint32_t chunk = (nw + 255u) / 256u;
  const uint32_t lo = min(li * chunk, nw), hi = min(lo + chunk, nw);
  uint32_t sum = 0u;
  for (uint32_t w = lo; w < hi; ++w) sum += __popc(index[4u + w]);
  part[li] = sum;
  __syncthreads();
  for (uint32_t off = 1u; off < 256u; off <<= 1) {
    const uint32_t v = li >= off ? part[li] + part[li - off] : part[li];
    __syncthreads();
    part[li] = v;
    __syncthreads();
  }
  uint32_t run = part[li] - sum;
  for (uint32_t w = lo; w < hi; ++w) {
    index[4u + cw + w] = run;
    run += __popc(index[4u + w]);
  }
  if (li == 255u) index[1] = part[255];
}

// pack_spray (pack.wgsl:103-124): the bits and prefix words, each slot's weight and time made
// into spray_w, and the live slots' words.
__global__ __launch_bounds__(256) void k_pack_spray(const float4* __restrict__ spray, uint32_t n, uint32_t cw,
                                                    const uint32_t* __restrict__ index, uint32_t* __restrict__ key,
                                                    uint32_t spray_word, float2* __restrict__ spray_w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t nw = bit_words(n), w = i / 32u;
  const uint32_t bits = index[4u + w], before = index[4u + cw + w];
  if (i % 32u == 0u) {
    key[spray_word + w] = bits;
    key[spray_word + nw + w] = before;
  }
  const float4 pos = spray[(size_t)i * 3], vel = spray[(size_t)i * 3 + 1], air = spray[(size_t)i * 3 + 2];
  if (spray_w) spray_w[i] = make_float2(vel.w, air.w);
  const uint32_t bit = 1u << (i % 32u);
  if ((bits & bit) == 0u) return;
  const uint32_t base = spray_word + 2u * nw + 3u * (before + (uint32_t)__popc(bits & (bit - 1u)));
  const uint2 q = pack_pos(pos.x, pos.y, pos.z, false);
  key[base] = q.x;
  key[base + 1u] = q.y;
  key[base + 2u] = __float_as_uint(pos.w);
}

// The 4 header words (keys.wgsl:1-8): spray slots and live slots from the spray index, or zeros
// without spray.
__global__ __launch_bounds__(64) void k_key_header(const uint32_t* __restrict__ index, uint32_t* __restrict__ key) {
  const uint32_t l = threadIdx.x;
  if (l < 4u) key[l] = (index && l < 2u) ? index[l] : 0u;
}

// ---- the spray index in a fixed device array (natray_sim_pack_keyframe_static) --------------------
//
// For callers that cannot allocate, the keyframe's spray index lives in this array, sized for
// the reference's spray capacity (NATRAY_SPRAY_CAP slots: 4 + 2 x 37,500 words, 300 KB). Each
// GPU holds its own copy until the process ends. Host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (kf_scratch below); the array's name in
// host code is not a device address.
static __device__ __align__(256) uint32_t scratch[NATRAY_KF_SCRATCH_WORDS];

int kf_scratch(int device, uint32_t** out) {
  static std::mutex lock;
  static uint32_t* addr[NATRAY_MAX_DEVICES] = {};
  if (device < 0 || device >= (int)NATRAY_MAX_DEVICES) return (int)hipErrorInvalidDevice;
  std::lock_guard<std::mutex> g(lock);
  if (!addr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return (int)e;
    if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return (int)e;
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) {
      const hipError_t e2 = hipSetDevice(cur);
      if (e == hipSuccess) e = e2;
    }
    if (e != hipSuccess) return (int)e;
    if (!p) return (int)hipErrorInvalidSymbol;
    addr[device] = static_cast<uint32_t*>(p);
  }
  *out = addr[device];
  return 0;
}

// ---- water diagnostics (debug_stats' water line, src/sim.rs:691-716) ----------------------------
//
// natray_sim_water_stats: the alive particles' count, box, mean position and velocity, kinetic
// energy per unit mass, fastest particle, and the grid cells that hold at least 4 of them now.
// Each workgroup sums its particles (a fixed grid-stride share) in double, in index order per
// thread, then over its threads as a fixed tree; the host folds the workgroups in order, so
// the sums do not depend on timing. The cells are counted in an open-addressing hash table of
// cell keys (integer counts, exact in any order).

constexpr unsigned long long CELL_EMPTY = ~0ull;

// The key of the cell holding (x, y, z): floor((p - origin) / h) per axis as the reference
// computes it in f32, each coordinate offset by 2^20 into 21 bits (coordinates beyond +-2^20
// cells are clamped).
__device__ __forceinline__ unsigned long long water_cell_key(float x, float y, float z, float ox, float oy, float oz,
                                                             float h) {
  const float c[3] = {floorf((x - ox) / h), floorf((y - oy) / h), floorf((z - oz) / h)};
  unsigned long long k = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = fminf(fmaxf(c[a], -1048576.0f), 1048575.0f);
    k = (k << 21) | (unsigned long long)((int)f + 1048576);
  }
  return k;
}

__global__ __launch_bounds__(256) void k_water_stats(const float4* P, uint32_t n, float ox, float oy, float oz, float h,
                                                     unsigned long long* keys, uint32_t* counts, uint32_t mask,
                                                     WaterPart* out) {
  double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  float vmax = 0.0f;
  uint32_t fast = NONE, alive = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    const float4 pos = P[(size_t)i * 5u], vel = P[(size_t)i * 5u + 1u];
    if (!(vel.w > 0.5f)) continue;
    alive += 1u;
    lo[0] = fminf(lo[0], pos.x);
    lo[1] = fminf(lo[1], pos.y);
    lo[2] = fminf(lo[2], pos.z);
    hi[0] = fmaxf(hi[0], pos.x);
    hi[1] = fmaxf(hi[1], pos.y);
    hi[2] = fmaxf(hi[2], pos.z);
    s[0] += (double)pos.x;
    s[1] += (double)pos.y;
    s[2] += (double)pos.z;
    s[3] += (double)vel.x;
    s[4] += (double)vel.y;
    s[5] += (double)vel.z;
    const float v2 = vel.x * vel.x + vel.y * vel.y + vel.z * vel.z;
    s[6] += 0.5 * (double)v2;
    const float sp = __builtin_sqrtf(v2);
    if (sp > vmax) {
      vmax = sp;
      fast = i;
    }
    // The table has at least twice as many slots as particles, so a free slot is always found.
    const unsigned long long key = water_cell_key(pos.x, pos.y, pos.z, ox, oy, oz, h);
    uint32_t slot = (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
      const unsigned long long prev = atomicCAS(&keys[slot], CELL_EMPTY, key);
      if (prev == CELL_EMPTY || prev == key) {
        atomicAdd(&counts[slot], 1u);
        break;
      }
      slot = (slot + 1u) & mask;
    }
  }
  // The workgroup's 256 threads as a fixed tree: lanes within each wave, then the 4 waves.
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 7; ++k) s[k] += __shfl_xor(s[k], off, 64);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
    }
    const float v2 = __shfl_xor(vmax, off, 64);
    const uint32_t f2 = (uint32_t)__shfl_xor((int)fast, off, 64);
    if (v2 > vmax || (v2 == vmax && f2 < fast)) {
      vmax = v2;
      fast = f2;
    }
    alive += (uint32_t)__shfl_xor((int)alive, off, 64);
  }
  __shared__ WaterPart w[4];
  const uint32_t wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63u) == 0u) {
    WaterPart& o = w[wv];
    for (int k = 0; k < 7; ++k) o.sum[k] = s[k];
    for (int a = 0; a < 3; ++a) {
      o.lo[a] = lo[a];
      o.hi[a] = hi[a];
    }
    o.vmax = vmax;
    o.fastest = fast;
    o.alive = alive;
    o.pad = 0u;
  }
  __syncthreads();
  if (threadIdx.x == 0u) {
    WaterPart o = w[0];
    for (int q = 1; q < 4; ++q) {
      const WaterPart& b = w[q];
      for (int k = 0; k < 7; ++k) o.sum[k] += b.sum[k];
      for (int a = 0; a < 3; ++a) {
        o.lo[a] = fminf(o.lo[a], b.lo[a]);
        o.hi[a] = fmaxf(o.hi[a], b.hi[a]);
      }
      if (b.vmax > o.vmax || (b.vmax == o.vmax && b.fastest < o.fastest)) {
        o.vmax = b.vmax;
        o.fastest = b.fastest;
      }
      o.alive += b.alive;
    }
    out[blockIdx.x] = o;
  }
}

// Table slots whose cell holds at least 4 particles, added into *full.
__global__ __launch_bounds__(256) void k_water_cells(const uint32_t* counts, uint32_t slots, uint32_t* full) {
  uint32_t c = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < slots; i += gridDim.x * 256u) c += counts[i] >= 4u ? 1u : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += (uint32_t)__shfl_xor((int)c, off, 64);
  if ((threadIdx.x & 63u) == 0u && c) atomicAdd(full, c);
}

// ---- launch wrappers -------------------------------------------------------------------------------

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Launches kernel template K on target t (one scene or a batch) with the workgroups of `kind`.
#define NATRAY_LAUNCH(K, kind, threads, t, stream, ...)                                                      \
  do {                                                                                                       \
    if ((t).d)                                                                                               \
      hipLaunchKernelGGL(K<One>, dim3((t).grid[kind]), dim3(threads), 0, stream, One{*(t).d}, ##__VA_ARGS__); \
    else                                                                                                     \
      hipLaunchKernelGGL(K<Many>, dim3((t).grid[kind]), dim3(threads), 0, stream, Many{(t).scenes, (t).map[kind]}, \
                         ##__VA_ARGS__);                                                                     \
  } while (0)

int launch_keys(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_keys, LK_PARTS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_begin(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_begin, LK_ONE, 64, t, S(stream));
  return (int)hipGetLastError();
}
int launch_bin(const Target& t, void* stream, bool rubber, bool air) {
  NATRAY_LAUNCH(k_clear, LK_CELLS, 256, t, S(stream));
  if (air) NATRAY_LAUNCH(k_keep_blocks, LK_ONE, 256, t, S(stream));
  // A single scene's marks take four particles per thread; a batch's scenes share the particle
  // workgroup map, one particle per thread.
  if (t.d)
    hipLaunchKernelGGL((k_mark<One, 4>), dim3((t.grid[LK_PARTS] + 3u) / 4u), dim3(256), 0, S(stream), One{*t.d});
  else
    hipLaunchKernelGGL((k_mark<Many, 1>), dim3(t.grid[LK_PARTS]), dim3(256), 0, S(stream),
                       Many{t.scenes, t.map[LK_PARTS]});
  if (rubber) NATRAY_LAUNCH(k_mark_rubber, LK_VERTS, 256, t, S(stream));
  NATRAY_LAUNCH(k_scan_reduce, LK_TILES, 256, t, S(stream));
  NATRAY_LAUNCH(k_scan_tiles, LK_ONE, 1024, t, S(stream));
  NATRAY_LAUNCH(k_scan_scatter, LK_TILES, 256, t, S(stream));
  if (t.d)
    hipLaunchKernelGGL((k_sort<One, 4>), dim3((t.grid[LK_PARTS] + 3u) / 4u), dim3(256), 0, S(stream), One{*t.d});
  else
    hipLaunchKernelGGL((k_sort<Many, 1>), dim3(t.grid[LK_PARTS]), dim3(256), 0, S(stream),
                       Many{t.scenes, t.map[LK_PARTS]});
  return (int)hipGetLastError();
}
int launch_p2g(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_p2g, LK_BLOCKS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_grid_prep(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_grid_prep, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int solve_max_grid(int kind, int* per_cu) {
  int occ = 0, occ_many = 0;
  hipError_t e;
  if (kind == (int)NATRAY_CG_PIPELINED) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_solve_pipe<One>, SOLVE_THREADS, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_many, k_solve_pipe<Many>, SOLVE_THREADS, 0);
  } else if (kind == (int)NATRAY_CG_MIXED) {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_solve<true, One>, SOLVE_THREADS, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_many, k_solve<true, Many>, SOLVE_THREADS, 0);
  } else {
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_solve<false, One>, SOLVE_THREADS, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_many, k_solve<false, Many>, SOLVE_THREADS, 0);
  }
  *per_cu = occ < occ_many ? occ : occ_many;
  return (int)e;
}
int launch_solve(const Target& t, void* stream, int kind) {
  const dim3 grid(t.grid[LK_SOLVE]), block(SOLVE_THREADS);
  One one{t.d ? *t.d : Dev{}};
  Many many{t.scenes, t.map[LK_SOLVE]};
  void* args_one[] = {&one};
  void* args_many[] = {&many};
  const void* fn;
  if (kind == (int)NATRAY_CG_PIPELINED)
    fn = t.d ? (const void*)k_solve_pipe<One> : (const void*)k_solve_pipe<Many>;
  else if (kind == (int)NATRAY_CG_MIXED)
    fn = t.d ? (const void*)k_solve<true, One> : (const void*)k_solve<true, Many>;
  else
    fn = t.d ? (const void*)k_solve<false, One> : (const void*)k_solve<false, Many>;
  // The solve is a plain launch unless its Dev (a single scene) or the Target (a batch) says
  // otherwise. Nothing promises
  // that a plain launch's workgroups run at the same time, but every wait of the solve is
  // bounded: were one not resident, the others would abort and the step would fail with
  // NATRAY_E_SOLVER rather than hang. A cooperative launch (every workgroup resident, or an
  // error) goes through the launch attribute so that stream capture can record it as a graph
  // node.
  if (t.d ? t.d->solve_plain != 0u : t.plain_solve)
    return (int)hipLaunchKernel(fn, grid, block, t.d ? args_one : args_many, 0, S(stream));
  hipLaunchAttribute attr{};
  attr.id = hipLaunchAttributeCooperative;
  attr.val.cooperative = 1;
  hipLaunchConfig_t cfg{};
  cfg.gridDim = grid;
  cfg.blockDim = block;
  cfg.dynamicSmemBytes = 0;
  cfg.stream = S(stream);
  cfg.attrs = &attr;
  cfg.numAttrs = 1;
  return (int)hipLaunchKernelExC(&cfg, fn, t.d ? args_one : args_many);
}
int launch_project(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_project, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_extend(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_extend1, LK_WAVES, 256, t, S(stream));
  NATRAY_LAUNCH(k_extend2, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_ws_inputs(const Target& t, void* stream, uint32_t cap, void* net_in) {
  NATRAY_LAUNCH(k_ws_prepare, LK_WAVES, 256, t, S(stream));
  NATRAY_LAUNCH(k_ws_tiles, LK_ONE, 1024, t, S(stream), cap);
  NATRAY_LAUNCH(k_ws_gather, LK_WAVES, 256, t, S(stream), static_cast<bf16x8*>(net_in));
  return (int)hipGetLastError();
}
int launch_ws_guess(const Target& t, void* stream, const float* pred) {
  NATRAY_LAUNCH(k_ws_scatter, LK_WAVES, 256, t, S(stream), pred);
  NATRAY_LAUNCH(k_ws_check, LK_WAVES, 256, t, S(stream));
  NATRAY_LAUNCH(k_ws_decide, LK_ONE, 1024, t, S(stream));
  return (int)hipGetLastError();
}
int launch_ws_zero_start(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_ws_zero_start, LK_ONE, 64, t, S(stream));
  return (int)hipGetLastError();
}
int launch_ws_samples(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_ws_samples, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_g2p(const Target& t, void* stream) {
  if (t.d && t.d->n_bullets)
    hipLaunchKernelGGL((k_g2p<One, true>), dim3(t.grid[LK_BLOCKS]), dim3(256), 0, S(stream), One{*t.d});
  else
    NATRAY_LAUNCH(k_g2p, LK_BLOCKS, 256, t, S(stream));
  NATRAY_LAUNCH(k_vmax, LK_ONE, 256, t, S(stream));
  return (int)hipGetLastError();
}

int launch_mem_splat(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_mem_splat, LK_VERTS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_mem_sample(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_mem_sample, LK_VERTS, 256, t, S(stream));
  return (int)hipGetLastError();
}

int launch_air_carry(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_air_carry, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_air_fill(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_air_fill, LK_WAVES, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_spray_update(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_spray_update, LK_CELLS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_break_up(const Target& t, void* stream) {
  NATRAY_LAUNCH(k_break_up, LK_CELLS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_bullets(const Target& t, void* stream) {
  if (!t.d || !t.d->n_bullets) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_bullet_patches<One>, dim3(t.d->n_bullets * (NATRAY_BULLET_PATCHES / 4u)), dim3(256), 0, S(stream),
                     One{*t.d});
  hipLaunchKernelGGL(k_bullet_body<One>, dim3(t.d->n_bullets), dim3(256), 0, S(stream), One{*t.d});
  NATRAY_LAUNCH(k_bullet_apply, LK_CELLS, 256, t, S(stream));
  return (int)hipGetLastError();
}
int launch_pack_keyframe(const Target& t, void* stream, const KeyframeArgs& a) {
  if (!t.d) return (int)hipErrorInvalidValue;
  const Dev& d = *t.d;
  uint32_t* key = static_cast<uint32_t*>(a.key);
  if (a.write_header) hipLaunchKernelGGL(k_key_header, dim3(1), dim3(64), 0, S(stream), a.index, key);
  NATRAY_LAUNCH(k_pack_water, LK_PARTS, 256, t, S(stream), reinterpret_cast<uint2*>(key + 4));
  const uint32_t mem_word = 4u + 2u * d.N, alive_word = mem_word + 2u * a.n_verts;
  if (a.n_verts)
    hipLaunchKernelGGL(k_pack_mem, dim3((a.n_verts + 255u) / 256u), dim3(256), 0, S(stream), d.mem_x, a.n_verts,
                       reinterpret_cast<uint2*>(key + mem_word));
  if (a.n_tris)
    hipLaunchKernelGGL(k_pack_alive, dim3((bit_words(a.n_tris) + 255u) / 256u), dim3(256), 0, S(stream), a.alive,
                       a.n_tris, key + alive_word);
  if (a.n_spray)
    hipLaunchKernelGGL(k_pack_spray, dim3((a.n_spray + 255u) / 256u), dim3(256), 0, S(stream),
                       reinterpret_cast<const float4*>(d.spray), a.n_spray, a.cw, a.index, key,
                       alive_word + bit_words(a.n_tris), reinterpret_cast<float2*>(a.spray_w));
  return (int)hipGetLastError();
}
int launch_spray_index(const Target& t, void* stream, uint32_t n, uint32_t cw, uint32_t* index) {
  if (!t.d) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_spray_index_bits, dim3((bit_words(n) + 255u) / 256u + 1u), dim3(256), 0, S(stream),
                     reinterpret_cast<const float4*>(t.d->spray), n, index);
  hipLaunchKernelGGL(k_spray_index_scan, dim3(1), dim3(256), 0, S(stream), n, cw, index);
  return (int)hipGetLastError();
}
int launch_pack_water(const Target& t, void* stream, void* out) {
  NATRAY_LAUNCH(k_pack_water, LK_PARTS, 256, t, S(stream), static_cast<uint2*>(out));
  return (int)hipGetLastError();
}

int launch_water_stats(const float* particles, uint32_t n, const float origin[3], float h, unsigned long long* keys,
                       uint32_t* counts, uint32_t slots, WaterPart* parts, uint32_t groups, uint32_t* full, void* stream) {
  if (!n || !groups || slots < 2u * n || (slots & (slots - 1u))) return (int)hipErrorInvalidValue;
  hipLaunchKernelGGL(k_water_stats, dim3(groups), dim3(256), 0, S(stream), reinterpret_cast<const float4*>(particles), n,
                     origin[0], origin[1], origin[2], h, keys, counts, slots - 1u, parts);
  hipLaunchKernelGGL(k_water_cells, dim3(std::min<uint32_t>(1024u, (slots + 255u) / 256u)), dim3(256), 0, S(stream), counts,
                     slots, full);
  return (int)hipGetLastError();
}

#undef NATRAY_LAUNCH

}  // namespace natray
