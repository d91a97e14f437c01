/* natray.h — C ABI of the MI355X (gfx950) water substep library, libnatray.so.
 *
 * The library replaces the GPU half of natray's `Sim` (src/sim.rs): the buffers that
 * `Sim::new` allocates (src/sim.rs:237-315), the kernel plan (src/sim.rs:390-426) and the
 * encode/submit/readback of `Sim::run` (src/sim.rs:585-648). The CPU half of `Sim::new`
 * (scene build, domain layout, Params packing, src/sim.rs:208-236,428-441) stays with the
 * caller, which hands the packed `natray_params` and the packed particle records over here.
 *
 * Layouts are byte-identical to the reference:
 *   natray_params   == Params    (src/sim.rs:49-64,   shaders/sim_defs.wgsl:8-21)   192 B
 *   natray_state    == State     (src/sim.rs:66-94,   shaders/sim_defs.wgsl:24-58)   96 B
 *   natray_particle == Particle  (shaders/sim_defs.wgsl:61, packed at src/sim.rs:227-237) 80 B
 *   acc / cf        == field-major grid sums and solver fields (shaders/sim_defs.wgsl:105-148)
 *
 * Conventions: every entry point returns 0 on success or a negative NATRAY_E* code and never
 * unwinds; natray_last_error() gives a thread-local message for the last failure. A handle is
 * not thread-safe; each handle owns its own HIP stream. All device memory is allocated in
 * natray_sim_create and freed in natray_sim_destroy; stepping allocates nothing.
 */
#ifndef NATRAY_H
#define NATRAY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- constants shared with the kernels (shaders/sim_defs.wgsl:98-157) ---------------- */

#define NATRAY_NONE 0xffffffffu       /* sim_defs.wgsl:98  */
#define NATRAY_PENDING 0xfffffffeu    /* sim_defs.wgsl:100 */
#define NATRAY_FLT_MAX_BITS 0x7f7fffffu

/* Transfer sums, ACC_FIELDS per cell, field-major (sim_defs.wgsl:110-125). */
#define NATRAY_A_MOM(d) (2u * (d))
#define NATRAY_A_WT(d) (2u * (d) + 1u)
#define NATRAY_A_WC 6u
#define NATRAY_A_PC 7u
#define NATRAY_A_N 8u
#define NATRAY_A_MF 9u
#define NATRAY_A_MA 10u
#define NATRAY_A_BU 11u
#define NATRAY_A_SRC 14u
#define NATRAY_A_MW 15u
#define NATRAY_A_MV 16u
#define NATRAY_ACC_FIELDS 19u

/* Fixed-point scales (sim_defs.wgsl:129-136). */
#define NATRAY_FP_W 16777216.0f
#define NATRAY_FP_M 65536.0f
#define NATRAY_FP_P 65536.0f
#define NATRAY_FP_F 4194304.0f
#define NATRAY_FP_A 65536.0f
#define NATRAY_FP_Q 16777216.0f

/* Solver fields, CELL_FIELDS per cell, field-major (sim_defs.wgsl:140-148; the velocity
 * extension reuses 4..7, fluid.wgsl:604-605). */
#define NATRAY_F_U 0u
#define NATRAY_F_P 3u
#define NATRAY_F_R 4u
#define NATRAY_F_Z 5u
#define NATRAY_F_D 6u
#define NATRAY_F_Q 8u
#define NATRAY_F_DIAG 9u
#define NATRAY_F_TYPE 10u
#define NATRAY_F_EXT 4u
#define NATRAY_F_EXT_SET 7u
#define NATRAY_CELL_FIELDS 11u

/* Cell types stored as f32 in F_TYPE (sim_defs.wgsl:153-157). */
#define NATRAY_GAS 0.0f
#define NATRAY_WATER 1.0f
#define NATRAY_SOLID 2.0f
#define NATRAY_COVERED 3.0f
#define NATRAY_RUBBER 4.0f

/* Physical constants of the reference (src/sim.rs:17-20, src/balloon.rs:10, air.wgsl:24-31). */
#define NATRAY_SOUND 1480.0f
#define NATRAY_CFL 0.8f
#define NATRAY_DT_MAX 1e-3f
#define NATRAY_SOLVE_TOL 1e-5f
#define NATRAY_MAX_ITERS 4096u
#define NATRAY_MAX_BATCH 256u
#define NATRAY_MAX_BLOCKS 65535u
#define NATRAY_RHO_WATER 1000.0f
#define NATRAY_SIGMA_W 0.072f
#define NATRAY_P_FLOOR (2339.0f - 101325.0f)

/* ---- layouts ----------------------------------------------------------------------------- */

/* == Params (src/sim.rs:49-64, sim_defs.wgsl:8-21). */
typedef struct natray_params {
  float origin[4];    /* domain corner, w = cell size h */
  uint32_t dims[4];   /* domain size in cells, w = particle count */
  uint32_t bdims[4];  /* domain size in 4^3 blocks, w = block capacity */
  float gravity[4];   /* xyz, w = sound speed in water */
  float fluid[4];     /* x density, y CFL, z longest step (s), w particle spacing */
  float wet[4];       /* x wet map half size (renderer side, unused here) */
  uint32_t limits[4]; /* x wet res, y spray cap, z CG iteration budget, w stop time (ns) */
  float solve[4];     /* x relative tolerance */
  uint32_t mem[4];    /* x rubber verts, y rubber substeps, z air substeps, w bullets */
  float rubber[4];
  float rubber2[4];
  float knot[4];      /* tie point xyz, w = neck radius */
} natray_params;

/* == State (src/sim.rs:66-94, sim_defs.wgsl:24-58). f32 members are stored as f32. */
typedef struct natray_state {
  uint32_t t_ns;
  uint32_t dt_ns;
  uint32_t step;
  uint32_t n_blocks;
  float max_speed;
  uint32_t n_blocks_prev;
  uint32_t overflow;
  uint32_t solve_iters;
  float rz;
  float alpha;
  float b_norm;
  float r_norm;
  uint32_t free_dt_ns;
  uint32_t iters_max;
  uint32_t unconverged;
  float beta;
  uint32_t spray_count;
  float mem_dt;
  float cap_dt;
  float mem_dt_seen;
  float dt_flow;
  float air_dt;
  float air_dt_seen;
  uint32_t mem_ticket;
} natray_state;

/* == Particle (sim_defs.wgsl:60-61). pos.w carried; vel.w = alive (1/0); c0/c1/c2.xyz = APIC
 * C rows of u/v/w; c0.w = carried gauge pressure (Pa); c1.w = thin-sheet age (s); c2.w = 0
 * after a step (the reference keeps the in-block rank there during its sort; this library
 * keeps ranks in a side array). */
typedef struct natray_particle {
  float pos[4];
  float vel[4];
  float c0[4];
  float c1[4];
  float c2[4];
} natray_particle;

/* Scene semantics beyond Params. The reference's only solids are the ground (y < 0) and
 * the knot (fluid.wgsl:305-314); the synthetic benchmark scenes add box walls, a moving
 * capsule solid, and a fixed spherical mask standing in for the membrane. */
#define NATRAY_SCENE_GROUND 1u   /* cells below y = 0 are SOLID (fluid.wgsl:305-307)     */
#define NATRAY_SCENE_KNOT 2u     /* the knot plug is SOLID (fluid.wgsl:309-314)          */
#define NATRAY_SCENE_WALLS 4u    /* the outermost cell layer of the domain is SOLID       */
#define NATRAY_SCENE_CAPSULE 8u  /* a capsule moving at constant velocity is SOLID        */
#define NATRAY_SCENE_MASK 16u    /* non-water cells inside a sphere are COVERED           */
#define NATRAY_SCENE_AIR 32u     /* the air boundary: gas carried between steps, cells that
                                    stop being water filled with gas or vapour, GAS and RUBBER
                                    pressures from it (air.wgsl:85-148,271-277); without it
                                    every GAS cell is at gauge 0 */
/* Spray and wet ground: a particle that leaves the resolved flow breaks up into drops
 * (break_up, fluid.wgsl:720-765) or, on the ground, wets it (ground_touch, sim_defs.wgsl:233-240),
 * and the drops fly until they land (spray_update, fluid.wgsl:686-711). Without it a leaving
 * particle only dies. */
#define NATRAY_SCENE_SPRAY 64u
#define NATRAY_SCENE_REFERENCE (NATRAY_SCENE_GROUND | NATRAY_SCENE_KNOT | NATRAY_SCENE_AIR | NATRAY_SCENE_SPRAY)

#define NATRAY_PRECOND_BLOCK_SGS 0u /* fluid.wgsl:384-410 */
#define NATRAY_PRECOND_JACOBI 1u

#define NATRAY_CG_FP32 0u  /* reference: f32 throughout */
#define NATRAY_CG_MIXED 1u /* f32 SpMV/axpy, f64 dots, norms, convergence and true-residual refresh */
/* Pipelined preconditioned CG (Ghysels and Vanroose, Parallel Computing 40, 2014) with the
 * precision of NATRAY_CG_MIXED: one reduction per iteration instead of two, and residual
 * replacement (Cools et al., SIAM J. Sci. Comput. 40, 2018) every refresh_every iterations.
 * Not in the reference; the same system and tolerance test as the other two. */
#define NATRAY_CG_PIPELINED 2u

typedef struct natray_scene {
  uint32_t flags;         /* NATRAY_SCENE_* */
  uint32_t precond;       /* NATRAY_PRECOND_* */
  uint32_t precision;     /* NATRAY_CG_* */
  uint32_t refresh_every; /* mixed and pipelined: recompute r = b - A p in f64 every this many iterations (0 = never) */
  float fixed_dt;         /* > 0: every step is this long (s), instead of the CFL choice */
  float mask_p;           /* gauge pressure (Pa) of COVERED mask cells */
  uint32_t solve_groups;  /* workgroups of the persistent pressure solve (0 = chosen from the block count) */
  uint32_t reserved0;
  float capsule_a[4];     /* capsule axis start at t = 0 (xyz), w = radius */
  float capsule_b[4];     /* capsule axis end at t = 0 (xyz) */
  float capsule_v[4];     /* capsule velocity (xyz) */
  float mask[4];          /* mask sphere centre (xyz), w = radius */
  uint32_t reserved1[4];
} natray_scene;

/* ---- return codes --------------------------------------------------------------------- */

#define NATRAY_OK 0
#define NATRAY_E_OVERFLOW (-1) /* more blocks than the capacity (src/sim.rs:635 asserts) */
#define NATRAY_E_NOMEM (-2)    /* allocation failed (src/gpu.rs:61-70 asserts) */
#define NATRAY_E_HIP (-3)      /* a HIP runtime error */
#define NATRAY_E_ARG (-4)      /* bad argument */
#define NATRAY_E_SOLVER (-5)   /* the persistent solver lost its grid barrier */

/* ---- phases of one step, in the reference's plan order (src/sim.rs:390-426) --------------- */

#define NATRAY_PHASE_BEGIN 0   /* begin_step                       fluid.wgsl:31-89           */
#define NATRAY_PHASE_BIN 1     /* clear/mark/scan/sort             fluid.wgsl:91-196          */
#define NATRAY_PHASE_P2G 2     /* p2g                              fluid.wgsl:254-303         */
#define NATRAY_PHASE_GRID 3    /* grid_prep (+ bullet_apply hook)  fluid.wgsl:319-356         */
#define NATRAY_PHASE_SOLVE 4   /* solve_init/start + iterations    fluid.wgsl:445-578         */
#define NATRAY_PHASE_PROJECT 5 /* project                          fluid.wgsl:580-598         */
#define NATRAY_PHASE_EXTEND 6  /* extend_1, extend_2               fluid.wgsl:600-674         */
#define NATRAY_PHASE_G2P 7     /* g2p                              fluid.wgsl:863-901         */
#define NATRAY_N_PHASES 8

/* Device buffers for zero-copy consumers (natray_sim_device_ptr). */
#define NATRAY_BUF_PARTICLES 0
#define NATRAY_BUF_STATE 1
#define NATRAY_BUF_ACC 2
#define NATRAY_BUF_CF 3
#define NATRAY_BUF_BLOCK_LIST 4
#define NATRAY_BUF_STREAM 5 /* the handle's hipStream_t */

typedef struct natray_sim natray_sim;

/* ---- entry points ----------------------------------------------------------------------- */

/* Sizes of the ABI structs, in the order params, state, particle, scene. Touches no GPU. */
void natray_abi_sizes(uint32_t out[4]);

/* Replaces the GPU half of Sim::new (src/sim.rs:237-315,428-441). `particles` holds n packed
 * 80-byte records (src/sim.rs:227-237); params->dims[3] must equal n and params->bdims[3] is
 * the block capacity. `scene` may be NULL (reference semantics: ground and knot solid,
 * block-SGS preconditioner, f32 CG). `init` may be NULL (the reference's first State,
 * src/sim.rs:240-241: mem_dt = cap_dt = FLT_MAX, air_dt = 0.9 h / (3 c_air)). */
int natray_sim_create(const natray_params* params, const float* particles, uint32_t n,
                      const natray_scene* scene, const natray_state* init, natray_sim** out);
void natray_sim_destroy(natray_sim* sim);

/* == Sim::run (src/sim.rs:585-648): n_steps steps with a CG budget of `iters` iterations per
 * step (0: the budget the last batches adapted), stopping at `stop_ns`; reads the State back
 * into `out` (may be NULL). Returns NATRAY_E_OVERFLOW where the reference asserts. With
 * bullets set, n_steps may not exceed NATRAY_BULLET_LOG: the pose log holds one entry per
 * step at step % NATRAY_BULLET_LOG and the reference reads it after every batch of at most
 * that many steps. */
int natray_sim_run(natray_sim* sim, uint32_t n_steps, uint32_t iters, uint32_t stop_ns, natray_state* out);

/* == Sim::advance (src/sim.rs:498-522), with the reference's batch sizing and CG budget
 * adaptation (src/sim.rs:636-645). Steps taken go to *taken. The bullet handoff time t0 of
 * the reference is 0 here (no exterior flight). */
int natray_sim_advance(natray_sim* sim, double t_target, uint32_t budget, uint32_t* taken);

/* The host rules of Sim::advance/run as pure functions (no GPU). natray_plan_batch: how many
 * steps the next batch encodes (src/sim.rs:513-515): as many as the last free step length
 * would take to reach stop_ns, at least 1, at most min(budget_left, 256, 4 ms / step_ms).
 * natray_next_iters: the CG budget after a batch (src/sim.rs:636-645): doubled (to at most
 * 4096) when a step ran out of iterations, else 1.25 x the most any step took + 2, within
 * [8, 4096], else unchanged. */
uint32_t natray_plan_batch(uint32_t stop_ns, uint32_t t_ns, uint32_t free_dt_ns, double step_ms, uint32_t budget_left);
uint32_t natray_next_iters(uint32_t iters, uint32_t iters_max, uint32_t unconverged);

/* One natray_sim_run call, as natray_sim_advance issues them (the batches NATRAY_PROF would
 * show, src/sim.rs:585-648): the steps asked for, the CG budget and stop time they ran with,
 * the steps taken, and the State's iters_max, unconverged and t_ns after them. */
typedef struct natray_run_record {
  uint32_t n_steps, iters, stop_ns, taken, iters_max, unconverged, t_ns, pad;
} natray_run_record;
#define NATRAY_RUN_LOG 256u
/* The last min(runs, max_records) records, oldest first; *count gets how many are kept (at
 * most NATRAY_RUN_LOG). out may be NULL when max_records is 0. A run that overflowed is
 * recorded too. */
int natray_sim_run_log(natray_sim* sim, uint32_t max_records, natray_run_record* out, uint32_t* count);

/* One step with the current budget and no stop time. */
int natray_sim_step(natray_sim* sim, natray_state* out);

/* One phase of a step on the handle's stream (for parity tests). */
int natray_sim_phase(natray_sim* sim, int phase);

/* Host <-> device copies (synchronous on the handle's stream). */
int natray_sim_read_state(natray_sim* sim, natray_state* out);
int natray_sim_write_state(natray_sim* sim, const natray_state* in);
int natray_sim_read_particles(natray_sim* sim, float* dst);
int natray_sim_write_particles(natray_sim* sim, const float* src);
int natray_sim_write_params(natray_sim* sim, const natray_params* params);
int natray_sim_read_params(natray_sim* sim, natray_params* out);

/* The grid of the blocks in use: *n_blocks, block_list[n], block_nb[27 n], block_start[n+1],
 * acc[19][64 n] and cf[11][64 n] (field-major, compacted to the blocks in use). Any output
 * pointer may be NULL. Call with everything NULL first to learn n_blocks. */
int natray_sim_read_grid(natray_sim* sim, uint32_t* n_blocks, uint32_t* block_list, uint32_t* block_nb,
                         uint32_t* block_start, int32_t* acc, float* cf);
/* Overwrites acc and cf of the blocks in use with the same compacted layout (parity tests
 * feed a phase from identical grid state). */
int natray_sim_write_grid(natray_sim* sim, const int32_t* acc, const float* cf);
/* block_ptr[S] (S = product of bdims) and sorted[N]. */
int natray_sim_read_bins(natray_sim* sim, uint32_t* block_ptr, uint32_t* sorted);

void* natray_sim_device_ptr(natray_sim* sim, int which);

/* Per-phase GPU timing with HIP events on the handle's stream (the analogue of NATRAY_PROF,
 * src/sim.rs:594-632). ms[NATRAY_N_PHASES] gets the summed milliseconds since profiling was
 * switched on, counts[] the launches. */
int natray_sim_set_profiling(natray_sim* sim, int on);
int natray_sim_phase_times(natray_sim* sim, double* ms, uint32_t* counts);

/* How natray_sim_run launches a step (no counterpart in the reference, whose step is one
 * encoded command buffer, src/sim.rs:602-627). On (the default): the step is one HIP graph,
 * captured once and replayed every step, whose solve is a cooperative kernel node; with
 * profiling or the warm start on, the kernels before the solve and those after it are two
 * graphs with the solve launched between them. Off: every kernel is launched on its own. All
 * give the same bits. With graphs on, profiling times only the solve phase. */
int natray_sim_set_graph(natray_sim* sim, int on);

/* How a single scene's grid-synchronised solve is launched (no counterpart in the reference).
 * NATRAY_SOLVE_PLAIN (the default): a plain launch of at most one workgroup per CU, a node of
 * the step graph like the other kernels. Nothing promises that its workgroups run at the same
 * time, but every wait of the solve is bounded: if another program holds part of the GPU and
 * one is not resident, the others abort after their spin limit (seconds) and the step fails
 * with NATRAY_E_SOLVER and a message naming this call; it never hangs.
 * NATRAY_SOLVE_COOPERATIVE: a cooperative launch (all resident, or the launch fails); it costs
 * about 27 us more per step on the 2M impact. NATRAY_SOLVE_PLAIN_TEST_CHECK: the plain launch
 * with its last workgroup staying out, so the solve aborts (for tests of that path). Both
 * launches give the same bits. A batch's solve: natray_batch_set_solve_launch. */
#define NATRAY_SOLVE_PLAIN 0
#define NATRAY_SOLVE_COOPERATIVE 1
#define NATRAY_SOLVE_PLAIN_TEST_CHECK 2
int natray_sim_set_solve_launch(natray_sim* sim, int mode);

/* Work counters since creation, summed over the steps that ran: [0] steps, [1] CG
 * iterations, [2] water-cell CG iterations, [3] blocks in use, [4] water cells, [5] particles
 * binned, [6] blocks cleared, [7] 0. They give the exact algorithmic bytes of a run. */
int natray_sim_counters(natray_sim* sim, uint64_t out[8]);

/* The water line of debug_stats (src/sim.rs:691-716) for the particles as they are now: over
 * the alive particles (record float 7 > 0.5), their count; their bounding box; mean position
 * and velocity; sum of 0.5 |v|^2 (per unit particle mass: times the particle mass
 * RHO_WATER spacing^3 for joules, src/sim.rs:853); the fastest one's speed and position (the
 * first in particle order among equal speeds, as the reference's strict comparison keeps; 0
 * with none); and the grid cells floor((x - origin) / dx) holding at least 4 of them, with
 * their volume in litres (cells x dx^3 x 1e3). The reference sums in f32 one particle after
 * another; here the sums are f64 in a fixed order that does not depend on timing. With no
 * alive particle, lo is FLT_MAX and hi -FLT_MAX, as the reference's fold leaves them.
 * Allocates a temporary cell table of 12 bytes x the next power of two >= 2 n (freed before
 * returning). */
typedef struct natray_water_stats {
  double mean[3];
  double vmean[3];
  double ke;
  double volume_l;
  float lo[3];
  float hi[3];
  float fastest[3];
  float vmax;
  uint32_t alive;
  uint32_t full_cells;
} natray_water_stats;
int natray_sim_water_stats(natray_sim* sim, natray_water_stats* out);

const char* natray_last_error(void);

/* ---- the CG warm-start network ------------------------------------------------------------
 * No counterpart in the reference: the learned initial guess of the pressure solve. A 4-layer
 * 3D convolution net on 16^3 tiles of cells (3^3 kernels; 2 -> 16 -> 16 -> 16 -> 1 channels,
 * ReLU between), BF16 weights and activations with FP32 accumulation on MFMA. Inputs per
 * voxel: the solve's initial residual scaled by its tile's largest magnitude, and the cell
 * kind (1 water, -1 solid, 0 otherwise); output: the pressure correction in the same scale.
 * Arrays passed to these functions are host memory, tile-major, voxel v = 256 x + 16 y + z. */
typedef struct natray_net natray_net;

/* Parameters, FP32, in order: W1[27][2][16], b1[16], W2[27][16][16], b2[16], W3[27][16][16],
 * b3[16], W4[27][16][1], b4[1] (W[tap][in][out], tap = 9 (dx + 1) + 3 (dy + 1) + dz + 1). */
#define NATRAY_NET_PARAMS 15169

/* A network with room for `max_tiles` tiles (<= 65536), He-uniform weights from `seed`. */
int natray_net_create(uint32_t max_tiles, uint64_t seed, natray_net** out);
void natray_net_destroy(natray_net* net);
int natray_net_params(natray_net* net, float* out);
int natray_net_set_params(natray_net* net, const float* in);
/* output[tiles][4096] = the network on input[tiles][4096][2]. */
int natray_net_forward(natray_net* net, const float* input, uint32_t tiles, float* output);
/* One Adam step (lr; lr = 0 only computes gradients) on the masked mean squared error of the
 * output against target[tiles][4096] over the voxels where mask[tiles][4096] is 1; *loss gets
 * the loss before the step. */
int natray_net_train_step(natray_net* net, const float* input, const float* target, const float* mask, uint32_t tiles,
                          float lr, float* loss);
/* One Adam step on the scaled residual the guess leaves: loss = mean over water voxels of
 * (x_v - (A o)_v)^2, x = input channel 0, o the output, (A o)_v = diag_v o_v - the sum of o
 * over v's water neighbours inside the tile. diag[tiles][4096] is the solve's diagonal (0 off
 * water), nbm[tiles][4096] its water-neighbour bits (bit 2 axis for the - side, 2 axis + 1 for
 * the + side; axes x, y, z), both as natray_sim_ws_samples gives them. */
int natray_net_train_step_residual(natray_net* net, const float* input, const float* diag, const uint32_t* nbm,
                                   uint32_t tiles, float lr, float* loss);
/* The parameter gradients of the last train step. */
int natray_net_grads(natray_net* net, float* out);
/* One convolution layer alone, for gradient checks: Y = act(conv(X, W) + b) and, from dY,
 * dX, dW and db, with X, W and the ReLU-masked dY rounded to BF16 as the network does.
 * X[tiles][4096][cin], W[27][cin][cout], dY and Y [tiles][4096][cout]; cin is 2 or 16, cout
 * 16 or 1, tiles <= 64. */
int natray_net_layer_check(uint32_t cin, uint32_t cout, int relu, const float* X, const float* W, const float* b,
                           const float* dY, uint32_t tiles, float* Y, float* dX, float* dW, float* db);
const char* natray_net_last_error(void);

/* ---- membrane coupling ----------------------------------------------------------------------
 * The rubber membrane's own step (Ogden forces, substeps: shaders/membrane.wgsl) is not part of
 * this library; what the water substep runs with its vertices is. n vertices as the membrane's
 * last substep left them: x[n][4] positions (m), v[n][4] velocities (m/s), mass[n] (0 where the
 * knot holds a vertex), push[n][2] its push on the water (N) and its area (m^2). Each step then
 * claims the blocks around rubber off the water (mark_rubber, membrane.wgsl:164-180), splats
 * the rubber's area, motion and push to the grid after P2G (mem_splat, membrane.wgsl:182-216;
 * grid_prep makes RUBBER and COVERED cells of them, fluid.wgsl:339-355) and samples the water at
 * every vertex after the extension (mem_sample, membrane.wgsl:239-254). Sets params.mem.x = n
 * (the reference's vertex count, src/sim.rs:428-441). Buffers grow on demand; the sampled water
 * of the previous step, which decides which vertices are on the water, is kept while the
 * vertex count does not grow. n = 0 leaves the membrane out. */
int natray_sim_set_membrane(natray_sim* sim, uint32_t n, const float* x, const float* v, const float* mass,
                            const float* push);
/* The water sampled at each vertex by the last step: water[n][4] = face velocity (m/s) and
 * fill (1 inside the water, 0 away from it), the reference's mem_water. */
int natray_sim_read_membrane(natray_sim* sim, float* water);

/* With NATRAY_SCENE_SPRAY: the spray records as break_up and spray_update left them, the
 * `spray` buffer of the reference (src/sim.rs:252, Spray at sim_defs.wgsl:62) and `landed`
 * (src/sim.rs:254). Copies the first min(count, cap, max_records) records: spray[k][12] =
 * position (m) and radius (m, 0 once gone), velocity (m/s) and weight (drops in the parcel),
 * 0 0 0 and the time it was made (s); landed[k][4] = where it landed (x, z), the radius of its
 * drops' volume and when (zero until it lands). *count = State.spray_count, which counts every
 * slot asked for, also past the cap (as the reference's). The slots follow the order in which
 * particles broke up, which is not fixed: compare records as a set. Either array may be null. */
int natray_sim_read_spray(natray_sim* sim, uint32_t max_records, float* spray, float* landed, uint32_t* count);
/* With NATRAY_SCENE_SPRAY: the wet map, `touch` of the reference (src/sim.rs:253, written by
 * ground_touch, sim_defs.wgsl:233-240), limits.x^2 words, row z, column x: the f32 bits of the
 * first time (s) water touched each ground cell, 0xffffffff where none has. */
int natray_sim_read_wet(natray_sim* sim, uint32_t* touch);

/* A whole keyframe of the live simulation (pack_water, pack_mem, pack_alive, index_bits,
 * index_scan and pack_spray, shaders/pack.wgsl:33-124, in the layout of shaders/keys.wgsl:1-54):
 * 4 header words (spray slots n = min(State.spray_count, cap), live spray slots, 0, 0); the
 * water's 2 words per particle; the membrane's 2 words per vertex (those natray_sim_set_membrane
 * gave); one alive bit per triangle (alive[n_tris], the membrane's own flags, may be null for
 * n_tris = 0); then the spray's live bits, the live count before each bit word, and 3 words per
 * live slot (packed position, radius bits) in slot order. *n_words = its size in words, which
 * Cache sizes as key_bytes (src/cache.rs:408-416, before rounding to 256 bytes). words = null
 * only asks the size; max_words smaller than it is NATRAY_E_ARG. spray_w[n][2] (may be null)
 * receives each slot's weight and the time it was made, the reference's spray_w. The spray's
 * slots follow the order in which particles broke up, which is not fixed. */
int natray_sim_pack_keyframe(natray_sim* sim, uint32_t n_tris, const uint32_t* alive, uint32_t* words,
                             uint32_t max_words, uint32_t* n_words, float* spray_w);

/* The reference's spray capacity SPRAY_CAP (src/sim.rs:13), and the words of the fixed spray
 * index natray_sim_pack_keyframe_static keeps on each GPU: 4 + 2 x ceil(SPRAY_CAP / 32). */
#define NATRAY_SPRAY_CAP 1200000u
#define NATRAY_KF_SCRATCH_WORDS (4u + 2u * ((NATRAY_SPRAY_CAP + 31u) / 32u))

/* No counterpart in the reference. natray_sim_pack_keyframe for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same keyframe, word for word, with no hipMalloc, hipFree or allocating
 * library call. Differences from natray_sim_pack_keyframe:
 * - alive (n_tris flags), words (max_words) and spray_w (n x 2 floats, may be null) are DEVICE
 *   memory the caller owns; the call returns once the keyframe is written;
 * - the spray index (live bits and their prefix) lives in a fixed device array of
 *   NATRAY_KF_SCRATCH_WORDS words (300 KB) that each GPU holds until the process ends. A scene
 *   whose spray capacity (params.limits.y) exceeds NATRAY_SPRAY_CAP slots is NATRAY_E_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (from several
 * host threads or handles at once) race, and must be serialised by the caller. */
int natray_sim_pack_keyframe_static(natray_sim* sim, uint32_t n_tris, const uint32_t* alive, uint32_t* words,
                                    uint32_t max_words, uint32_t* n_words, float* spray_w);

/* ---- bullets -------------------------------------------------------------------------------
 * The water's side of the reference's bullet model (shaders/bullet.wgsl): each step, after the
 * grid is prepared, every bullet in the grid takes the pressure of the water or gas at each
 * patch of its surface, the water takes the opposite push on its faces (bullet_step and
 * bullet_apply, bullet.wgsl:187-386), the cavity the nose opens feeds the pressure solve as a
 * source of volume while it is thinner than the grid (A_SRC, fluid.wgsl:481-482), and G2P moves
 * particles out of it (carve, bullet.wgsl:390-424, fluid.wgsl:818). The bullet's clock joins
 * begin_step's choice of dt (fluid.wgsl:36-50). Not part of this library: the flight through
 * the air before the bullet arrives (src/flight.rs), the rubber's contact with it
 * (rubber_contact, bullet.wgsl:519-611, part of the membrane's own substeps; its impulses and
 * held rubber are zero here). The renderer's pose log is kept (natray_sim_read_pose_log). */
#define NATRAY_MAX_BULLETS 4u       /* MAX_SHOTS (src/params.rs:49) */
#define NATRAY_BULLET_PATCHES 576u  /* MAX_PATCHES (src/bullet.rs:9) */
#define NATRAY_BULLET_SECTIONS 32u  /* N_SEC (src/bullet.rs:10) */
#define NATRAY_BULLET_LOG 256u      /* LOG_CAP (src/bullet.rs:11): pose log entries per bullet */

/* == Bullet (bullet.wgsl:16-34, GpuBullet src/bullet.rs:19-28), 112 bytes. */
typedef struct natray_bullet {
  float x[4];       /* centre of mass (m), w = mass (kg) */
  float v[4];       /* velocity (m/s), w = roll angle (rad) */
  float q[4];       /* body to world rotation (x, y, z, w); the body's +z runs base to nose */
  float l[4];       /* angular momentum about the centre of mass, w = the pace dt allows for (m/s) */
  float shape[4];   /* x axial moment of inertia, y transverse, z reach of the surface from the
                       centre of mass (m), w radius (m) */
  float reach[4];   /* centre and radius of a sphere around its thin cavity sections */
  uint32_t info[4]; /* x arrival (sim ns), y patch count, z 0 on its way / 1 in the grid / 2 gone,
                       w sections made so far */
} natray_bullet;

/* Sets n <= NATRAY_MAX_BULLETS bullets: bullets[n] as the flight hands them over, and
 * patches[n][NATRAY_BULLET_PATCHES][8] each bullet's surface in its body frame (position xyz
 * and area, outward normal xyz and 0; the first info.y are used), as src/bullet.rs:97-124
 * builds it. Clears the cavity sections. Sets params.mem.w = n (src/sim.rs:299-311). One scene
 * only (not in a batch). n = 0 removes them. */
int natray_sim_set_bullets(natray_sim* sim, uint32_t n, const natray_bullet* bullets, const float* patches);
/* The bullets as the last step left them, and each one's cavity sections,
 * sections[n][NATRAY_BULLET_SECTIONS][12] (the reference's Section, bullet.wgsl:42-50: a.xyz
 * and when made (s), b.xyz and d(R^2)/dt, R^2 at the start). Either array may be null. */
int natray_sim_read_bullets(natray_sim* sim, natray_bullet* bullets, float* sections);
/* The bullets' pose log (bullet.wgsl:52-57, 206-209, 355): log[n][NATRAY_BULLET_LOG][12], the
 * entry of step s at s % NATRAY_BULLET_LOG: position and the f32 bits of t_ns, rotation,
 * velocity and roll. A step a bullet is not in the grid zeroes its entry's first four words.
 * The renderer reads it as Sim::read_poses does (src/sim.rs:650-667). */
int natray_sim_read_pose_log(natray_sim* sim, float* log);

/* With NATRAY_SCENE_AIR: each cell's gas after the last phase that wrote it, air[5][64 n] for
 * the n blocks in use (density kg/m^3, momentum x, y, z, total energy per volume J/m^3; 0 in
 * water and solid cells). The air's own substeps (HLL, air.wgsl) are not part of this library:
 * the gas moves only by being carried between blocks and filled into cells that stop being
 * water. */
int natray_sim_read_air(natray_sim* sim, float* air);

/* The water's part of a keyframe (pack_water, shaders/pack.wgsl:33-39 with pack_pos,
 * shaders/keys.wgsl:21-24): words[2 N], two per particle in particle order, 21 bits per axis
 * over an 8 m box from (-4, -0.5, -4), bit 31 of the second word set for a live particle. The
 * keyframe places them after its 4 header words (keys.wgsl KEY_HEAD). */
int natray_sim_pack_water(natray_sim* sim, uint32_t* words);

/* ---- batches of independent scenes ----------------------------------------------------------
 * No counterpart in the reference, which steps one Sim per scene (src/sim.rs). A batch steps
 * many scenes, each a natray_sim of its own (particles, grid, clock, dt and CG of its own), with
 * one launch per kernel for all of them: a workgroup map gives every workgroup a scene and its
 * place among that scene's workgroups. Workgroups are shared among the scenes in proportion to
 * their particle counts; the cooperative solve gives every scene workgroups of its own, which
 * reduce among themselves. A scene gives the same bits in a batch as alone. The scenes must
 * share the CG precision, be on one device and not use the warm start; they stay owned by the
 * caller and must outlive the batch. A batch holds at most one scene per compute unit. */
typedef struct natray_batch natray_batch;
/* Launch shape of a batch: workgroups per compute unit shared among the scenes for P2G and G2P
 * (blocks), for the one-wave-per-block kernels (waves) and for clearing (cells). */
typedef struct {
  uint32_t blocks_per_cu, waves_per_cu, cells_per_cu, pad;
} natray_batch_shape;
int natray_batch_create(natray_sim* const* sims, uint32_t n, natray_batch** out);
void natray_batch_destroy(natray_batch* batch);
/* n_steps steps of every scene, with CG iteration cap `iters` (0: each scene's params.limits.z
 * as it stands); each scene's State is read back afterwards (natray_sim_read_state). Once every
 * scene has stepped, the solve's workgroups are shared by each scene's blocks times its last
 * CG iteration count rather than by particles. */
int natray_batch_run(natray_batch* batch, uint32_t n_steps, uint32_t iters);
/* Steps as replayed graphs (on, the default): one graph of every phase of a step, and one of 8
 * steps for each run of 8 in natray_batch_run, captured again when the solve's share of
 * workgroups changes; or kernel by kernel. */
int natray_batch_set_graph(natray_batch* batch, int on);
/* How a batch's solve is launched (no counterpart in the reference): NATRAY_SOLVE_PLAIN (the
 * default), a node of the step graphs whose every wait is bounded, as for a single scene; or
 * NATRAY_SOLVE_COOPERATIVE. Both give the same bits. */
int natray_batch_set_solve_launch(natray_batch* batch, int mode);
int natray_batch_set_shape(natray_batch* batch, const natray_batch_shape* shape);
/* Per-phase GPU time of the batch's kernel-by-kernel steps (graphs off), summed over the runs
 * since profiling was switched on, as natray_sim_phase_times. */
int natray_batch_set_profiling(natray_batch* batch, int on);
int natray_batch_phase_times(natray_batch* batch, double* ms, uint32_t* counts);
int natray_batch_get_shape(natray_batch* batch, natray_batch_shape* out);
/* The autotuner: times each candidate shape (blocks and waves per CU in {4, 8, 16}) on real
 * steps of the scenes, one step per candidate per round, the rounds visiting the candidates
 * forwards and backwards in turn so that the scenes' change over the steps weighs all alike
 * (`steps` rounds, rounded up to even; the scenes advance by rounds x candidates steps). Keeps
 * the fastest and returns it in *best, with the candidates' mean ms per step in ms[] (up to
 * max_candidates) and their number. */
int natray_batch_tune(natray_batch* batch, uint32_t steps, uint32_t iters, natray_batch_shape* best, double* ms,
                      uint32_t max_candidates, uint32_t* n_candidates);
/* The workgroup shares the batch uses: out[k] = scene k's part of `budget`, in proportion to
 * weights[k] (largest remainders), at least min_each and at most max_each[k] (no bound when
 * max_each is NULL). */
int natray_batch_plan(const uint64_t* weights, const uint32_t* max_each, uint32_t n, uint32_t budget, uint32_t min_each,
                      uint32_t* out);

/* ---- the warm start in the step ------------------------------------------------------------
 * No counterpart in the reference, whose solve always starts from the particle-carried pressure
 * (solve_init, shaders/fluid.wgsl:450-489). Each step, before the solve: r0 = b - A p_carried on
 * the water cells; the 16^3 tiles holding water are listed in tile order; each listed tile's
 * inputs are r0 / s (s = the tile's largest |r0|) and the cell kind.
 * mode 0: off (the default). mode 1: only the inputs, kept for natray_sim_ws_samples; the solve
 * starts as without the warm start. mode 2: `net` (up to its max_tiles tiles) maps the inputs to
 * o, guess = p_carried + s o on the tiles' water cells (p_carried elsewhere), and the solve
 * starts from the guess when |b - A guess|^2 < |r0|^2 (FP64 sums), else from p_carried. The
 * right-hand side b and the tolerance are the same either way. The network must stay alive
 * while the mode uses it. mode 3: every solve starts from p = 0 (the zero guess), b unchanged;
 * `net` is not used. Resets the statistics of natray_sim_ws_stats. */
int natray_sim_set_warm_start(natray_sim* sim, natray_net* net, int mode);
/* Training samples of the last step's listed tiles (mode 1 or 2), at most max_tiles of them:
 * inputs[tiles][4096][2] as the network got them, targets[tiles][4096] = (p_solved -
 * p_carried) / s and masks[tiles][4096] = 1 on water cells (0 in a tile whose s is 0), and the
 * solve's operator: diag[tiles][4096] (0 off those cells) and nbm[tiles][4096], the bits of the
 * water neighbours (bit 2 axis for the - side, 2 axis + 1 for the + side). *tiles gets the
 * count. Any of the arrays may be NULL. */
int natray_sim_ws_samples(natray_sim* sim, float* inputs, float* targets, float* masks, float* diag, uint32_t* nbm,
                          uint32_t max_tiles, uint32_t* tiles);
/* Since the mode was set: [0] solves offered a guess, [1] guesses taken, [2] |b - A guess|^2
 * and [3] |r0|^2 of the last step, [4] tiles with water in the last step (also those beyond
 * the network's capacity), [5..7] 0. */
int natray_sim_ws_stats(natray_sim* sim, double out[8]);
/* The last step's guess (bdims.w * 64 cells, the grid's cell order) and whether the solve
 * started from it. */
int natray_sim_ws_guess(natray_sim* sim, float* guess, uint32_t* accepted);

#ifdef __cplusplus
}
#endif

#endif /* NATRAY_H */
