// Device-side view of one simulation: pointers into the buffers natray_sim_create allocates.
// Shared by the kernels (natray_kernels.hip) and the host library (natray_sim.cpp); the struct
// is passed to every kernel by value.
#pragma once
#include <stdint.h>

#include "natray.h"

namespace natray {

// A particle's own cell packed as x | y << 10 | z << 20 (domain sides up to 1024 cells), or
// NATRAY_NONE for a dead particle or one outside the domain. The G2P kernel writes it after
// advecting, so binning reads 4 bytes per particle instead of the 80-byte record.
constexpr uint32_t KEY_BITS = 10;
constexpr uint32_t KEY_MASK = (1u << KEY_BITS) - 1u;
constexpr uint32_t MAX_DIM = 1u << KEY_BITS;

// Binning scans the domain slots in tiles of this many.
constexpr uint32_t SCAN_TILE = 1024;

// Solver reductions fold blocks in chunks of K = ceil(n / MAX_CHUNKS), so there are at most
// this many partial sums.
constexpr uint32_t MAX_CHUNKS = 256;
// The solve's tagged chunk sums: one 256-byte record per chunk (2 buffers x 8 slots of 8
// bytes, padded), so the 256 workgroups polling every chunk spread over many memory channels
// instead of queueing on a few lines.
constexpr uint32_t RED_RECORD = 32;  // words per chunk record
constexpr uint32_t RED_WORDS = RED_RECORD * MAX_CHUNKS;


struct Dev {
  const natray_params* U;  // device copy of Params (the reference's uniform)
  const natray_scene* sc;  // device copy of the scene
  natray_state* st;        // State
  float* P;                // particles, N x 20
  uint32_t* key;           // N: own cell of each particle, see KEY_BITS
  uint32_t* rank;          // N: place among the particles of its block
  uint32_t* block_ptr;     // Spad: block index of each domain slot, or NONE
  uint32_t* slot_count;    // Spad: particles whose own block is the slot
  uint32_t* block_list;    // cap: slot of each block in use
  uint32_t* block_nb;      // cap x 27: neighbour blocks (itself in the middle)
  uint32_t* block_start;   // cap + 1: first entry of each block in `sorted`
  uint32_t* sorted;        // N: particle indices grouped by block
  int32_t* acc;            // 19 x capc fixed-point transfer sums
  float* cf;               // 11 x capc solver fields
  float* bvec;             // capc: CG right-hand side (true-residual refresh)
  uint32_t* nbm;           // capc: bit k = face neighbour k is water, bit 6 = the cell is water
  double* part;            // cap x 4: per-block partial sums of the solver (memory form)
  uint32_t* tile_sum;      // 2 x ntiles: scan tile totals (slots claimed, particles)
  uint32_t* tile_off;      // 2 x ntiles: their exclusive prefix
  uint32_t* bar;           // control words: [2] the solve's abort flag a stuck wait raises,
                           // [4] its last reduction tag, [10] warm-start tiles this step, [11] 1
                           // when the solve starts from the warm start's guess, [12] particles
                           // that broke up this step; [256 + 32 x] and [512 + 32 x], x < 8: the
                           // counters that deal P2G's and G2P's blocks (reset by k_begin)
  unsigned long long* rwords;  // RED_WORDS: tagged chunk sums of the solve's reductions
  unsigned long long* hz;      // 3 x capc: tagged halo words of each water cell (the register
                               // forms and the memory form's streaming passes: z in the first
                               // third, the streaming passes' directions alternating between the
                               // other two; the pipelined form alternates between the first two)
  float* pcg;                  // 4 x capc: z, q and two m buffers of the pipelined solve's memory
                               // form (NATRAY_CG_PIPELINED only; else null)
  float* vmax_part;        // fastest particle of each g2p workgroup (<= cap workgroups)
  double* solver;          // double copies of rz, alpha, beta, b_norm, r_norm
  unsigned long long* stats;  // counters, see natray_sim_counters
  // Warm start of the pressure solve (the learned initial guess).
  float* guess;        // capc: the starting pressure the warm start offers
  float* ws_r0;        // capc: b - A p of the particle-carried pressure, per water cell
  float* ws_pcar;      // capc: the particle-carried pressure before the solve
  uint32_t* tile_map;  // tdx x tdy x tdz: 1 + the bits of the largest |r0| of a tile's water
                       // cells (16^3 cells per tile), 0 for a tile without water
  uint32_t* tile_list; // ws_max: the tiles with water this step, in tile order
  float* tile_scale;   // ws_max: each listed tile's largest |r0|
  float* ws_in;        // ws_max x 4096 x 2: the network's inputs
  float* ws_tgt;       // ws_max x 4096: training targets of the last solve (see natray_sim_ws_samples)
  float* ws_mask;      // ws_max x 4096: their water mask
  float* ws_diag;      // ws_max x 4096: the solve's diagonal there (0 where the mask is 0)
  uint32_t* ws_nbm;    // ws_max x 4096: its water-neighbour bits
  double* ws_part;     // cap x 2: per-block squared residuals of the guess and the carried pressure
  double* ws_stat;     // 8: see natray_sim_ws_stats
  uint32_t tdx, tdy, tdz, ws_max;
  // The rubber membrane's vertices as its own step left them (natray_sim_set_membrane): the
  // coupling kernels read positions, velocities, masses and each vertex's push on the water
  // (N) and area (m^2); mem_sample writes mem_water. mem_cap is 0 without a membrane.
  // The air boundary (NATRAY_SCENE_AIR): each cell's gas, field-major [5][capc] (density,
  // momentum x, y, z, total energy per volume), as the last step left it (air_x) and as
  // carried into this step's blocks (air_tmp); the last step's block list and, per domain slot,
  // its index in that list (stale entries are told apart by the list); blocks whose air was
  // active (set by the air's own substeps, which are not part of this library).
  float* air_x;
  float* air_tmp;
  uint32_t* prev_ptr;
  uint32_t* block_list_prev;
  uint32_t* block_active;
  // Spray and wet ground (NATRAY_SCENE_SPRAY): spray_cap records of 12 floats (position and
  // radius, velocity and weight, 0 0 0 and the time made), where each landed (4 floats), and the
  // wet map (wet_res^2 words: f32 bits of the first time water touched each ground cell,
  // 0xffffffff untouched). Null without the flag.
  float* spray;
  float* landed;
  float* breakups;  // N x 8: the particles G2P let break up this step (index bits, x, v, 0),
                    // bar[12] of them, which k_break_up turns into spray
  uint32_t* touch;
  uint32_t spray_cap, wet_res;
  const float4* mem_x;
  const float4* mem_v;
  const float* mem_mass;
  const float2* mem_push;
  float4* mem_water;
  uint32_t mem_cap;
  // Bullets (natray_sim_set_bullets): n_bullets of them, their patches ([n][576][8]) and cavity
  // sections ([n][32][12]); null and 0 without bullets.
  natray_bullet* bullets;
  const float* patches;
  float* sections;
  float* bullet_rec;  // [4][576][12] + [4][576]: each patch's share of bullet_step (k_bullet_patches)
  float* pose_log;    // [4][256][12]: each bullet's pose per step (the renderer's log)
  uint32_t n_bullets;
  const Dev* self;  // a device copy of this struct, for device functions that take it by pointer
  uint32_t N;
  uint32_t S;      // domain slots (product of bdims)
  uint32_t Spad;   // S rounded up to SCAN_TILE
  uint32_t ntiles; // Spad / SCAN_TILE
  uint32_t cap;    // block capacity (bdims.w)
  uint32_t capc;   // cap x 64 cells per field
  uint32_t fixed_ns;  // fixed step in ns, or 0
  uint32_t g2p_wgs;   // workgroups of the G2P launch (entries of vmax_part it writes)
  // The single scene's solve: 0 a cooperative launch, 1 a plain one, 2 a plain one whose last
  // workgroup stays out (the test of a workgroup that is not resident).
  uint32_t solve_plain;
};

// How a kernel finds its scene. One: a single scene, its Dev passed by value. Many: a batch of
// scenes in one launch; workgroup w works for scene map[w].scene as its map[w].wg-th of
// map[w].wgs workgroups, with that scene's Dev at scenes[map[w].scene] (device memory).
struct WgMap {
  uint32_t scene, wg, wgs, pad;
};
struct One {
  Dev d;
};
struct Many {
  const Dev* scenes;
  const WgMap* map;
};

// The workgroup counts of a launch, by kind of kernel: one per scene; one per 256 particles;
// one per scan tile; grid-stride over the cells to clear; blocks of P2G and G2P; one wave per
// block (grid, project, extend, warm start); the cooperative solve; one per 256 membrane
// vertices.
enum LaunchKind { LK_ONE, LK_PARTS, LK_TILES, LK_CELLS, LK_BLOCKS, LK_WAVES, LK_SOLVE, LK_VERTS, LK_KINDS };

// What a launch wrapper launches on: one scene (`d`, a host copy of its Dev) or a batch
// (`scenes` and one workgroup map per kind, in device memory), with the workgroups per kind.
struct Target {
  const Dev* d = nullptr;
  const Dev* scenes = nullptr;
  const WgMap* map[LK_KINDS] = {};
  uint32_t grid[LK_KINDS] = {};
  bool plain_solve = false;  // a batch's solve as a plain launch (a single scene: Dev::solve_plain)
};

}  // namespace natray

// Launch wrappers (natray_kernels.hip), on one scene or a batch (Target). Each returns the
// hipError_t of its launches as int.
namespace natray {
int launch_keys(const Target& t, void* stream);
int launch_begin(const Target& t, void* stream);
int launch_bin(const Target& t, void* stream, bool rubber = false, bool air = false);
int launch_p2g(const Target& t, void* stream);
int launch_grid_prep(const Target& t, void* stream);
// The solve of precision `kind` (NATRAY_CG_FP32, _MIXED or _PIPELINED).
int launch_solve(const Target& t, void* stream, int kind);
int solve_max_grid(int kind, int* per_cu);
int launch_project(const Target& t, void* stream);
int launch_extend(const Target& t, void* stream);
int launch_g2p(const Target& t, void* stream);
// Membrane coupling (one scene): the rubber's pull splatted to the grid after P2G, and the
// water sampled at its vertices after the velocity extension.
int launch_mem_splat(const Target& t, void* stream);
int launch_mem_sample(const Target& t, void* stream);
// The air boundary (one scene): the gas carried into this step's blocks after P2G, and the
// cells filled and the GAS and RUBBER pressures set after grid_prep.
int launch_air_carry(const Target& t, void* stream);
int launch_air_fill(const Target& t, void* stream);
// Spray (one scene): the drops fly, evaporate and land; before G2P, which makes new ones.
int launch_spray_update(const Target& t, void* stream);
// Spray (one scene): the particles G2P let break up become drops; after G2P.
int launch_break_up(const Target& t, void* stream);
// Bullets (one scene): their step and the water's push (after air_fill, before the solve).
int launch_bullets(const Target& t, void* stream);
// The keyframe words of the water: two per particle into `out` (device memory, 8 N bytes).
int launch_pack_water(const Target& t, void* stream, void* out);
// The spray index of the first n slots into `index` (4 + 2 cw words, cw = bit words of the
// spray capacity): [0] n, [1] live slots, the live bits at 4, their exclusive prefix at 4 + cw.
int launch_spray_index(const Target& t, void* stream, uint32_t n, uint32_t cw, uint32_t* index);
// A whole keyframe into `key` (device memory), its header words already written.
struct KeyframeArgs {
  void* key;
  uint32_t n_verts, n_tris, n_spray, cw;
  const uint32_t* alive;  // n_tris flags (device memory)
  const uint32_t* index;  // the spray index launch_spray_index built
  float* spray_w;         // n_spray x 2: each slot's weight and time made (may be null)
  int write_header;       // write the 4 header words from `index` (zeros without spray)
};
int launch_pack_keyframe(const Target& t, void* stream, const KeyframeArgs& a);
// The device address of this GPU's fixed spray-index array (NATRAY_KF_SCRATCH_WORDS words),
// looked up once per device with that device current.
#define NATRAY_MAX_DEVICES 64u
int kf_scratch(int device, uint32_t** out);

// One workgroup's share of natray_sim_water_stats (the host folds them in workgroup order).
struct WaterPart {
  double sum[7];  // x, y, z, vx, vy, vz, 0.5 |v|^2 of the alive particles
  float lo[3], hi[3];
  float vmax;
  uint32_t fastest;  // particle index of the fastest (the first among equal speeds), NONE with none
  uint32_t alive;
  uint32_t pad;
};
// The water diagnostics' kernels: particles[20 n] on `groups` workgroups, the cell table of
// `slots` keys and counts (a power of two >= 2 n, keys cleared to all ones and counts to 0
// beforehand), parts[groups], and *full (cleared beforehand) gets the cells with >= 4.
int launch_water_stats(const float* particles, uint32_t n, const float origin[3], float h, unsigned long long* keys,
                       uint32_t* counts, uint32_t slots, WaterPart* parts, uint32_t groups, uint32_t* full, void* stream);
// Warm start (one scene), before the solve: the network's inputs for at most `cap` tiles (also
// into the network's BF16 input buffer `net_in`, [cap][4096][8]), then, after the network has
// run on them (its output `pred`, [cap][4096][16] FP32), the guess and its acceptance. After a
// step, launch_ws_samples writes the training samples of its listed tiles.
int launch_ws_inputs(const Target& t, void* stream, uint32_t cap, void* net_in);
int launch_ws_guess(const Target& t, void* stream, const float* pred);
int launch_ws_samples(const Target& t, void* stream);
// Warm-start mode 3: the solve starts from the guess buffer (held at zero by the host).
int launch_ws_zero_start(const Target& t, void* stream);
}  // namespace natray
