// Host half of the MI355X water substep: the C ABI of include/natray.h.
//
// natray_sim_create allocates every device buffer once (the buffers of Sim::new,
// src/sim.rs:237-315); stepping enqueues the kernels of one step in plan order on the
// handle's stream (src/sim.rs:390-426) and reads the 96-byte State back once per batch
// (src/sim.rs:623-627). natray_sim_run and natray_sim_advance keep the reference's batch sizing
// and CG budget adaptation (src/sim.rs:498-522,636-645).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "natray.h"
#include "natray_dev.h"
#include "natray_net.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                                     \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) return fail(NATRAY_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define LAUNCHCHK(x)                                                                                           \
  do {                                                                                                         \
    int e_ = (x);                                                                                              \
    if (e_ != 0) return fail(NATRAY_E_HIP, std::string(#x) + ": " + hipGetErrorString((hipError_t)e_)); \
  } while (0)

constexpr double BATCH_MS = 4.0;  // src/sim.rs:24
constexpr size_t EVENT_POOL = 8192;

}  // namespace

struct natray_sim {
  natray_params U{};
  natray_scene sc{};
  natray::Dev D{};
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cus = 0;
  int grid_cells = 0, grid_blocks = 0, grid_waves = 0, solve_grid = 0;
  natray::Target tgt;  // this scene as the kernels' launch target
  uint32_t* packed = nullptr;  // 2 N words: the keyframe words of the water (natray_sim_pack_water)
  // natray_sim_pack_keyframe: the keyframe (kf_cap words), the spray index, the triangles' alive
  // flags (alive_cap) and each spray slot's weight and time made.
  uint32_t* kf = nullptr;
  size_t kf_cap = 0;
  uint32_t* kf_index = nullptr;
  uint32_t* kf_alive = nullptr;
  size_t alive_cap = 0;
  float* kf_spray_w = nullptr;
  int mixed = 0;  // the CG precision (NATRAY_CG_*)
  // Upper bound of the reduction tags the solves enqueued since the tagged words were last
  // cleared (see clear_tags).
  uint64_t tags_used = 0;
  std::vector<void*> allocs;
  // Pinned staging for the small host <-> device copies of a batch.
  natray_params* h_params = nullptr;
  natray_state* h_state = nullptr;
  uint32_t* h_words = nullptr;
  // Sim::advance / Sim::run host state (src/sim.rs:187-203,480).
  natray_state last{};
  uint32_t iters = 64;
  uint32_t unconverged = 0;
  double step_ms = 0.0;
  bool started = false;
  double t = 0.0;
  // The last NATRAY_RUN_LOG natray_sim_run calls (natray_sim_run_log).
  natray_run_record run_log[NATRAY_RUN_LOG] = {};
  uint64_t runs = 0;
  // Step graphs: the phases before the solve and those after it, captured on `stream`.
  bool use_graph = true;
  bool capturing = false;  // no timing events go into a captured graph
  hipGraphExec_t graph_pre = nullptr, graph_post = nullptr, graph_all = nullptr;
  hipGraphExec_t graph_many = nullptr;  // kGraphSteps steps of graph_all's kernels
  bool skip_tags = false;  // the step's caller did the solve's tag bookkeeping (graph_all)
  // Learned warm start of the solve (natray_sim_set_warm_start).
  int ws_mode = 0;
  natray_net* ws_net = nullptr;
  bool ws_ready = false;  // its buffers are allocated
  // Per-phase timing.
  bool prof = false;
  std::vector<hipEvent_t> events;
  std::vector<int> ev_phase;
  size_t ev_used = 0;
  double ms[NATRAY_N_PHASES] = {};
  uint32_t cnt[NATRAY_N_PHASES] = {};
};

namespace {

template <class T>
int dalloc(natray_sim* s, T** p, size_t count) {
  void* q = nullptr;
  size_t bytes = count * sizeof(T);
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(NATRAY_E_NOMEM, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
  s->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return NATRAY_OK;
}

void free_all(natray_sim* s) {
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->graph_pre) (void)hipGraphExecDestroy(s->graph_pre);
  if (s->graph_post) (void)hipGraphExecDestroy(s->graph_post);
  if (s->graph_all) (void)hipGraphExecDestroy(s->graph_all);
  if (s->graph_many) (void)hipGraphExecDestroy(s->graph_many);
  s->graph_all = s->graph_many = nullptr;
  s->graph_pre = s->graph_post = nullptr;
  for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
  s->events.clear();
  for (void* p : s->allocs) (void)hipFree(p);
  s->allocs.clear();
  if (s->h_params) (void)hipHostFree(s->h_params);
  if (s->h_state) (void)hipHostFree(s->h_state);
  if (s->h_words) (void)hipHostFree(s->h_words);
  s->h_params = nullptr;
  s->h_state = nullptr;
  s->h_words = nullptr;
  if (s->stream) (void)hipStreamDestroy(s->stream);
  s->stream = nullptr;
}

int harvest_events(natray_sim* s) {
  for (size_t k = 0; k < s->ev_used; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, s->events[2 * k], s->events[2 * k + 1]));
    s->ms[s->ev_phase[k]] += ms;
    s->cnt[s->ev_phase[k]] += 1;
  }
  s->ev_used = 0;
  return NATRAY_OK;
}

// Workgroups of the cooperative solve: one per CU. The block count of a step is only known
// on the device, so the launch does not follow it: with MAX_CHUNKS <= CUs every workgroup
// holds at most one chunk and the solve keeps its blocks in registers up to 32 blocks per
// chunk; a workgroup left without a chunk only takes part in the reductions. Any count gives
// the same result.
int solve_grid_for(const natray_sim* s) {
  if (s->sc.solve_groups) return (int)std::min<uint32_t>(s->sc.solve_groups, (uint32_t)s->solve_grid);
  return s->solve_grid;
}

// Zeroes every tagged word of the solve and its tag counter, so tags start again from 1.
int clear_tags(natray_sim* s) {
  HIPCHK(hipMemsetAsync(s->D.rwords, 0, natray::RED_WORDS * 8, s->stream));
  HIPCHK(hipMemsetAsync(s->D.hz, 0, (size_t)s->D.capc * 24, s->stream));
  HIPCHK(hipMemsetAsync(s->D.bar + 4, 0, 4, s->stream));
  s->tags_used = 0;
  return NATRAY_OK;
}

// A solve launch uses at most one tag per reduction: one before the iterations and up to
// four per iteration (the iteration cap is min(limits.z, 2^24) on the device). The pipelined
// form numbers its halo words 8 t + k from reduction tag t, so tags stay below 2^28 (the
// words' 32-bit tag field then never wraps).
constexpr uint64_t kTagLimit = 1ull << 28;
uint64_t solve_tag_bound(const natray_sim* s) {
  return 4ull * std::min<uint64_t>(s->U.limits[2], 1ull << 24) + 4ull;
}

// The warm start before a solve: the network's inputs, and in mode 2 the network and the guess
// the solve starts from when it lowers the residual.
int enqueue_warm_start(natray_sim* s) {
  const natray::Dev& D = s->D;
  void* st = s->stream;
  if (s->ws_mode == 3) {
    LAUNCHCHK(natray::launch_ws_zero_start(s->tgt, st));
    return NATRAY_OK;
  }
  const bool guess = s->ws_mode == 2;
  const uint32_t cap = guess ? std::min(D.ws_max, natray::net::max_tiles(s->ws_net)) : D.ws_max;
  LAUNCHCHK(natray::launch_ws_inputs(s->tgt, st, cap, guess ? natray::net::input_buffer(s->ws_net) : nullptr));
  if (!guess) return NATRAY_OK;
  int e = natray::net::forward_device(s->ws_net, cap, D.bar + 10, st);
  if (e) return fail(e, std::string("warm-start network: ") + natray_net_last_error());
  LAUNCHCHK(natray::launch_ws_guess(s->tgt, st, natray::net::output_buffer(s->ws_net)));
  return NATRAY_OK;
}

// The warm start's buffers, allocated on first use; the device copy of Dev follows and the
// step graphs are captured again with it.
int alloc_warm_start(natray_sim* s) {
  natray::Dev& D = s->D;
  D.tdx = (s->U.bdims[0] + 3u) / 4u;
  D.tdy = (s->U.bdims[1] + 3u) / 4u;
  D.tdz = (s->U.bdims[2] + 3u) / 4u;
  const size_t T = (size_t)D.tdx * D.tdy * D.tdz;
  // Tiles that can hold water: every tile of the domain, and no more than 64 per block
  // of capacity (a block lies in one tile).
  D.ws_max = (uint32_t)std::min<size_t>(T, D.cap);
  const size_t capc = D.capc, W = D.ws_max;
  int e;
  if ((e = dalloc(s, &D.guess, capc)) || (e = dalloc(s, &D.ws_r0, capc)) || (e = dalloc(s, &D.ws_pcar, capc)) ||
      (e = dalloc(s, &D.tile_map, T)) || (e = dalloc(s, &D.tile_list, W)) ||
      (e = dalloc(s, &D.tile_scale, W)) || (e = dalloc(s, &D.ws_in, W * 4096 * 2)) || (e = dalloc(s, &D.ws_tgt, W * 4096)) ||
      (e = dalloc(s, &D.ws_mask, W * 4096)) || (e = dalloc(s, &D.ws_diag, W * 4096)) ||
      (e = dalloc(s, &D.ws_nbm, W * 4096)) || (e = dalloc(s, &D.ws_part, (size_t)D.cap * 2)) ||
      (e = dalloc(s, &D.ws_stat, 8)))
    return e;
  HIPCHK(hipMemsetAsync(D.guess, 0, capc * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.ws_r0, 0, capc * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.ws_pcar, 0, capc * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.tile_map, 0, T * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.tile_list, 0, W * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.tile_scale, 0, W * 4, s->stream));
  HIPCHK(hipMemsetAsync(D.ws_part, 0, (size_t)D.cap * 2 * 8, s->stream));
  HIPCHK(hipMemcpyAsync(const_cast<natray::Dev*>(D.self), &D, sizeof(natray::Dev), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->graph_pre) (void)hipGraphExecDestroy(s->graph_pre);
  if (s->graph_post) (void)hipGraphExecDestroy(s->graph_post);
  if (s->graph_all) (void)hipGraphExecDestroy(s->graph_all);
  if (s->graph_many) (void)hipGraphExecDestroy(s->graph_many);
  s->graph_all = s->graph_many = nullptr;
  s->graph_pre = s->graph_post = nullptr;
  s->ws_ready = true;
  return NATRAY_OK;
}

int enqueue_phase(natray_sim* s, int phase) {
  void* st = s->stream;
  size_t slot = 0;
  bool timed = s->prof && !s->capturing && s->ev_used < s->events.size() / 2;
  if (timed) {
    slot = s->ev_used++;
    s->ev_phase[slot] = phase;
    HIPCHK(hipEventRecord(s->events[2 * slot], s->stream));
  }
  switch (phase) {
    case NATRAY_PHASE_BEGIN: LAUNCHCHK(natray::launch_begin(s->tgt, st)); break;
    case NATRAY_PHASE_BIN: LAUNCHCHK(natray::launch_bin(s->tgt, st, s->D.mem_cap > 0, s->D.air_x != nullptr)); break;
    case NATRAY_PHASE_P2G:
      LAUNCHCHK(natray::launch_p2g(s->tgt, st));
      if (s->D.air_x) LAUNCHCHK(natray::launch_air_carry(s->tgt, st));
      if (s->D.mem_cap) LAUNCHCHK(natray::launch_mem_splat(s->tgt, st));
      break;
    case NATRAY_PHASE_GRID:
      LAUNCHCHK(natray::launch_grid_prep(s->tgt, st));
      if (s->D.air_x) LAUNCHCHK(natray::launch_air_fill(s->tgt, st));
      // bullet_step and bullet_apply come after air_fill and before solve_init (src/sim.rs:406-409).
      if (s->D.n_bullets) LAUNCHCHK(natray::launch_bullets(s->tgt, st));
      break;
    case NATRAY_PHASE_SOLVE: {
      // Tags are 32-bit: clear the tagged words well before a tag could repeat.
      if (!s->skip_tags) {
        const uint64_t bound = solve_tag_bound(s);
        if (s->tags_used + bound >= kTagLimit) {
          int e = clear_tags(s);
          if (e) return e;
        }
        s->tags_used += bound;
      }
      if (s->ws_mode) {
        int e = enqueue_warm_start(s);
        if (e) return e;
      }
      LAUNCHCHK(natray::launch_solve(s->tgt, st, s->mixed));
      break;
    }
    case NATRAY_PHASE_PROJECT: LAUNCHCHK(natray::launch_project(s->tgt, st)); break;
    case NATRAY_PHASE_EXTEND:
      LAUNCHCHK(natray::launch_extend(s->tgt, st));
      if (s->D.mem_cap) LAUNCHCHK(natray::launch_mem_sample(s->tgt, st));
      break;
    case NATRAY_PHASE_G2P:
      // spray_update runs before g2p (src/sim.rs:425), so drops made in a step move in the next.
      if (s->D.spray) LAUNCHCHK(natray::launch_spray_update(s->tgt, st));
      LAUNCHCHK(natray::launch_g2p(s->tgt, st));
      if (s->D.spray) LAUNCHCHK(natray::launch_break_up(s->tgt, st));
      break;
    default: return fail(NATRAY_E_ARG, "unknown phase " + std::to_string(phase));
  }
  if (timed) HIPCHK(hipEventRecord(s->events[2 * slot + 1], s->stream));
  return NATRAY_OK;
}

// Captures phases [lo, hi) as a graph on the handle's stream.
int capture_phases(natray_sim* s, int lo, int hi, hipGraphExec_t* out, uint32_t reps = 1) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int e = NATRAY_OK;
  s->capturing = true;
  for (uint32_t r = 0; r < reps && e == NATRAY_OK; ++r)
    for (int ph = lo; ph < hi && e == NATRAY_OK; ++ph) e = enqueue_phase(s, ph);
  s->capturing = false;
  hipError_t ec = hipStreamEndCapture(s->stream, &g);
  if (e) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  if (ec != hipSuccess) return fail(NATRAY_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  ec = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ec != hipSuccess) return fail(NATRAY_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
  return NATRAY_OK;
}

// The tags of `steps` solves, cleared first if a tag could repeat within them.
int reserve_tags(natray_sim* s, uint32_t steps) {
  const uint64_t bound = solve_tag_bound(s) * steps;
  if (s->tags_used + bound >= kTagLimit) {
    int e = clear_tags(s);
    if (e) return e;
  }
  s->tags_used += bound;
  return NATRAY_OK;
}

// Steps replay one graph of every phase (no events, no warm start).
bool one_graph(const natray_sim* s) { return s->use_graph && !s->ws_mode && !s->prof; }

// One step: one captured graph of every phase, its solve a cooperative kernel node; or, with
// profiling or the warm start on, two graphs around the solve launched between them (its
// events, and the warm start's host work, stay outside the graphs; profiling then times only
// the solve); or every phase launched on its own. The solve's tag bookkeeping is host work, so
// with the single graph it is done here before each replay.
int enqueue_step(natray_sim* s) {
  if (!s->use_graph) {
    for (int ph = 0; ph < NATRAY_N_PHASES; ++ph) {
      int e = enqueue_phase(s, ph);
      if (e) return e;
    }
    return NATRAY_OK;
  }
  if (one_graph(s)) {
    int e = reserve_tags(s, 1);
    if (e) return e;
    if (!s->graph_all) {
      s->skip_tags = true;
      int e = capture_phases(s, NATRAY_PHASE_BEGIN, NATRAY_N_PHASES, &s->graph_all);
      s->skip_tags = false;
      if (e) return e;
    }
    HIPCHK(hipGraphLaunch(s->graph_all, s->stream));
    return NATRAY_OK;
  }
  if (!s->graph_pre) {
    int e = capture_phases(s, NATRAY_PHASE_BEGIN, NATRAY_PHASE_SOLVE, &s->graph_pre);
    if (e) return e;
    e = capture_phases(s, NATRAY_PHASE_PROJECT, NATRAY_N_PHASES, &s->graph_post);
    if (e) return e;
  }
  HIPCHK(hipGraphLaunch(s->graph_pre, s->stream));
  int e = enqueue_phase(s, NATRAY_PHASE_SOLVE);
  if (e) return e;
  HIPCHK(hipGraphLaunch(s->graph_post, s->stream));
  return NATRAY_OK;
}

// n steps: with the single-graph launch, runs of kGraphSteps steps replay one graph that holds
// them all (no launch gap between its steps), the rest one step at a time.
constexpr uint32_t kGraphSteps = 8;
int enqueue_steps(natray_sim* s, uint32_t n) {
  if (one_graph(s) && n >= kGraphSteps) {
    if (!s->graph_many) {
      s->skip_tags = true;
      int e = capture_phases(s, NATRAY_PHASE_BEGIN, NATRAY_N_PHASES, &s->graph_many, kGraphSteps);
      s->skip_tags = false;
      if (e) return e;
    }
    for (; n >= kGraphSteps; n -= kGraphSteps) {
      int e = reserve_tags(s, kGraphSteps);
      if (e) return e;
      HIPCHK(hipGraphLaunch(s->graph_many, s->stream));
    }
  }
  for (; n > 0; --n) {
    int e = enqueue_step(s);
    if (e) return e;
  }
  return NATRAY_OK;
}

int upload_params(natray_sim* s) {
  *s->h_params = s->U;
  HIPCHK(hipMemcpyAsync(const_cast<natray_params*>(s->D.U), s->h_params, sizeof(natray_params), hipMemcpyHostToDevice, s->stream));
  return NATRAY_OK;
}

// Reads the State back and checks the solver's barrier; the stream is idle afterwards. A lost
// barrier's message names `launch_call`, the call that switches the launch, when the solve was
// a plain launch (`plain`; < 0: the scene's own setting).
int sync_state(natray_sim* s, const char* launch_call = "natray_sim_set_solve_launch(sim, ", int plain = -1) {
  HIPCHK(hipMemcpyAsync(s->h_state, s->D.st, sizeof(natray_state), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_words, s->D.bar, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->last = *s->h_state;
  if (s->prof) {
    int e = harvest_events(s);
    if (e) return e;
  }
  if (s->h_words[2] != 0u) {
    // Workgroups may have stopped at different reductions: start the tags afresh.
    HIPCHK(hipMemsetAsync(s->D.bar, 0, 1024 * sizeof(uint32_t), s->stream));
    int e = clear_tags(s);
    if (e) return e;
    HIPCHK(hipStreamSynchronize(s->stream));
    if (plain < 0) plain = s->D.solve_plain != 0u;
    return fail(NATRAY_E_SOLVER,
                plain ? std::string("the pressure solve lost its grid barrier: not all of its workgroups were resident "
                                    "(another program on the GPU?); ") +
                            launch_call + "NATRAY_SOLVE_COOPERATIVE) launches it cooperatively"
                      : std::string("the pressure solve lost its grid barrier (not all workgroups were resident)"));
  }
  return NATRAY_OK;
}

}  // namespace

extern "C" {

void natray_abi_sizes(uint32_t out[4]) {
  out[0] = sizeof(natray_params);
  out[1] = sizeof(natray_state);
  out[2] = sizeof(natray_particle);
  out[3] = sizeof(natray_scene);
}

const char* natray_last_error(void) { return g_err.c_str(); }

int natray_sim_create(const natray_params* params, const float* particles, uint32_t n, const natray_scene* scene,
                      const natray_state* init, natray_sim** out) {
  if (!params || !out || (n && !particles)) return fail(NATRAY_E_ARG, "null argument");
  *out = nullptr;
  if (params->dims[3] != n) return fail(NATRAY_E_ARG, "params.dims.w must equal the particle count");
  for (int k = 0; k < 3; ++k) {
    if (params->dims[k] == 0 || params->dims[k] > natray::MAX_DIM)
      return fail(NATRAY_E_ARG, "domain sides must be 1..1024 cells");
    if (params->bdims[k] * 4u != params->dims[k]) return fail(NATRAY_E_ARG, "bdims must be dims / 4 (dims multiples of 4)");
  }
  if (params->bdims[3] == 0) return fail(NATRAY_E_ARG, "block capacity bdims.w must be positive");
  if (!(params->origin[3] > 0.0f)) return fail(NATRAY_E_ARG, "cell size origin.w must be positive");
  uint64_t S64 = (uint64_t)params->bdims[0] * params->bdims[1] * params->bdims[2];
  if (S64 >= 0xfffff000ull) return fail(NATRAY_E_ARG, "domain too large");
  if ((uint64_t)params->bdims[3] * 64ull * NATRAY_ACC_FIELDS >= (1ull << 40)) return fail(NATRAY_E_ARG, "capacity too large");

  natray_sim* s = new natray_sim();
  s->U = *params;
  if (scene) {
    s->sc = *scene;
  } else {
    std::memset(&s->sc, 0, sizeof s->sc);
    s->sc.flags = NATRAY_SCENE_REFERENCE;
  }
  if (s->sc.precond > NATRAY_PRECOND_JACOBI || s->sc.precision > NATRAY_CG_PIPELINED) {
    delete s;
    return fail(NATRAY_E_ARG, "unknown preconditioner or precision");
  }
  s->mixed = (int)s->sc.precision;
  auto bail = [&](int code) {
    free_all(s);
    delete s;
    return code;
  };
#define TRY(x)              \
  do {                      \
    int e__ = (x);          \
    if (e__) return bail(e__); \
  } while (0)
#define HIPTRY(x)                                                                                   \
  do {                                                                                              \
    hipError_t e__ = (x);                                                                           \
    if (e__ != hipSuccess) return bail(fail(NATRAY_E_HIP, std::string(#x) + ": " + hipGetErrorString(e__))); \
  } while (0)

  HIPTRY(hipGetDevice(&s->device));
  HIPTRY(hipDeviceGetAttribute(&s->num_cus, hipDeviceAttributeMultiprocessorCount, s->device));
  int coop = 0;
  HIPTRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, s->device));
  if (!coop) return bail(fail(NATRAY_E_HIP, "the device cannot launch cooperative kernels"));
  HIPTRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  HIPTRY(hipHostMalloc((void**)&s->h_params, sizeof(natray_params), hipHostMallocDefault));
  HIPTRY(hipHostMalloc((void**)&s->h_state, sizeof(natray_state), hipHostMallocDefault));
  HIPTRY(hipHostMalloc((void**)&s->h_words, 16 * sizeof(uint32_t), hipHostMallocDefault));

  natray::Dev& D = s->D;
  D.N = n;
  D.S = (uint32_t)S64;
  D.ntiles = (uint32_t)((S64 + natray::SCAN_TILE - 1) / natray::SCAN_TILE);
  D.Spad = D.ntiles * natray::SCAN_TILE;
  D.cap = params->bdims[3];
  D.capc = D.cap * 64u;
  D.fixed_ns = 0;
  if (s->sc.fixed_dt > 0.0f) {
    double ns = std::floor((double)s->sc.fixed_dt * 1e9 + 0.5);
    D.fixed_ns = ns < 20.0 ? 20u : (ns > 4e9 ? 4000000000u : (uint32_t)ns);
  }
  const size_t capc = D.capc;
  natray_params* dU = nullptr;
  natray_scene* dSc = nullptr;
  TRY(dalloc(s, &dU, 1));
  TRY(dalloc(s, &dSc, 1));
  D.U = dU;
  D.sc = dSc;
  TRY(dalloc(s, &D.st, 1));
  TRY(dalloc(s, &D.P, (size_t)n * 20));
  TRY(dalloc(s, &D.key, (size_t)n + 1));
  TRY(dalloc(s, &D.rank, (size_t)n + 1));
  TRY(dalloc(s, &D.block_ptr, D.Spad));
  TRY(dalloc(s, &D.slot_count, D.Spad));
  TRY(dalloc(s, &D.block_list, (size_t)D.cap + 1));
  TRY(dalloc(s, &D.block_nb, (size_t)D.cap * 27));
  TRY(dalloc(s, &D.block_start, (size_t)D.cap + 2));
  TRY(dalloc(s, &D.sorted, (size_t)n + 1));
  TRY(dalloc(s, &D.acc, NATRAY_ACC_FIELDS * capc));
  TRY(dalloc(s, &D.cf, NATRAY_CELL_FIELDS * capc));
  TRY(dalloc(s, &D.bvec, capc));
  TRY(dalloc(s, &D.nbm, capc));
  TRY(dalloc(s, &D.part, (size_t)D.cap * 4));
  TRY(dalloc(s, &D.rwords, natray::RED_WORDS));
  TRY(dalloc(s, &D.hz, 3 * capc));
  if (s->sc.precision == NATRAY_CG_PIPELINED) TRY(dalloc(s, &D.pcg, 4 * capc));
  TRY(dalloc(s, &D.tile_sum, (size_t)D.ntiles * 2));
  TRY(dalloc(s, &D.tile_off, (size_t)D.ntiles * 2));
  TRY(dalloc(s, &D.bar, 1024));
  TRY(dalloc(s, &D.vmax_part, (size_t)D.cap));
  TRY(dalloc(s, &D.solver, 8));
  TRY(dalloc(s, &D.stats, 8));
  if (s->sc.flags & NATRAY_SCENE_AIR) {
    TRY(dalloc(s, &D.air_x, 5 * capc));
    TRY(dalloc(s, &D.air_tmp, 5 * capc));
    TRY(dalloc(s, &D.prev_ptr, D.Spad));
    TRY(dalloc(s, &D.block_list_prev, (size_t)D.cap + 1));
    TRY(dalloc(s, &D.block_active, (size_t)D.cap + 1));
    HIPTRY(hipMemsetAsync(D.air_x, 0, 5 * capc * 4, s->stream));
    HIPTRY(hipMemsetAsync(D.air_tmp, 0, 5 * capc * 4, s->stream));
    HIPTRY(hipMemsetAsync(D.prev_ptr, 0xff, (size_t)D.Spad * 4, s->stream));
    HIPTRY(hipMemsetAsync(D.block_list_prev, 0, ((size_t)D.cap + 1) * 4, s->stream));
    HIPTRY(hipMemsetAsync(D.block_active, 0, ((size_t)D.cap + 1) * 4, s->stream));
  }
  if (s->sc.flags & NATRAY_SCENE_SPRAY) {
    // src/sim.rs:252-254: spray and landed start zeroed, the wet map at u32::MAX.
    D.spray_cap = s->U.limits[1];
    D.wet_res = s->U.limits[0];
    const size_t ww = (size_t)D.wet_res * D.wet_res;
    TRY(dalloc(s, &D.spray, (size_t)D.spray_cap * 12));
    TRY(dalloc(s, &D.landed, (size_t)D.spray_cap * 4));
    TRY(dalloc(s, &D.touch, ww));
    TRY(dalloc(s, &D.breakups, (size_t)std::max<uint32_t>(D.N, 1u) * 8));
    HIPTRY(hipMemsetAsync(D.spray, 0, (size_t)D.spray_cap * 48, s->stream));
    HIPTRY(hipMemsetAsync(D.landed, 0, (size_t)D.spray_cap * 16, s->stream));
    HIPTRY(hipMemsetAsync(D.touch, 0xff, ww * 4, s->stream));
  }
  natray::Dev* dself = nullptr;
  TRY(dalloc(s, &dself, 1));
  D.self = dself;

  // Launch shapes: per-block kernels loop over the blocks in use, which only the GPU knows.
  {
    const uint32_t cap = D.cap;
    s->grid_cells = s->num_cus * 4;
    s->grid_blocks = (int)std::min<uint32_t>(cap, (uint32_t)s->num_cus * 8u);
    s->grid_waves = (int)std::min<uint32_t>((cap + 3u) / 4u, (uint32_t)s->num_cus * 8u);
    int per_cu = 0;
    int e = natray::solve_max_grid(s->mixed, &per_cu);
    if (e || per_cu <= 0) return bail(fail(NATRAY_E_HIP, "cannot size the cooperative solve"));
    s->solve_grid = std::min(per_cu * s->num_cus, s->num_cus);
    D.g2p_wgs = (uint32_t)s->grid_blocks;
    natray::Target& t = s->tgt;
    t.d = &s->D;
    t.grid[natray::LK_ONE] = 1;
    t.grid[natray::LK_PARTS] = n ? (n + 255u) / 256u : 1u;
    t.grid[natray::LK_TILES] = D.ntiles;
    t.grid[natray::LK_CELLS] = (uint32_t)s->grid_cells;
    t.grid[natray::LK_BLOCKS] = (uint32_t)s->grid_blocks;
    t.grid[natray::LK_WAVES] = (uint32_t)s->grid_waves;
    t.grid[natray::LK_SOLVE] = (uint32_t)solve_grid_for(s);
    t.grid[natray::LK_VERTS] = 1;
    D.solve_plain = 1u;  // natray_sim_set_solve_launch
  }

  HIPTRY(hipMemsetAsync(D.block_ptr, 0xff, (size_t)D.Spad * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.slot_count, 0, (size_t)D.Spad * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.block_list, 0, ((size_t)D.cap + 1) * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.block_nb, 0xff, (size_t)D.cap * 27 * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.block_start, 0, ((size_t)D.cap + 2) * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.acc, 0, NATRAY_ACC_FIELDS * capc * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.cf, 0, NATRAY_CELL_FIELDS * capc * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.bvec, 0, capc * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.nbm, 0, capc * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.part, 0, (size_t)D.cap * 4 * 8, s->stream));
  HIPTRY(hipMemsetAsync(D.bar, 0, 1024 * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.rwords, 0, natray::RED_WORDS * 8, s->stream));
  HIPTRY(hipMemsetAsync(D.hz, 0, capc * 24, s->stream));
  HIPTRY(hipMemsetAsync(D.vmax_part, 0, (size_t)D.cap * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.solver, 0, 8 * 8, s->stream));
  HIPTRY(hipMemsetAsync(D.stats, 0, 8 * 8, s->stream));
  HIPTRY(hipMemsetAsync(D.sorted, 0, ((size_t)n + 1) * 4, s->stream));
  HIPTRY(hipMemsetAsync(D.rank, 0, ((size_t)n + 1) * 4, s->stream));
  if (n) HIPTRY(hipMemcpyAsync(D.P, particles, (size_t)n * 80, hipMemcpyHostToDevice, s->stream));
  HIPTRY(hipMemcpyAsync(dself, &D, sizeof(natray::Dev), hipMemcpyHostToDevice, s->stream));
  HIPTRY(hipMemcpyAsync(dSc, &s->sc, sizeof(natray_scene), hipMemcpyHostToDevice, s->stream));

  natray_state st0{};
  if (init) {
    st0 = *init;
  } else {
    // src/sim.rs:240-241
    float c_air = std::sqrt(1.4f * 101325.0f / 1.204f);
    st0.mem_dt = FLT_MAX;
    st0.cap_dt = FLT_MAX;
    st0.air_dt = 0.9f * params->origin[3] / (3.0f * c_air);
  }
  *s->h_state = st0;
  HIPTRY(hipMemcpyAsync(D.st, s->h_state, sizeof(natray_state), hipMemcpyHostToDevice, s->stream));
  TRY(upload_params(s));
  int e = natray::launch_keys(s->tgt, s->stream);
  if (e) return bail(fail(NATRAY_E_HIP, std::string("keys: ") + hipGetErrorString((hipError_t)e)));
  HIPTRY(hipStreamSynchronize(s->stream));
  s->last = st0;
  *out = s;
  return NATRAY_OK;
#undef TRY
#undef HIPTRY
}

void natray_sim_destroy(natray_sim* s) {
  if (!s) return;
  free_all(s);
  delete s;
}

int natray_sim_phase(natray_sim* s, int phase) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  int e = enqueue_phase(s, phase);
  if (e) return e;
  e = sync_state(s);
  if (e) return e;
  return s->last.overflow ? fail(NATRAY_E_OVERFLOW, "the water needs more grid blocks than the capacity") : NATRAY_OK;
}

int natray_sim_run(natray_sim* s, uint32_t n_steps, uint32_t iters, uint32_t stop_ns, natray_state* out) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  // The pose log keeps one entry per step at step % NATRAY_BULLET_LOG; the reference reads it
  // after every batch of at most MAX_BATCH = LOG_CAP steps (src/sim.rs:22,514,628).
  if (s->D.n_bullets && n_steps > NATRAY_BULLET_LOG)
    return fail(NATRAY_E_ARG, "with bullets set, at most NATRAY_BULLET_LOG steps per run (the pose log would wrap)");
  // src/sim.rs:586-593
  s->U.limits[2] = iters ? iters : s->iters;
  s->U.limits[3] = stop_ns;
  int e = upload_params(s);
  if (e) return e;
  s->h_words[4] = 0;
  s->h_words[5] = 0;
  HIPCHK(hipMemcpyAsync(&s->D.st->iters_max, &s->h_words[4], 2 * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
  const uint32_t step_before = s->last.step;
  auto t0 = std::chrono::steady_clock::now();
  e = enqueue_steps(s, n_steps);
  if (e) return e;
  e = sync_state(s);
  if (e) return e;
  if (n_steps) s->step_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / n_steps;
  if (out) *out = s->last;
  const natray_state& st = s->last;
  natray_run_record& rec = s->run_log[s->runs % NATRAY_RUN_LOG];
  rec = natray_run_record{n_steps, s->U.limits[2], stop_ns, st.step - step_before, st.iters_max, st.unconverged, st.t_ns, 0};
  ++s->runs;
  if (st.overflow) return fail(NATRAY_E_OVERFLOW, "the water needs more grid blocks than the capacity");
  // src/sim.rs:636-645
  s->unconverged += st.unconverged;
  s->iters = natray_next_iters(s->iters, st.iters_max, st.unconverged);
  return NATRAY_OK;
}

int natray_sim_run_log(natray_sim* s, uint32_t max_records, natray_run_record* out, uint32_t* count) {
  if (!s || !count) return fail(NATRAY_E_ARG, "null argument");
  const uint64_t kept = std::min<uint64_t>(s->runs, NATRAY_RUN_LOG);
  const uint32_t n = (uint32_t)std::min<uint64_t>(kept, max_records);
  *count = (uint32_t)kept;
  if (n && !out) return fail(NATRAY_E_ARG, "null argument");
  // The last n records, oldest first.
  for (uint32_t k = 0; k < n; ++k) out[k] = s->run_log[(s->runs - n + k) % NATRAY_RUN_LOG];
  return NATRAY_OK;
}

uint32_t natray_next_iters(uint32_t iters, uint32_t iters_max, uint32_t unconverged) {
  if (unconverged > 0) return std::min(iters * 2u, NATRAY_MAX_ITERS);
  if (iters_max > 0) return std::min(std::max(iters_max + iters_max / 4u + 2u, 8u), NATRAY_MAX_ITERS);
  return iters;
}

uint32_t natray_plan_batch(uint32_t stop_ns, uint32_t t_ns, uint32_t free_dt_ns, double step_ms, uint32_t budget_left) {
  if (t_ns >= stop_ns || budget_left == 0) return 0;
  const uint64_t dt = free_dt_ns > 0 ? free_dt_ns : (uint64_t)(NATRAY_DT_MAX * 1e9);
  const uint64_t fit = step_ms > 0.0 ? (uint64_t)(BATCH_MS / step_ms) : 1;
  uint64_t hi = std::min<uint64_t>(budget_left, NATRAY_MAX_BATCH);
  hi = std::min<uint64_t>(hi, std::max<uint64_t>(fit, 1));
  const uint64_t want = ((uint64_t)stop_ns - t_ns + dt - 1) / dt;
  return (uint32_t)std::max<uint64_t>(1, std::min(want, hi));
}

int natray_sim_advance(natray_sim* s, double t_target, uint32_t budget, uint32_t* taken_out) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  // src/sim.rs:498-522 with the handoff time t0 = 0.
  const double t0 = 0.0;
  if (!s->started) s->started = true;
  double span = t_target - t0;
  if (span > 4.0) span = 4.0;
  if (span < 0.0) span = 0.0;
  const uint32_t stop = (uint32_t)std::llround(span * 1e9);
  uint32_t taken = 0;
  while (taken < budget && s->last.t_ns < stop) {
    const uint32_t nb = natray_plan_batch(stop, s->last.t_ns, s->last.free_dt_ns, s->step_ms, budget - taken);
    const uint32_t before = s->last.step;
    int e = natray_sim_run(s, nb, 0, stop, nullptr);
    if (e) return e;
    taken += s->last.step - before;
  }
  s->t = t0 + s->last.t_ns * 1e-9;
  if (taken_out) *taken_out = taken;
  return NATRAY_OK;
}

int natray_sim_step(natray_sim* s, natray_state* out) { return natray_sim_run(s, 1, 0, 0xffffffffu, out); }

int natray_sim_read_state(natray_sim* s, natray_state* out) {
  if (!s || !out) return fail(NATRAY_E_ARG, "null argument");
  int e = sync_state(s);
  if (e) return e;
  *out = s->last;
  return NATRAY_OK;
}

int natray_sim_write_state(natray_sim* s, const natray_state* in) {
  if (!s || !in) return fail(NATRAY_E_ARG, "null argument");
  *s->h_state = *in;
  HIPCHK(hipMemcpyAsync(s->D.st, s->h_state, sizeof(natray_state), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->last = *in;
  return NATRAY_OK;
}

int natray_sim_read_particles(natray_sim* s, float* dst) {
  if (!s || (!dst && s->D.N)) return fail(NATRAY_E_ARG, "null argument");
  if (s->D.N) HIPCHK(hipMemcpyAsync(dst, s->D.P, (size_t)s->D.N * 80, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_write_particles(natray_sim* s, const float* src) {
  if (!s || (!src && s->D.N)) return fail(NATRAY_E_ARG, "null argument");
  if (s->D.N) HIPCHK(hipMemcpyAsync(s->D.P, src, (size_t)s->D.N * 80, hipMemcpyHostToDevice, s->stream));
  LAUNCHCHK(natray::launch_keys(s->tgt, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_write_params(natray_sim* s, const natray_params* params) {
  if (!s || !params) return fail(NATRAY_E_ARG, "null argument");
  for (int k = 0; k < 4; ++k)
    if (params->dims[k] != s->U.dims[k] || params->bdims[k] != s->U.bdims[k])
      return fail(NATRAY_E_ARG, "the domain and capacity are fixed at creation");
  s->U = *params;
  int e = upload_params(s);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_params(natray_sim* s, natray_params* out) {
  if (!s || !out) return fail(NATRAY_E_ARG, "null argument");
  *out = s->U;
  return NATRAY_OK;
}

int natray_sim_read_grid(natray_sim* s, uint32_t* n_blocks, uint32_t* block_list, uint32_t* block_nb, uint32_t* block_start,
                         int32_t* acc, float* cf) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  int e = sync_state(s);
  if (e) return e;
  const uint32_t n = std::min(s->last.n_blocks, s->U.bdims[3]);
  const size_t nc = (size_t)n * 64u, capc = s->D.capc;
  if (n_blocks) *n_blocks = n;
  if (block_list && n) HIPCHK(hipMemcpyAsync(block_list, s->D.block_list, (size_t)n * 4, hipMemcpyDeviceToHost, s->stream));
  if (block_nb && n) HIPCHK(hipMemcpyAsync(block_nb, s->D.block_nb, (size_t)n * 27 * 4, hipMemcpyDeviceToHost, s->stream));
  if (block_start) HIPCHK(hipMemcpyAsync(block_start, s->D.block_start, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost, s->stream));
  if (acc && n)
    for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f)
      HIPCHK(hipMemcpyAsync(acc + f * nc, s->D.acc + f * capc, nc * 4, hipMemcpyDeviceToHost, s->stream));
  if (cf && n)
    for (uint32_t f = 0; f < NATRAY_CELL_FIELDS; ++f)
      HIPCHK(hipMemcpyAsync(cf + f * nc, s->D.cf + f * capc, nc * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_write_grid(natray_sim* s, const int32_t* acc, const float* cf) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  int e = sync_state(s);
  if (e) return e;
  const uint32_t n = std::min(s->last.n_blocks, s->U.bdims[3]);
  const size_t nc = (size_t)n * 64u, capc = s->D.capc;
  if (acc && n)
    for (uint32_t f = 0; f < NATRAY_ACC_FIELDS; ++f)
      HIPCHK(hipMemcpyAsync(s->D.acc + f * capc, acc + f * nc, nc * 4, hipMemcpyHostToDevice, s->stream));
  if (cf && n)
    for (uint32_t f = 0; f < NATRAY_CELL_FIELDS; ++f)
      HIPCHK(hipMemcpyAsync(s->D.cf + f * capc, cf + f * nc, nc * 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_bins(natray_sim* s, uint32_t* block_ptr, uint32_t* sorted) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  if (block_ptr) HIPCHK(hipMemcpyAsync(block_ptr, s->D.block_ptr, (size_t)s->D.S * 4, hipMemcpyDeviceToHost, s->stream));
  if (sorted && s->D.N) HIPCHK(hipMemcpyAsync(sorted, s->D.sorted, (size_t)s->D.N * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

void* natray_sim_device_ptr(natray_sim* s, int which) {
  if (!s) return nullptr;
  switch (which) {
    case NATRAY_BUF_PARTICLES: return s->D.P;
    case NATRAY_BUF_STATE: return s->D.st;
    case NATRAY_BUF_ACC: return s->D.acc;
    case NATRAY_BUF_CF: return s->D.cf;
    case NATRAY_BUF_BLOCK_LIST: return s->D.block_list;
    case NATRAY_BUF_STREAM: return (void*)s->stream;
    default: return nullptr;
  }
}

int natray_sim_set_profiling(natray_sim* s, int on) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (on && s->events.empty()) {
    s->events.resize(EVENT_POOL);
    s->ev_phase.resize(EVENT_POOL / 2);
    for (size_t k = 0; k < EVENT_POOL; ++k) HIPCHK(hipEventCreate(&s->events[k]));
  }
  s->prof = on != 0;
  s->ev_used = 0;
  for (int k = 0; k < NATRAY_N_PHASES; ++k) {
    s->ms[k] = 0.0;
    s->cnt[k] = 0;
  }
  return NATRAY_OK;
}

int natray_sim_set_solve_launch(natray_sim* s, int mode) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  if (mode != NATRAY_SOLVE_PLAIN && mode != NATRAY_SOLVE_COOPERATIVE && mode != NATRAY_SOLVE_PLAIN_TEST_CHECK)
    return fail(NATRAY_E_ARG, "unknown solve launch mode");
  s->D.solve_plain = mode == NATRAY_SOLVE_COOPERATIVE ? 0u : (mode == NATRAY_SOLVE_PLAIN ? 1u : 2u);
  // The graphs hold the Dev by value: capture them again.
  HIPCHK(hipMemcpyAsync(const_cast<natray::Dev*>(s->D.self), &s->D, sizeof(natray::Dev), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->graph_pre) (void)hipGraphExecDestroy(s->graph_pre);
  if (s->graph_post) (void)hipGraphExecDestroy(s->graph_post);
  if (s->graph_all) (void)hipGraphExecDestroy(s->graph_all);
  if (s->graph_many) (void)hipGraphExecDestroy(s->graph_many);
  s->graph_all = s->graph_many = nullptr;
  s->graph_pre = s->graph_post = nullptr;
  return NATRAY_OK;
}

int natray_sim_set_graph(natray_sim* s, int on) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  s->use_graph = on != 0;
  return NATRAY_OK;
}

int natray_sim_phase_times(natray_sim* s, double* ms, uint32_t* counts) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->prof) {
    int e = harvest_events(s);
    if (e) return e;
  }
  for (int k = 0; k < NATRAY_N_PHASES; ++k) {
    if (ms) ms[k] = s->ms[k];
    if (counts) counts[k] = s->cnt[k];
  }
  return NATRAY_OK;
}

int natray_sim_set_warm_start(natray_sim* s, natray_net* net, int mode) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  if (mode < 0 || mode > 3) return fail(NATRAY_E_ARG, "warm-start mode must be 0, 1, 2 or 3");
  if (mode == 2 && !net) return fail(NATRAY_E_ARG, "warm-start mode 2 needs a network");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (mode && !s->ws_ready) {
    int e = alloc_warm_start(s);
    if (e) return e;
  }
  s->ws_mode = mode;
  s->ws_net = mode == 2 ? net : nullptr;
  // Mode 3 starts every solve from the guess buffer held at zero (nothing else writes it then).
  if (mode == 3) HIPCHK(hipMemsetAsync(s->D.guess, 0, (size_t)s->D.capc * 4, s->stream));
  if (s->ws_ready) {
    HIPCHK(hipMemsetAsync(s->D.ws_stat, 0, 8 * 8, s->stream));
    HIPCHK(hipMemsetAsync(s->D.bar + 10, 0, 2 * 4, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  return NATRAY_OK;
}

int natray_sim_ws_samples(natray_sim* s, float* inputs, float* targets, float* masks, float* diag, uint32_t* nbm,
                          uint32_t max_tiles, uint32_t* tiles) {
  if (!s || !tiles) return fail(NATRAY_E_ARG, "null argument");
  if (!s->ws_ready) return fail(NATRAY_E_ARG, "the warm start was never enabled");
  LAUNCHCHK(natray::launch_ws_samples(s->tgt, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_words + 10, s->D.bar + 10, 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const uint32_t n = std::min(s->h_words[10], max_tiles);
  const size_t vox = (size_t)n * 4096;
  if (inputs && n) HIPCHK(hipMemcpyAsync(inputs, s->D.ws_in, vox * 2 * 4, hipMemcpyDeviceToHost, s->stream));
  if (targets && n) HIPCHK(hipMemcpyAsync(targets, s->D.ws_tgt, vox * 4, hipMemcpyDeviceToHost, s->stream));
  if (masks && n) HIPCHK(hipMemcpyAsync(masks, s->D.ws_mask, vox * 4, hipMemcpyDeviceToHost, s->stream));
  if (diag && n) HIPCHK(hipMemcpyAsync(diag, s->D.ws_diag, vox * 4, hipMemcpyDeviceToHost, s->stream));
  if (nbm && n) HIPCHK(hipMemcpyAsync(nbm, s->D.ws_nbm, vox * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *tiles = n;
  return NATRAY_OK;
}

int natray_sim_set_membrane(natray_sim* s, uint32_t n, const float* x, const float* v, const float* mass,
                            const float* push) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  if (n && (!x || !v || !mass || !push)) return fail(NATRAY_E_ARG, "null argument");
  HIPCHK(hipStreamSynchronize(s->stream));
  natray::Dev& D = s->D;
  if (n > D.mem_cap) {
    float4 *mx = nullptr, *mv = nullptr, *mw = nullptr;
    float* mm = nullptr;
    float2* mp = nullptr;
    int e;
    if ((e = dalloc(s, &mx, n)) || (e = dalloc(s, &mv, n)) || (e = dalloc(s, &mm, n)) || (e = dalloc(s, &mp, n)) ||
        (e = dalloc(s, &mw, n)))
      return e;
    HIPCHK(hipMemsetAsync(mw, 0, (size_t)n * sizeof(float4), s->stream));
    D.mem_x = mx;
    D.mem_v = mv;
    D.mem_mass = mm;
    D.mem_push = mp;
    D.mem_water = mw;
    D.mem_cap = n;
    s->tgt.grid[natray::LK_VERTS] = (n + 255u) / 256u;
    HIPCHK(hipMemcpyAsync(const_cast<natray::Dev*>(D.self), &D, sizeof(natray::Dev), hipMemcpyHostToDevice, s->stream));
    if (s->graph_pre) (void)hipGraphExecDestroy(s->graph_pre);
    if (s->graph_post) (void)hipGraphExecDestroy(s->graph_post);
    if (s->graph_all) (void)hipGraphExecDestroy(s->graph_all);
    if (s->graph_many) (void)hipGraphExecDestroy(s->graph_many);
    s->graph_all = s->graph_many = nullptr;
    s->graph_pre = s->graph_post = nullptr;
  }
  if (n) {
    HIPCHK(hipMemcpyAsync(const_cast<float4*>(D.mem_x), x, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(const_cast<float4*>(D.mem_v), v, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(const_cast<float*>(D.mem_mass), mass, (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(const_cast<float2*>(D.mem_push), push, (size_t)n * 8, hipMemcpyHostToDevice, s->stream));
  }
  s->U.mem[0] = n;
  int e = upload_params(s);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_set_bullets(natray_sim* s, uint32_t n, const natray_bullet* bullets, const float* patches) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  if (n > NATRAY_MAX_BULLETS) return fail(NATRAY_E_ARG, "at most NATRAY_MAX_BULLETS bullets");
  if (n && (!bullets || !patches)) return fail(NATRAY_E_ARG, "null argument");
  for (uint32_t k = 0; k < n; ++k)
    if (bullets[k].info[1] > NATRAY_BULLET_PATCHES) return fail(NATRAY_E_ARG, "a bullet has more than NATRAY_BULLET_PATCHES patches");
  HIPCHK(hipStreamSynchronize(s->stream));
  natray::Dev& D = s->D;
  if (n && !(D.bullets && D.patches && D.sections && D.bullet_rec && D.pose_log)) {
    // All five buffers or none: a failed allocation leaves the handle without bullets, and a
    // later call allocates them again (what did succeed is freed with the handle).
    int e;
    natray_bullet* b = nullptr;
    float *pt = nullptr, *sec = nullptr, *rec = nullptr, *log = nullptr;
    if ((e = dalloc(s, &b, NATRAY_MAX_BULLETS)) || (e = dalloc(s, &pt, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_PATCHES * 8)) ||
        (e = dalloc(s, &sec, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_SECTIONS * 12)) ||
        (e = dalloc(s, &rec, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_PATCHES * 13)) ||
        (e = dalloc(s, &log, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_LOG * 12)))
      return e;
    D.bullets = b;
    D.patches = pt;
    D.sections = sec;
    D.bullet_rec = rec;
    D.pose_log = log;
  }
  if (n) {
    HIPCHK(hipMemcpyAsync(D.bullets, bullets, (size_t)n * sizeof(natray_bullet), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(const_cast<float*>(D.patches), patches, (size_t)n * NATRAY_BULLET_PATCHES * 32,
                          hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemsetAsync(D.sections, 0, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_SECTIONS * 48, s->stream));
    HIPCHK(hipMemsetAsync(D.pose_log, 0, (size_t)NATRAY_MAX_BULLETS * NATRAY_BULLET_LOG * 48, s->stream));
  }
  // Without bullets nothing clears their acc fields any more, and the solve reads A_SRC.
  if (!n && D.n_bullets)
    HIPCHK(hipMemsetAsync(D.acc + (size_t)NATRAY_A_BU * D.capc, 0, (size_t)4 * D.capc * 4, s->stream));
  // The bullet count is a launch argument (G2P's carve, the bullets' grid): the step graphs
  // are captured again.
  D.n_bullets = n;
  HIPCHK(hipMemcpyAsync(const_cast<natray::Dev*>(D.self), &D, sizeof(natray::Dev), hipMemcpyHostToDevice, s->stream));
  if (s->graph_pre) (void)hipGraphExecDestroy(s->graph_pre);
  if (s->graph_post) (void)hipGraphExecDestroy(s->graph_post);
  if (s->graph_all) (void)hipGraphExecDestroy(s->graph_all);
  if (s->graph_many) (void)hipGraphExecDestroy(s->graph_many);
  s->graph_all = s->graph_many = nullptr;
  s->graph_pre = s->graph_post = nullptr;
  s->U.mem[3] = n;
  int e = upload_params(s);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_bullets(natray_sim* s, natray_bullet* bullets, float* sections) {
  if (!s) return fail(NATRAY_E_ARG, "null handle");
  const uint32_t n = s->D.n_bullets;
  if (n && bullets) HIPCHK(hipMemcpyAsync(bullets, s->D.bullets, (size_t)n * sizeof(natray_bullet), hipMemcpyDeviceToHost, s->stream));
  if (n && sections)
    HIPCHK(hipMemcpyAsync(sections, s->D.sections, (size_t)n * NATRAY_BULLET_SECTIONS * 48, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_pose_log(natray_sim* s, float* log) {
  if (!s || !log) return fail(NATRAY_E_ARG, "null argument");
  const uint32_t n = s->D.n_bullets;
  if (n) HIPCHK(hipMemcpyAsync(log, s->D.pose_log, (size_t)n * NATRAY_BULLET_LOG * 48, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_air(natray_sim* s, float* air) {
  if (!s || !air) return fail(NATRAY_E_ARG, "null argument");
  if (!s->D.air_x) return fail(NATRAY_E_ARG, "the scene has no air boundary (NATRAY_SCENE_AIR)");
  const size_t nc = (size_t)std::min(s->last.n_blocks, s->D.cap) * 64;
  for (int f = 0; f < 5; ++f)
    if (nc) HIPCHK(hipMemcpyAsync(air + f * nc, s->D.air_x + (size_t)f * s->D.capc, nc * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_spray(natray_sim* s, uint32_t max_records, float* spray, float* landed, uint32_t* count) {
  if (!s || !count) return fail(NATRAY_E_ARG, "null argument");
  if (!s->D.spray) return fail(NATRAY_E_ARG, "the scene has no spray (NATRAY_SCENE_SPRAY)");
  HIPCHK(hipStreamSynchronize(s->stream));
  int e = sync_state(s);
  if (e) return e;
  *count = s->last.spray_count;
  const size_t n = std::min(std::min(s->last.spray_count, s->D.spray_cap), max_records);
  if (n && spray) HIPCHK(hipMemcpyAsync(spray, s->D.spray, n * 48, hipMemcpyDeviceToHost, s->stream));
  if (n && landed) HIPCHK(hipMemcpyAsync(landed, s->D.landed, n * 16, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_wet(natray_sim* s, uint32_t* touch) {
  if (!s || !touch) return fail(NATRAY_E_ARG, "null argument");
  if (!s->D.touch) return fail(NATRAY_E_ARG, "the scene has no wet map (NATRAY_SCENE_SPRAY)");
  const size_t ww = (size_t)s->D.wet_res * s->D.wet_res;
  if (ww) HIPCHK(hipMemcpyAsync(touch, s->D.touch, ww * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_pack_water(natray_sim* s, uint32_t* words) {
  if (!s || !words) return fail(NATRAY_E_ARG, "null argument");
  if (!s->D.N) return NATRAY_OK;
  if (!s->packed) {
    int e = dalloc(s, &s->packed, (size_t)s->D.N * 2);
    if (e) return e;
  }
  LAUNCHCHK(natray::launch_pack_water(s->tgt, s->stream, s->packed));
  HIPCHK(hipMemcpyAsync(words, s->packed, (size_t)s->D.N * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_pack_keyframe(natray_sim* s, uint32_t n_tris, const uint32_t* alive, uint32_t* words, uint32_t max_words,
                             uint32_t* n_words, float* spray_w) {
  if (!s || !n_words) return fail(NATRAY_E_ARG, "null argument");
  if (n_tris && !alive) return fail(NATRAY_E_ARG, "null alive flags");
  natray::Dev& D = s->D;
  HIPCHK(hipStreamSynchronize(s->stream));
  int e = sync_state(s);
  if (e) return e;
  const uint32_t n_verts = std::min(s->U.mem[0], D.mem_cap);
  const uint32_t n_spray = D.spray ? std::min(s->last.spray_count, D.spray_cap) : 0u;
  const uint32_t cw = (D.spray_cap + 31u) / 32u;
  uint32_t head[4] = {n_spray, 0u, 0u, 0u};
  if (D.spray) {
    if (!s->kf_index) {
      if ((e = dalloc(s, &s->kf_index, 4 + 2 * (size_t)cw)) || (e = dalloc(s, &s->kf_spray_w, 2 * (size_t)std::max(D.spray_cap, 1u))))
        return e;
    }
    LAUNCHCHK(natray::launch_spray_index(s->tgt, s->stream, n_spray, cw, s->kf_index));
    HIPCHK(hipMemcpyAsync(head, s->kf_index, 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  const uint64_t total = 4ull + 2ull * ((uint64_t)D.N + n_verts) + (n_tris + 31ull) / 32ull + 2ull * ((n_spray + 31ull) / 32ull) +
                         3ull * head[1];
  if (total > 0xffffffffull) return fail(NATRAY_E_ARG, "keyframe larger than 2^32 words");
  *n_words = (uint32_t)total;
  if (!words) return NATRAY_OK;
  if (max_words < total) return fail(NATRAY_E_ARG, "the keyframe needs " + std::to_string(total) + " words");
  if (s->kf_cap < total) {
    const size_t cap = std::max<size_t>(total, s->kf_cap + s->kf_cap / 2);
    if (s->kf) {
      s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), static_cast<void*>(s->kf)));
      HIPCHK(hipFree(s->kf));
      s->kf = nullptr;
    }
    if ((e = dalloc(s, &s->kf, cap))) return e;
    s->kf_cap = cap;
  }
  if (n_tris > s->alive_cap) {
    if (s->kf_alive) {
      s->allocs.erase(std::find(s->allocs.begin(), s->allocs.end(), static_cast<void*>(s->kf_alive)));
      HIPCHK(hipFree(s->kf_alive));
      s->kf_alive = nullptr;
    }
    if ((e = dalloc(s, &s->kf_alive, n_tris))) return e;
    s->alive_cap = n_tris;
  }
  if (n_tris) HIPCHK(hipMemcpyAsync(s->kf_alive, alive, (size_t)n_tris * 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->kf, head, 16, hipMemcpyHostToDevice, s->stream));
  natray::KeyframeArgs a{s->kf, n_verts, n_tris, n_spray, cw, s->kf_alive, s->kf_index, s->kf_spray_w};
  LAUNCHCHK(natray::launch_pack_keyframe(s->tgt, s->stream, a));
  HIPCHK(hipMemcpyAsync(words, s->kf, (size_t)total * 4, hipMemcpyDeviceToHost, s->stream));
  if (spray_w && n_spray) HIPCHK(hipMemcpyAsync(spray_w, s->kf_spray_w, (size_t)n_spray * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_pack_keyframe_static(natray_sim* s, uint32_t n_tris, const uint32_t* alive, uint32_t* words,
                                    uint32_t max_words, uint32_t* n_words, float* spray_w) {
  if (!s || !n_words) return fail(NATRAY_E_ARG, "null argument");
  if (n_tris && !alive) return fail(NATRAY_E_ARG, "null alive flags");
  natray::Dev& D = s->D;
  const uint32_t cw = (D.spray_cap + 31u) / 32u;
  if (D.spray && 4ull + 2ull * cw > NATRAY_KF_SCRATCH_WORDS)
    return fail(NATRAY_E_ARG, "the scene's spray capacity is larger than the fixed spray index holds (NATRAY_SPRAY_CAP slots)");
  uint32_t* index = nullptr;
  int e = natray::kf_scratch(s->device, &index);
  if (e) return fail(NATRAY_E_HIP, std::string("spray index array: ") + hipGetErrorString((hipError_t)e));
  HIPCHK(hipStreamSynchronize(s->stream));
  e = sync_state(s);
  if (e) return e;
  const uint32_t n_verts = std::min(s->U.mem[0], D.mem_cap);
  const uint32_t n_spray = D.spray ? std::min(s->last.spray_count, D.spray_cap) : 0u;
  uint32_t live = 0;
  if (D.spray) {
    // The index is rebuilt on every call: the array keeps whatever the last call on this GPU left.
    LAUNCHCHK(natray::launch_spray_index(s->tgt, s->stream, n_spray, cw, index));
    HIPCHK(hipMemcpyAsync(s->h_words + 12, index, 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    live = s->h_words[13];
  }
  const uint64_t total = 4ull + 2ull * ((uint64_t)D.N + n_verts) + (n_tris + 31ull) / 32ull + 2ull * ((n_spray + 31ull) / 32ull) +
                         3ull * live;
  if (total > 0xffffffffull) return fail(NATRAY_E_ARG, "keyframe larger than 2^32 words");
  *n_words = (uint32_t)total;
  if (!words) return NATRAY_OK;
  if (max_words < total) return fail(NATRAY_E_ARG, "the keyframe needs " + std::to_string(total) + " words");
  natray::KeyframeArgs a{words, n_verts, n_tris, n_spray, cw, alive, D.spray ? index : nullptr, spray_w, 1};
  LAUNCHCHK(natray::launch_pack_keyframe(s->tgt, s->stream, a));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_read_membrane(natray_sim* s, float* water) {
  if (!s || !water) return fail(NATRAY_E_ARG, "null argument");
  const uint32_t n = std::min(s->U.mem[0], s->D.mem_cap);
  if (n) HIPCHK(hipMemcpyAsync(water, s->D.mem_water, (size_t)n * 16, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_ws_stats(natray_sim* s, double out[8]) {
  if (!s || !out) return fail(NATRAY_E_ARG, "null argument");
  for (int k = 0; k < 8; ++k) out[k] = 0.0;
  if (!s->ws_ready) return NATRAY_OK;
  HIPCHK(hipMemcpyAsync(out, s->D.ws_stat, 8 * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return NATRAY_OK;
}

int natray_sim_ws_guess(natray_sim* s, float* guess, uint32_t* accepted) {
  if (!s || !guess || !accepted) return fail(NATRAY_E_ARG, "null argument");
  if (!s->ws_ready) return fail(NATRAY_E_ARG, "the warm start was never enabled");
  HIPCHK(hipMemcpyAsync(guess, s->D.guess, (size_t)s->D.capc * 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(s->h_words + 11, s->D.bar + 11, 4, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *accepted = s->h_words[11];
  return NATRAY_OK;
}

int natray_sim_counters(natray_sim* s, uint64_t out[8]) {
  if (!s || !out) return fail(NATRAY_E_ARG, "null argument");
  unsigned long long h[8] = {};
  HIPCHK(hipMemcpyAsync(h, s->D.stats, sizeof h, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 8; ++k) out[k] = h[k];
  return NATRAY_OK;
}

int natray_sim_water_stats(natray_sim* s, natray_water_stats* out) {
  if (!s || !out) return fail(NATRAY_E_ARG, "null argument");
  natray_water_stats r{};
  for (int a = 0; a < 3; ++a) {
    r.lo[a] = FLT_MAX;
    r.hi[a] = -FLT_MAX;
  }
  const uint32_t n = s->D.N;
  if (n == 0) {
    *out = r;
    return NATRAY_OK;
  }
  uint32_t slots = 1024;
  while (slots < 2u * n) slots <<= 1;
  const uint32_t groups = std::min<uint32_t>(1024u, (n + 255u) / 256u);
  // One allocation: the keys, the counts, the workgroups' parts and the full-cell counter.
  const size_t off_counts = (size_t)slots * 8, off_parts = off_counts + (size_t)slots * 4,
               off_full = off_parts + (size_t)groups * sizeof(natray::WaterPart), bytes = off_full + 256;
  char* buf = nullptr;
  hipError_t e = hipMalloc(&buf, bytes);
  if (e != hipSuccess) return fail(NATRAY_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  std::vector<natray::WaterPart> parts(groups);
  uint32_t full = 0;
  int rc = NATRAY_OK;
  auto step = [&](hipError_t x) {
    if (x != hipSuccess && rc == NATRAY_OK) rc = fail(NATRAY_E_HIP, std::string("HIP: ") + hipGetErrorString(x));
    return rc == NATRAY_OK;
  };
  auto* keys = reinterpret_cast<unsigned long long*>(buf);
  auto* counts = reinterpret_cast<uint32_t*>(buf + off_counts);
  auto* dparts = reinterpret_cast<natray::WaterPart*>(buf + off_parts);
  auto* dfull = reinterpret_cast<uint32_t*>(buf + off_full);
  const float origin[3] = {s->U.origin[0], s->U.origin[1], s->U.origin[2]};
  if (step(hipMemsetAsync(keys, 0xff, (size_t)slots * 8, s->stream)) &&
      step(hipMemsetAsync(counts, 0, (size_t)slots * 4, s->stream)) && step(hipMemsetAsync(dfull, 0, 4, s->stream)) &&
      step((hipError_t)natray::launch_water_stats(s->D.P, n, origin, s->U.origin[3], keys, counts, slots, dparts, groups,
                                                  dfull, s->stream)) &&
      step(hipMemcpyAsync(parts.data(), dparts, (size_t)groups * sizeof(natray::WaterPart), hipMemcpyDeviceToHost,
                          s->stream)) &&
      step(hipMemcpyAsync(&full, dfull, 4, hipMemcpyDeviceToHost, s->stream)))
    step(hipStreamSynchronize(s->stream));
  const hipError_t ef = hipFree(buf);
  if (rc != NATRAY_OK) return rc;
  if (ef != hipSuccess) return fail(NATRAY_E_HIP, std::string("hipFree: ") + hipGetErrorString(ef));
  // The workgroups in order: sums, box, and the fastest (the lowest index among equal speeds).
  double sum[7] = {};
  uint32_t fast = 0xffffffffu;
  for (const natray::WaterPart& p : parts) {
    for (int k = 0; k < 7; ++k) sum[k] += p.sum[k];
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = std::fmin(r.lo[a], p.lo[a]);
      r.hi[a] = std::fmax(r.hi[a], p.hi[a]);
    }
    if (p.vmax > r.vmax || (p.vmax == r.vmax && p.fastest < fast)) {
      r.vmax = p.vmax;
      fast = p.fastest;
    }
    r.alive += p.alive;
  }
  if (r.alive) {
    for (int a = 0; a < 3; ++a) {
      r.mean[a] = sum[a] / r.alive;
      r.vmean[a] = sum[3 + a] / r.alive;
    }
  }
  r.ke = sum[6];
  if (fast != 0xffffffffu && r.vmax > 0.0f) {
    float rec[4];
    HIPCHK(hipMemcpyAsync(rec, s->D.P + (size_t)fast * 20, 16, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (int a = 0; a < 3; ++a) r.fastest[a] = rec[a];
  } else {
    r.vmax = 0.0f;
  }
  r.full_cells = full;
  const double dx = s->U.origin[3];
  r.volume_l = (double)full * dx * dx * dx * 1e3;
  *out = r;
  return NATRAY_OK;
}

}  // extern "C"

// ---- batches of independent scenes --------------------------------------------------------------

struct natray_batch {
  std::vector<natray_sim*> sims;
  hipStream_t stream = nullptr;
  int num_cus = 0;
  int mixed = 0;
  bool air = false;      // some scene has the air boundary (the kernels skip the others)
  int solve_budget = 0;  // co-resident workgroups of the cooperative solve
  natray_batch_shape shape{};
  natray::Dev* scenes = nullptr;                   // device table, one Dev per scene
  natray::WgMap* maps[natray::LK_KINDS] = {};       // device workgroup maps, per kind
  size_t map_cap[natray::LK_KINDS] = {};            // their capacities (entries)
  std::vector<uint32_t> counts[natray::LK_KINDS];  // workgroups per scene, per kind
  natray::Target tgt;
  bool use_graph = true;
  // The step's graph (every phase) and one of kGraphSteps steps; the host does the solve's tag
  // bookkeeping before each replay (skip_tags while capturing).
  hipGraphExec_t graph_all = nullptr, graph_many = nullptr;
  bool skip_tags = false;
  std::vector<void*> allocs;
  // Per-phase timing (natray_batch_phase_times): events around each phase of kernel-by-kernel
  // steps, summed when the run ends.
  bool prof = false;
  std::vector<hipEvent_t> events;
  std::vector<int> ev_phase;
  size_t ev_used = 0;
  double ms[NATRAY_N_PHASES] = {};
  uint32_t cnt[NATRAY_N_PHASES] = {};
};

namespace {

void batch_drop_graphs(natray_batch* b) {
  if (b->graph_all) (void)hipGraphExecDestroy(b->graph_all);
  if (b->graph_many) (void)hipGraphExecDestroy(b->graph_many);
  b->graph_all = b->graph_many = nullptr;
}

void batch_free(natray_batch* b) {
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  b->events.clear();
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graphs(b);
  for (void* p : b->allocs) (void)hipFree(p);
  b->allocs.clear();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  b->stream = nullptr;
}

// Shares of `budget` in proportion to `w`, each within [lo, hi[k]]: x_k = clamp(t w_k, lo, hi_k)
// with t chosen (bisection; x is monotone in t) so that the x_k sum to the budget, then the
// floors of x_k plus one more for the largest fractional parts (ties to the lower index). When
// the bounds cannot meet the budget, every share sits at its bound.
void plan_shares(const std::vector<uint64_t>& w, const std::vector<uint32_t>& hi, uint32_t budget, uint32_t lo,
                 std::vector<uint32_t>& out) {
  const size_t n = w.size();
  out.assign(n, 0);
  std::vector<double> x(n);
  auto fill = [&](double t) {
    double sum = 0.0;
    for (size_t k = 0; k < n; ++k) {
      const double h = (double)hi[k], l = std::min((double)lo, h);
      x[k] = std::min(h, std::max(l, t * (double)std::max<uint64_t>(w[k], 1)));
      sum += x[k];
    }
    return sum;
  };
  double t_lo = 0.0, t_hi = 1.0;
  if (fill(0.0) >= (double)budget) {
    fill(0.0);
  } else {
    while (fill(t_hi) < (double)budget && t_hi < 1e30) t_hi *= 2.0;
    for (int it = 0; it < 200; ++it) {
      const double mid = 0.5 * (t_lo + t_hi);
      (fill(mid) < (double)budget ? t_lo : t_hi) = mid;
    }
    fill(t_hi);
  }
  uint64_t used = 0;
  std::vector<std::pair<double, size_t>> frac;
  for (size_t k = 0; k < n; ++k) {
    out[k] = (uint32_t)std::floor(x[k] + 1e-9);
    used += out[k];
    if (out[k] < hi[k]) frac.push_back({x[k] - (double)out[k], k});
  }
  std::stable_sort(frac.begin(), frac.end(),
                   [](const std::pair<double, size_t>& a, const std::pair<double, size_t>& b) { return a.first > b.first; });
  for (size_t i = 0; i < frac.size() && used < budget; ++i) {
    out[frac[i].second] += 1;
    used += 1;
  }
}

// Workgroup counts per scene and kind for the batch's shape, the scenes' Dev copies (with their
// G2P workgroup counts) and the workgroup maps, uploaded; the step graphs are dropped.
int batch_layout(natray_batch* b) {
  const size_t n = b->sims.size();
  std::vector<uint64_t> wpart(n);
  std::vector<uint32_t> cap_blocks(n), cap_waves(n), unbounded(n, 0xffffffffu);
  for (size_t k = 0; k < n; ++k) {
    const natray_sim* s = b->sims[k];
    wpart[k] = std::max<uint64_t>(s->D.N, 1);
    cap_blocks[k] = s->D.cap;
    cap_waves[k] = (s->D.cap + 3u) / 4u;
  }
  const uint32_t cus = (uint32_t)b->num_cus;
  b->counts[natray::LK_ONE].assign(n, 1u);
  b->counts[natray::LK_VERTS].assign(n, 1u);
  auto& parts = b->counts[natray::LK_PARTS];
  auto& tiles = b->counts[natray::LK_TILES];
  parts.resize(n);
  tiles.resize(n);
  for (size_t k = 0; k < n; ++k) {
    const natray_sim* s = b->sims[k];
    parts[k] = s->D.N ? (s->D.N + 255u) / 256u : 1u;
    tiles[k] = s->D.ntiles;
  }
  plan_shares(wpart, unbounded, std::max<uint32_t>(b->shape.cells_per_cu * cus, (uint32_t)n), 1, b->counts[natray::LK_CELLS]);
  plan_shares(wpart, cap_blocks, std::max<uint32_t>(b->shape.blocks_per_cu * cus, (uint32_t)n), 1, b->counts[natray::LK_BLOCKS]);
  plan_shares(wpart, cap_waves, std::max<uint32_t>(b->shape.waves_per_cu * cus, (uint32_t)n), 1, b->counts[natray::LK_WAVES]);
  plan_shares(wpart, unbounded, (uint32_t)b->solve_budget, 1, b->counts[natray::LK_SOLVE]);
  uint64_t solve_total = 0;
  for (uint32_t c : b->counts[natray::LK_SOLVE]) solve_total += c;
  if (solve_total > (uint64_t)b->solve_budget)
    return fail(NATRAY_E_ARG, "a batch holds at most one scene per co-resident solve workgroup");
  // Scene table.
  std::vector<natray::Dev> table(n);
  for (size_t k = 0; k < n; ++k) {
    table[k] = b->sims[k]->D;
    table[k].g2p_wgs = b->counts[natray::LK_BLOCKS][k];
  }
  HIPCHK(hipMemcpyAsync(b->scenes, table.data(), n * sizeof(natray::Dev), hipMemcpyHostToDevice, b->stream));
  // Workgroup maps.
  for (int kind = 0; kind < natray::LK_KINDS; ++kind) {
    std::vector<natray::WgMap> m;
    for (size_t k = 0; k < n; ++k)
      for (uint32_t w = 0; w < b->counts[kind][k]; ++w) m.push_back({(uint32_t)k, w, b->counts[kind][k], 0u});
    if (m.size() > b->map_cap[kind]) {
      void* q = nullptr;
      hipError_t e = hipMalloc(&q, m.size() * sizeof(natray::WgMap));
      if (e != hipSuccess) return fail(NATRAY_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
      b->allocs.push_back(q);
      b->maps[kind] = static_cast<natray::WgMap*>(q);
      b->map_cap[kind] = m.size();
    }
    HIPCHK(hipMemcpyAsync(b->maps[kind], m.data(), m.size() * sizeof(natray::WgMap), hipMemcpyHostToDevice, b->stream));
    b->tgt.map[kind] = b->maps[kind];
    b->tgt.grid[kind] = (uint32_t)m.size();
  }
  b->tgt.d = nullptr;
  b->tgt.scenes = b->scenes;
  HIPCHK(hipStreamSynchronize(b->stream));
  batch_drop_graphs(b);
  return NATRAY_OK;
}

int batch_phase_launch(natray_batch* b, int phase);

// The solve's workgroups from each scene's last solve, once every scene has run a step (a new
// plan drops the step graphs, which hold the solve's grid). A scene's solve lasts
// about (iterations + 1) x (kSolveFixed + the most blocks one of its workgroups holds): every
// iteration pays its two reductions whatever the workgroup holds (kSolveFixed, in blocks of
// pass time: about 15 us against 0.3 us per block, measured), and a workgroup holds whole
// chunks of K blocks. Starting from one workgroup each, the scene that would take longest gets
// the workgroups that take one chunk off each of its workgroups, until they do not fit in the
// co-resident workgroups.
constexpr uint64_t kSolveFixed = 48;
int batch_plan_solve(natray_batch* b) {
  const size_t n = b->sims.size();
  std::vector<uint64_t> it(n);
  std::vector<uint32_t> K(n), m(n), shares(n, 1u);
  for (size_t k = 0; k < n; ++k) {
    const natray_state& st = b->sims[k]->last;
    if (st.step == 0 || st.n_blocks == 0) return NATRAY_OK;
    const uint32_t nb = std::min(st.n_blocks, b->sims[k]->D.cap);
    K[k] = (nb + natray::MAX_CHUNKS - 1u) / natray::MAX_CHUNKS;
    m[k] = (nb + K[k] - 1u) / K[k];
    it[k] = (uint64_t)st.solve_iters + 1u;
  }
  auto cost = [&](size_t k, uint32_t G) {
    return it[k] * (kSolveFixed + (uint64_t)((m[k] + G - 1u) / G) * K[k]);
  };
  uint32_t used = (uint32_t)n;
  while (true) {
    size_t worst = 0;
    for (size_t k = 1; k < n; ++k)
      if (cost(k, shares[k]) > cost(worst, shares[worst])) worst = k;
    // The fewest workgroups that hold one chunk fewer each (a workgroup holds whole chunks);
    // none once every workgroup holds one, or when they do not fit.
    const uint32_t most = (m[worst] + shares[worst] - 1u) / shares[worst];
    if (most <= 1u) break;
    const uint32_t G = (m[worst] + most - 2u) / (most - 1u);
    if (used + (G - shares[worst]) > (uint32_t)b->solve_budget) break;
    used += G - shares[worst];
    shares[worst] = G;
  }
  if (shares == b->counts[natray::LK_SOLVE]) return NATRAY_OK;
  std::vector<natray::WgMap> map;
  for (size_t k = 0; k < n; ++k)
    for (uint32_t g = 0; g < shares[k]; ++g) map.push_back({(uint32_t)k, g, shares[k], 0u});
  if (map.size() > b->map_cap[natray::LK_SOLVE]) return NATRAY_OK;  // keeps the current plan
  HIPCHK(hipMemcpyAsync(b->maps[natray::LK_SOLVE], map.data(), map.size() * sizeof(natray::WgMap), hipMemcpyHostToDevice,
                        b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->counts[natray::LK_SOLVE] = shares;
  b->tgt.grid[natray::LK_SOLVE] = (uint32_t)map.size();
  batch_drop_graphs(b);  // they hold the solve's grid
  return NATRAY_OK;
}

int batch_phase(natray_batch* b, int phase) {
  const bool timed = b->prof && !b->use_graph && b->ev_used < b->events.size() / 2;
  size_t slot = 0;
  if (timed) {
    slot = b->ev_used++;
    b->ev_phase[slot] = phase;
    HIPCHK(hipEventRecord(b->events[2 * slot], b->stream));
  }
  int e = batch_phase_launch(b, phase);
  if (e) return e;
  if (timed) HIPCHK(hipEventRecord(b->events[2 * slot + 1], b->stream));
  return NATRAY_OK;
}

// Every scene's tags for `steps` solves, its tagged words cleared first if a tag could repeat.
int batch_reserve_tags(natray_batch* b, uint32_t steps) {
  for (natray_sim* s : b->sims) {
    const uint64_t bound = solve_tag_bound(s) * steps;
    if (s->tags_used + bound >= kTagLimit) {
      HIPCHK(hipMemsetAsync(s->D.rwords, 0, natray::RED_WORDS * 8, b->stream));
      HIPCHK(hipMemsetAsync(s->D.hz, 0, (size_t)s->D.capc * 24, b->stream));
      HIPCHK(hipMemsetAsync(s->D.bar + 4, 0, 4, b->stream));
      s->tags_used = 0;
    }
    s->tags_used += bound;
  }
  return NATRAY_OK;
}

int batch_phase_launch(natray_batch* b, int phase) {
  const natray::Target& t = b->tgt;
  void* st = b->stream;
  switch (phase) {
    case NATRAY_PHASE_BEGIN: LAUNCHCHK(natray::launch_begin(t, st)); break;
    case NATRAY_PHASE_BIN: LAUNCHCHK(natray::launch_bin(t, st, false, b->air)); break;
    case NATRAY_PHASE_P2G:
      LAUNCHCHK(natray::launch_p2g(t, st));
      if (b->air) LAUNCHCHK(natray::launch_air_carry(t, st));
      break;
    case NATRAY_PHASE_GRID:
      LAUNCHCHK(natray::launch_grid_prep(t, st));
      if (b->air) LAUNCHCHK(natray::launch_air_fill(t, st));
      break;
    case NATRAY_PHASE_SOLVE: {
      if (!b->skip_tags) {
        int e = batch_reserve_tags(b, 1);
        if (e) return e;
      }
      LAUNCHCHK(natray::launch_solve(t, st, b->mixed));
      break;
    }
    case NATRAY_PHASE_PROJECT: LAUNCHCHK(natray::launch_project(t, st)); break;
    case NATRAY_PHASE_EXTEND: LAUNCHCHK(natray::launch_extend(t, st)); break;
    case NATRAY_PHASE_G2P: LAUNCHCHK(natray::launch_g2p(t, st)); break;
    default: return fail(NATRAY_E_ARG, "unknown phase " + std::to_string(phase));
  }
  return NATRAY_OK;
}

int batch_capture(natray_batch* b, hipGraphExec_t* out, uint32_t steps) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int e = NATRAY_OK;
  b->skip_tags = true;
  for (uint32_t r = 0; r < steps && e == NATRAY_OK; ++r)
    for (int ph = 0; ph < NATRAY_N_PHASES && e == NATRAY_OK; ++ph) e = batch_phase(b, ph);
  b->skip_tags = false;
  hipError_t ec = hipStreamEndCapture(b->stream, &g);
  if (e) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  if (ec != hipSuccess) return fail(NATRAY_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
  ec = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ec != hipSuccess) return fail(NATRAY_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
  return NATRAY_OK;
}

// Captures the step graph of the current shape if it is missing (recording runs nothing).
int batch_graphs(natray_batch* b) {
  if (!b->use_graph || b->graph_all) return NATRAY_OK;
  return batch_capture(b, &b->graph_all, 1);
}

int batch_step(natray_batch* b) {
  if (!b->use_graph) {
    for (int ph = 0; ph < NATRAY_N_PHASES; ++ph) {
      int e = batch_phase(b, ph);
      if (e) return e;
    }
    return NATRAY_OK;
  }
  int e = batch_graphs(b);
  if (e) return e;
  e = batch_reserve_tags(b, 1);
  if (e) return e;
  HIPCHK(hipGraphLaunch(b->graph_all, b->stream));
  return NATRAY_OK;
}

// n steps: with graphs, runs of kGraphSteps steps replay one graph of them all, the rest one
// step at a time.
int batch_steps(natray_batch* b, uint32_t n) {
  if (b->use_graph && n >= kGraphSteps) {
    if (!b->graph_many) {
      int e = batch_capture(b, &b->graph_many, kGraphSteps);
      if (e) return e;
    }
    for (; n >= kGraphSteps; n -= kGraphSteps) {
      int e = batch_reserve_tags(b, kGraphSteps);
      if (e) return e;
      HIPCHK(hipGraphLaunch(b->graph_many, b->stream));
    }
  }
  for (; n > 0; --n) {
    int e = batch_step(b);
    if (e) return e;
  }
  return NATRAY_OK;
}

// n_steps steps of every scene; the scenes' States are read back and checked afterwards.
// Events t0 and t1, when given, are recorded on the batch's stream around the steps.
int batch_run(natray_batch* b, uint32_t n_steps, uint32_t iters, hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr) {
  int rc = batch_plan_solve(b);
  if (rc) return rc;
  for (natray_sim* s : b->sims) {
    s->U.limits[2] = iters ? iters : s->U.limits[2];
    s->U.limits[3] = 0xffffffffu;
    *s->h_params = s->U;
    HIPCHK(hipMemcpyAsync(const_cast<natray_params*>(s->D.U), s->h_params, sizeof(natray_params), hipMemcpyHostToDevice,
                          b->stream));
    s->h_words[4] = 0;
    s->h_words[5] = 0;
    HIPCHK(hipMemcpyAsync(&s->D.st->iters_max, &s->h_words[4], 2 * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
  }
  if (t0) HIPCHK(hipEventRecord(t0, b->stream));
  int e = batch_steps(b, n_steps);
  if (e) return e;
  if (t1) HIPCHK(hipEventRecord(t1, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t k = 0; k < b->ev_used; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->events[2 * k], b->events[2 * k + 1]));
    b->ms[b->ev_phase[k]] += ms;
    b->cnt[b->ev_phase[k]] += 1;
  }
  b->ev_used = 0;
  for (size_t k = 0; k < b->sims.size(); ++k) {
    int e = sync_state(b->sims[k], "natray_batch_set_solve_launch(batch, ", b->tgt.plain_solve ? 1 : 0);
    if (e) return e;
    if (b->sims[k]->last.overflow)
      return fail(NATRAY_E_OVERFLOW, "scene " + std::to_string(k) + " needs more grid blocks than its capacity");
  }
  return NATRAY_OK;
}

}  // namespace

extern "C" {

int natray_batch_plan(const uint64_t* weights, const uint32_t* max_each, uint32_t n, uint32_t budget, uint32_t min_each,
                      uint32_t* out) {
  if (!weights || !out || n == 0) return fail(NATRAY_E_ARG, "null argument");
  std::vector<uint64_t> w(weights, weights + n);
  std::vector<uint32_t> hi(n, 0xffffffffu), shares;
  if (max_each) hi.assign(max_each, max_each + n);
  plan_shares(w, hi, budget, min_each, shares);
  for (uint32_t k = 0; k < n; ++k) out[k] = shares[k];
  return NATRAY_OK;
}

int natray_batch_create(natray_sim* const* sims, uint32_t n, natray_batch** out) {
  if (!sims || !out || n == 0) return fail(NATRAY_E_ARG, "a batch needs at least one scene");
  *out = nullptr;
  for (uint32_t k = 0; k < n; ++k) {
    if (!sims[k]) return fail(NATRAY_E_ARG, "null scene");
    if (sims[k]->mixed != sims[0]->mixed) return fail(NATRAY_E_ARG, "the scenes of a batch must share the CG precision");
    if (sims[k]->device != sims[0]->device) return fail(NATRAY_E_ARG, "the scenes of a batch must be on one device");
    if (sims[k]->ws_mode) return fail(NATRAY_E_ARG, "the warm start runs on single scenes, not in a batch");
    if (sims[k]->D.mem_cap) return fail(NATRAY_E_ARG, "the membrane coupling runs on single scenes, not in a batch");
    if (sims[k]->D.spray) return fail(NATRAY_E_ARG, "the spray runs on single scenes, not in a batch");
    if (sims[k]->D.n_bullets) return fail(NATRAY_E_ARG, "bullets run on single scenes, not in a batch");
    for (uint32_t j = 0; j < k; ++j)
      if (sims[j] == sims[k]) return fail(NATRAY_E_ARG, "a scene appears twice in the batch");
  }
  natray_batch* b = new natray_batch();
  b->sims.assign(sims, sims + n);
  b->num_cus = sims[0]->num_cus;
  b->mixed = sims[0]->mixed;
  for (uint32_t k = 0; k < n; ++k) b->air = b->air || sims[k]->D.air_x != nullptr;
  b->solve_budget = sims[0]->solve_grid;
  b->tgt.plain_solve = true;  // natray_batch_set_solve_launch
  b->shape = {8u, 8u, 4u, 0u};
  auto bail = [&](int code) {
    batch_free(b);
    delete b;
    return code;
  };
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) return bail(fail(NATRAY_E_HIP, std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e)));
  void* q = nullptr;
  e = hipMalloc(&q, (size_t)n * sizeof(natray::Dev));
  if (e != hipSuccess) return bail(fail(NATRAY_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)));
  b->allocs.push_back(q);
  b->scenes = static_cast<natray::Dev*>(q);
  // The scenes' own streams must be idle before the batch's stream works on their buffers.
  for (uint32_t k = 0; k < n; ++k) {
    e = hipStreamSynchronize(sims[k]->stream);
    if (e != hipSuccess) return bail(fail(NATRAY_E_HIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e)));
  }
  int rc = batch_layout(b);
  if (rc) return bail(rc);
  *out = b;
  return NATRAY_OK;
}

void natray_batch_destroy(natray_batch* b) {
  if (!b) return;
  batch_free(b);
  delete b;
}

int natray_batch_run(natray_batch* b, uint32_t n_steps, uint32_t iters) {
  if (!b) return fail(NATRAY_E_ARG, "null handle");
  for (natray_sim* s : b->sims) HIPCHK(hipStreamSynchronize(s->stream));
  return batch_run(b, n_steps, iters);
}

int natray_batch_set_solve_launch(natray_batch* b, int mode) {
  if (!b) return fail(NATRAY_E_ARG, "null handle");
  if (mode != NATRAY_SOLVE_PLAIN && mode != NATRAY_SOLVE_COOPERATIVE)
    return fail(NATRAY_E_ARG, "a batch's solve launch is NATRAY_SOLVE_PLAIN or NATRAY_SOLVE_COOPERATIVE");
  HIPCHK(hipStreamSynchronize(b->stream));
  b->tgt.plain_solve = mode == NATRAY_SOLVE_PLAIN;
  batch_drop_graphs(b);
  return NATRAY_OK;
}

int natray_batch_set_graph(natray_batch* b, int on) {
  if (!b) return fail(NATRAY_E_ARG, "null handle");
  b->use_graph = on != 0;
  return NATRAY_OK;
}

int natray_batch_set_profiling(natray_batch* b, int on) {
  if (!b) return fail(NATRAY_E_ARG, "null handle");
  HIPCHK(hipStreamSynchronize(b->stream));
  if (on && b->events.empty()) {
    b->events.resize(EVENT_POOL);
    b->ev_phase.resize(EVENT_POOL / 2);
    for (size_t k = 0; k < EVENT_POOL; ++k) HIPCHK(hipEventCreate(&b->events[k]));
  }
  b->prof = on != 0;
  b->ev_used = 0;
  for (int k = 0; k < NATRAY_N_PHASES; ++k) {
    b->ms[k] = 0.0;
    b->cnt[k] = 0;
  }
  return NATRAY_OK;
}

int natray_batch_phase_times(natray_batch* b, double* ms, uint32_t* counts) {
  if (!b) return fail(NATRAY_E_ARG, "null handle");
  for (int k = 0; k < NATRAY_N_PHASES; ++k) {
    if (ms) ms[k] = b->ms[k];
    if (counts) counts[k] = b->cnt[k];
  }
  return NATRAY_OK;
}

int natray_batch_set_shape(natray_batch* b, const natray_batch_shape* shape) {
  if (!b || !shape) return fail(NATRAY_E_ARG, "null argument");
  if (shape->blocks_per_cu < 1 || shape->blocks_per_cu > 64 || shape->waves_per_cu < 1 || shape->waves_per_cu > 64 ||
      shape->cells_per_cu < 1 || shape->cells_per_cu > 64)
    return fail(NATRAY_E_ARG, "workgroups per CU must be 1..64");
  HIPCHK(hipStreamSynchronize(b->stream));
  b->shape = *shape;
  return batch_layout(b);
}

int natray_batch_get_shape(natray_batch* b, natray_batch_shape* out) {
  if (!b || !out) return fail(NATRAY_E_ARG, "null argument");
  *out = b->shape;
  return NATRAY_OK;
}

int natray_batch_tune(natray_batch* b, uint32_t steps, uint32_t iters, natray_batch_shape* best, double* ms,
                      uint32_t max_candidates, uint32_t* n_candidates) {
  if (!b || steps == 0) return fail(NATRAY_E_ARG, "null handle or no steps");
  const uint32_t blocks[] = {4u, 8u, 16u}, waves[] = {4u, 8u, 16u};
  std::vector<natray_batch_shape> cand;
  for (uint32_t bl : blocks)
    for (uint32_t wv : waves) cand.push_back({bl, wv, 4u, 0u});
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // The scenes change from step to step (an impact's water spreads and its steps get dearer),
  // so the candidates are timed one step at a time in rounds that visit them forwards, then
  // backwards: over each pair of rounds a drift linear in the step count adds the same time to
  // every candidate. `steps` is rounded up to an even number of rounds. A candidate's graphs
  // are captured before its timed step, which is the only step it runs per round.
  const uint32_t rounds = (steps + 1u) & ~1u;
  std::vector<double> total(cand.size(), 0.0);
  int rc = NATRAY_OK;
  for (uint32_t r = 0; r < rounds && rc == NATRAY_OK; ++r) {
    for (size_t i = 0; i < cand.size() && rc == NATRAY_OK; ++i) {
      const size_t k = (r & 1u) ? cand.size() - 1 - i : i;
      rc = natray_batch_set_shape(b, &cand[k]);
      if (rc) break;
      rc = batch_graphs(b);
      if (rc) break;
      rc = batch_run(b, 1, iters, e0, e1);
      if (rc) break;
      float t = 0.0f;
      (void)hipEventElapsedTime(&t, e0, e1);
      total[k] += (double)t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  size_t best_k = 0;
  for (size_t k = 0; k < cand.size(); ++k) {
    if (ms && k < max_candidates) ms[k] = total[k] / rounds;
    if (total[k] < total[best_k]) best_k = k;
  }
  if (n_candidates) *n_candidates = (uint32_t)cand.size();
  if (best) *best = cand[best_k];
  return natray_batch_set_shape(b, &cand[best_k]);
}

}  // extern "C"
