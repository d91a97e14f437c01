/* CPU oracle of the natray water substep — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's WGSL kernels (shaders/fluid.wgsl,
 * shaders/sim_defs.wgsl) used by the tests, by __graft_entry__.smoke() and by bench.py's
 * CPU-baseline leg, as the checker and never as the product. It is compiled twice from one
 * source: with REAL = float ("o32_", FP32-faithful: same operation order, no contraction,
 * ties-to-even rounding) and with REAL = double ("o64_", the FP64 oracle of the error
 * metrics). Both share the integer fixed-point transfer sums.
 *
 * Which parts are pinned, and by what, is listed at the top of natray_oracle.c.
 */
#ifndef NATRAY_ORACLE_H
#define NATRAY_ORACLE_H

#include "natray.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct osim osim;

#define ORACLE_DECL(P)                                                                               \
  osim* P##create(const natray_params* params, const float* particles, uint32_t n,                  \
                  const natray_scene* scene, const natray_state* init);                             \
  void P##destroy(osim* s);                                                                         \
  int P##phase(osim* s, int phase);                                                                 \
  int P##run(osim* s, uint32_t n_steps, uint32_t iters, uint32_t stop_ns);                          \
  void P##read_state(const osim* s, natray_state* out);                                             \
  void P##write_state(osim* s, const natray_state* in);                                             \
  void P##read_particles(const osim* s, float* dst);                                                \
  void P##read_particles_f64(const osim* s, double* dst);                                           \
  void P##write_particles(osim* s, const float* src);                                               \
  uint32_t P##n_blocks(const osim* s);                                                              \
  void P##read_grid(const osim* s, uint32_t* block_list, uint32_t* block_nb, uint32_t* block_start, \
                    int32_t* acc, float* cf);                                                       \
  void P##read_cf_f64(const osim* s, double* cf);                                                   \
  void P##write_grid(osim* s, const int32_t* acc, const float* cf);                                 \
  void P##read_bins(const osim* s, uint32_t* block_ptr, uint32_t* sorted);                          \
  void P##solver_scalars(const osim* s, double out[8]);                                             \
  void P##precondition(const float* r, const float* diag, const int32_t* water, float* z, uint32_t mode); \
  void P##set_guess(osim* s, const float* guess);                                                   \
  void P##residual(osim* s, const float* p, float* r);                                              \
  int P##set_membrane(osim* s, uint32_t n, const float* x, const float* v, const float* mass,        \
                      const float* push);                                                           \
  void P##read_membrane(const osim* s, float* water);                                               \
  void P##pack_water(const osim* s, uint32_t* out);                                                 \
  void P##read_air(const osim* s, float* out);                                                      \
  uint32_t P##read_spray(const osim* s, uint32_t max_records, double* spray, double* landed);       \
  void P##read_wet(const osim* s, uint32_t* touch);                                                 \
  uint32_t P##pack_keyframe(const osim* s, uint32_t n_tris, const uint32_t* alive, uint32_t* words,   \
                            uint32_t max_words, float* spray_w);                                    \
  int P##set_bullets(osim* s, uint32_t n, const natray_bullet* bullets, const float* patches);      \
  void P##read_bullets(const osim* s, natray_bullet* bullets, double* sections);                    \
  void P##read_pose_log(const osim* s, float* log);                                                 \
  void P##elementary(int which, const float* a, const float* b, float* out, uint32_t n);           \
  void P##set_threads(int n);

ORACLE_DECL(o32_)
ORACLE_DECL(o64_)

#ifdef __cplusplus
}
#endif

#endif
