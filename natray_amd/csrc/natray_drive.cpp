// natray_drive: steps a scene file through the C ABI of libnatray.so, from native code.
//
//   natray_drive <scene file> [steps=20] [warmup=2] [cg iteration cap=1000] [ws]
//
// With `ws` the solve is warm-started by a network of seeded random weights (mode 2 of
// natray_sim_set_warm_start): its kernels run as with trained weights, for timing them.
//
// The scene file comes from `python bench.py --write-scene <path>` (natray_amd.scenes
// write_scene_file): "NATRAYSC", u32 version 1, u32 particle count, then natray_params,
// natray_scene, natray_state and the 80-byte particle records, then optionally b"BULLETS\0", u32 n,
// n natray_bullet records and their patches (n x 576 x 8 floats), which the driver hands to
// natray_sim_set_bullets (`bench.py --bullets-only --write-scene`). The driver is what a profiler
// wraps (`rocprofv3 ... -- natray_amd/natray_drive <file>`): a native program, so the profiler
// sees the kernels without an interpreter in between. It prints one line with the timing.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "natray.h"

static bool read_all(FILE* f, void* dst, size_t bytes) { return fread(dst, 1, bytes, f) == bytes; }

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s <scene file> [steps] [warmup] [iters]\n", argv[0]);
    return 2;
  }
  const unsigned steps = argc > 2 ? (unsigned)atoi(argv[2]) : 20u;
  const unsigned warmup = argc > 3 ? (unsigned)atoi(argv[3]) : 2u;
  const unsigned iters = argc > 4 ? (unsigned)atoi(argv[4]) : 1000u;
  const bool warm = argc > 5 && strcmp(argv[5], "ws") == 0;
  FILE* f = fopen(argv[1], "rb");
  if (!f) {
    perror(argv[1]);
    return 2;
  }
  char magic[8];
  uint32_t version = 0, n = 0;
  natray_params params;
  natray_scene scene;
  natray_state state;
  if (!read_all(f, magic, 8) || memcmp(magic, "NATRAYSC", 8) != 0 || !read_all(f, &version, 4) || version != 1 ||
      !read_all(f, &n, 4) || !read_all(f, &params, sizeof params) || !read_all(f, &scene, sizeof scene) ||
      !read_all(f, &state, sizeof state)) {
    fprintf(stderr, "%s: not a version-1 scene file\n", argv[1]);
    fclose(f);
    return 2;
  }
  std::vector<float> particles((size_t)n * 20);
  if (!read_all(f, particles.data(), particles.size() * sizeof(float))) {
    fprintf(stderr, "%s: truncated particle records\n", argv[1]);
    fclose(f);
    return 2;
  }
  // The optional bullets trailer.
  uint32_t nb = 0;
  std::vector<natray_bullet> bullets;
  std::vector<float> patches;
  char tag[8];
  if (read_all(f, tag, 8)) {
    if (memcmp(tag, "BULLETS", 8) != 0 || !read_all(f, &nb, 4) || nb > NATRAY_MAX_BULLETS) {
      fprintf(stderr, "%s: bad bullets trailer\n", argv[1]);
      fclose(f);
      return 2;
    }
    bullets.resize(nb);
    patches.resize((size_t)nb * NATRAY_BULLET_PATCHES * 8);
    if (!read_all(f, bullets.data(), nb * sizeof(natray_bullet)) || !read_all(f, patches.data(), patches.size() * 4)) {
      fprintf(stderr, "%s: truncated bullets\n", argv[1]);
      fclose(f);
      return 2;
    }
  }
  fclose(f);
  natray_sim* sim = nullptr;
  int rc = natray_sim_create(&params, particles.data(), n, &scene, &state, &sim);
  if (rc) {
    fprintf(stderr, "natray_sim_create: %d %s\n", rc, natray_last_error());
    return 1;
  }
  if (nb && (rc = natray_sim_set_bullets(sim, nb, bullets.data(), patches.data())) != 0) {
    fprintf(stderr, "natray_sim_set_bullets: %d %s\n", rc, natray_last_error());
    natray_sim_destroy(sim);
    return 1;
  }
  natray_net* net = nullptr;
  if (warm) {
    if ((rc = natray_net_create(1024, 1, &net)) != 0 || (rc = natray_sim_set_warm_start(sim, net, 2)) != 0) {
      fprintf(stderr, "warm start: %d %s %s\n", rc, natray_last_error(), natray_net_last_error());
      natray_sim_destroy(sim);
      if (net) natray_net_destroy(net);
      return 1;
    }
  }
  natray_state st;
  if (warmup && (rc = natray_sim_run(sim, warmup, iters, 0xffffffffu, &st)) != 0) {
    fprintf(stderr, "warmup: %d %s\n", rc, natray_last_error());
    natray_sim_destroy(sim);
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  rc = natray_sim_run(sim, steps, iters, 0xffffffffu, &st);
  auto t1 = std::chrono::steady_clock::now();
  if (rc) {
    fprintf(stderr, "run: %d %s\n", rc, natray_last_error());
    natray_sim_destroy(sim);
    return 1;
  }
  const double s = std::chrono::duration<double>(t1 - t0).count();
  printf("natray_drive: %u particles, %u steps in %.3f ms (%.3f ms/step, %.4g particle-substeps/s), step %u, %u blocks, "
         "%u CG iterations in the last step, %u bullets\n",
         n, steps, s * 1e3, s * 1e3 / (steps ? steps : 1), steps ? (double)n * steps / s : 0.0, st.step, st.n_blocks,
         st.solve_iters, nb);
  fflush(stdout);
  natray_sim_destroy(sim);
  if (net) natray_net_destroy(net);
  return 0;
}
