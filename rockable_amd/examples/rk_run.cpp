// rk_run -- the C++ caller of the engine that INTEGRATION.md shows: the
// integrate() loop of Rockable's console run (src/Apps/run.cpp ->
// Rockable::console_run, Rockable.cpp:1232-1269) on the GPU, through the C
// ABI of include/rockable_amd.h alone.
//
//   rk_run <conf file> <output dir> [max steps (-1: until tmax)] [driving file]
//
// It loads the deck (its accelerations() "fake step" included), reads the
// driving file if one is named, builds the neighbour list and integrates,
// writing conf<i> files into the output directory as Rockable does.  Exit
// status 0 on success, 1 with the engine's message on error.
#include <cstdio>
#include <cstdlib>

#include "rockable_amd.h"

static int run_on_gpu(const char* confFile, const char* outDir, int64_t maxSteps, const char* driving,
                      int64_t* done) {
  rk_engine* eng = nullptr;
  const int rc0 = rk_engine_create(&eng, /*device*/ 0);  // ~ Rockable()
  if (rc0 != 0) {
    fprintf(stderr, "rk_run: no usable device (%d)\n", rc0);
    return -1;
  }
  int rc = rk_load_conf(eng, confFile);                           // ~ loadConf + fake step
  if (rc == 0 && driving) rc = rk_load_driving_system(eng, driving);  // ~ System.read(true)
  if (rc == 0) rc = rk_update_nl(eng);                            // ~ UpdateNL() before integrate
  if (rc == 0) rc = rk_integrate(eng, outDir, maxSteps, done);    // ~ integrate(): conf files in outDir
  if (rc < 0) fprintf(stderr, "rk_run: %s\n", rk_last_error(eng));
  rk_engine_destroy(eng);
  return rc;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <conf file> <output dir> [max steps] [driving file]\n", argv[0]);
    return 2;
  }
  const int64_t maxSteps = argc > 3 ? strtoll(argv[3], nullptr, 10) : -1;
  const char* driving = argc > 4 ? argv[4] : nullptr;
  int64_t done = 0;
  if (run_on_gpu(argv[1], argv[2], maxSteps, driving, &done) < 0) return 1;
  printf("rk_run: %lld steps\n", (long long)done);
  return 0;
}
