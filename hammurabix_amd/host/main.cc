// hamx <params.xml>: the reference's command-line driver (main_std.cc:10-44)
// running the LOS integrator and random-field generators on the GPU.
#include <chrono>
#include <cstdio>
#include <exception>

#include "hamx_host.h"

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s <parameter file>\n", argv[0]);
    return 1;
  }
  try {
    const auto t0 = std::chrono::steady_clock::now();
    hamx_host::Pipeline run(argv[1]);
    run.assemble_grid();
    run.assemble_tereg();
    run.assemble_breg();
    run.assemble_ternd();
    run.assemble_brnd();
    run.assemble_cre();
    run.assemble_obs();
    const double s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("hamx: done in %.3f s\n", s);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "hamx: %s\n", e.what());
    return 1;
  }
  return 0;
}
