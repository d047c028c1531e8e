// `amca3d -i <input deck> [-o <log file>]`: the command line of the reference's
// `cafe` driver (Cafe.cpp:39-138) over the GPU library.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>

#include "../../../include/amca3d.h"

int main(int argc, char** argv) {
  std::string input = "cafe.i", log;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "-h") || !std::strcmp(argv[i], "--help")) {
      std::fprintf(stderr,
                   "amca3d supported options:\n  -h [ --help ]        Help message\n"
                   "  -v [ --version ]     Code version\n"
                   "  -i [ --input-deck ]  Analysis input file (default cafe.i)\n"
                   "  -o [ --log-file ]    Analysis log file\n");
      return 0;
    }
    if (!std::strcmp(argv[i], "-v") || !std::strcmp(argv[i], "--version")) {
      std::fprintf(stderr, "Version: %s\n", amca_version());
      return 0;
    }
    if ((!std::strcmp(argv[i], "-i") || !std::strcmp(argv[i], "--input-deck")) && i + 1 < argc)
      input = argv[++i];
    else if ((!std::strcmp(argv[i], "-o") || !std::strcmp(argv[i], "--log-file")) && i + 1 < argc)
      log = argv[++i];
  }
  std::ifstream fin(input);
  if (!fin.good()) {
    std::fprintf(stderr,
                 "Input file is not specified or does not exist: user specified or (default) name= %s\n",
                 input.c_str());
    return 0;
  }
  amca_sim* sim = nullptr;
  if (amca_sim_create_from_file(input.c_str(), &sim) != AMCA_OK) {
    std::fprintf(stderr, "error: %s\n", amca_sim_last_error());
    return 1;
  }
  // the reference's VTR time series, as its solver writes it
  if (amca_sim_enable_vtr(sim, 1) != AMCA_OK) {
    std::fprintf(stderr, "error: %s\n", amca_sim_last_error());
    amca_sim_destroy(sim);
    return 1;
  }
  auto t0 = std::chrono::steady_clock::now();
  if (amca_sim_initialize(sim) != AMCA_OK) {
    std::fprintf(stderr, "error: %s\n", amca_sim_last_error());
    amca_sim_destroy(sim);
    return 1;
  }
  amca_counters c{};
  if (amca_sim_run(sim, -1, &c) != AMCA_OK) {
    std::fprintf(stderr, "error: %s\n", amca_sim_last_error());
    amca_sim_destroy(sim);
    return 1;
  }
  // output_microstructure_information of a coupled deck
  if (amca_sim_finalize(sim) != AMCA_OK) {
    std::fprintf(stderr, "error: %s\n", amca_sim_last_error());
    amca_sim_destroy(sim);
    return 1;
  }
  double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("steps %lld time %.9g remaining liquid cells %lld grains %d orientations %d wall %.3f s\n",
              (long long)c.step, c.time, (long long)c.n_liquid, c.n_grains, c.n_orient, sec);
  amca_sim_destroy(sim);
  return 0;
}
