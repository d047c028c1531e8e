// Host check of the environment-setting parser (amca3d_amd/csrc/env.h):
// prints "<code> <value>" for each name given on the command line.
#include <cstdio>

#include "../amca3d_amd/csrc/env.h"

int main(int argc, char** argv) {
  for (int i = 1; i < argc; i++) {
    int v = -12345;
    const int code = amca::env_int(argv[i], 0, 1024, &v);
    std::printf("%d %d\n", code, v);
  }
  return 0;
}
