#include "rng.h"

#include <cstring>
#include <stdexcept>

namespace amca_host {

void GlibcRand::reseed(unsigned seed) {
  std::memset(state_, 0, sizeof(state_));
  std::memset(&data_, 0, sizeof(data_));
  // a 128-byte state selects glibc's TYPE_3 additive generator, the one
  // behind rand()/srand()
  if (initstate_r(seed, state_, sizeof(state_), &data_) != 0)
    throw std::runtime_error("initstate_r failed");
}

int GlibcRand::next() {
  int32_t r = 0;
  random_r(&data_, &r);
  return (int)r;
}

}  // namespace amca_host
