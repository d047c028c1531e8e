// Host random streams of the reference.
//
// The reference draws from two generators (known-answer values measured on
// glibc 2.35 / GCC 11.4 pin both, tests/test_oracle_kat.py):
//  * glibc rand() after srand(rank*10000) (CellularAutomataManager_3D.cpp:40),
//    one global stream shared by site draws and nucleation angles;
//  * a fresh default-seeded std::default_random_engine feeding
//    std::normal_distribution<double> for each create_nucleation_sites call
//    (CellularAutomataManager.cpp:174-175).
// GlibcRand reproduces rand() with a private state (glibc's random_r, which is
// the generator rand() uses), so independent domains keep independent streams.
#pragma once
#include <stdint.h>
#include <stdlib.h>

#include <random>

namespace amca_host {

class GlibcRand {
 public:
  explicit GlibcRand(unsigned seed = 0) { reseed(seed); }
  void reseed(unsigned seed);
  int next();  // same sequence as srand(seed); rand(); rand(); ...
  static constexpr int kRandMax = 2147483647;  // RAND_MAX

 private:
  char state_[128];
  struct random_data data_;
};

// The normal-deviate stream of one create_nucleation_sites call.
class SiteNormal {
 public:
  SiteNormal(double mean, double sd) : dist_(mean, sd) {}
  double next() { return dist_(gen_); }

 private:
  std::default_random_engine gen_;
  std::normal_distribution<double> dist_;
};

}  // namespace amca_host
