// Integer settings read from the environment when a domain or batch is made.
#pragma once
#include <cstdlib>

namespace amca {

// 0: `name` is unset; 1: it holds an integer in [lo, hi], stored in *out;
// -1: it is set to anything else.
inline int env_int(const char* name, long lo, long hi, int* out) {
  const char* s = std::getenv(name);
  if (!s) return 0;
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  if (end == s || *end != '\0' || v < lo || v > hi) return -1;
  *out = (int)v;
  return 1;
}

}  // namespace amca
