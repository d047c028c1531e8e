#include "generators.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace amca_host {

namespace {

int pool_size(int threads) {
  if (threads > 0) return threads;
  return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

template <class F>
void parallel_for(int64_t n, int threads, F&& f) {
  const int nt = (int)std::min<int64_t>(pool_size(threads), std::max<int64_t>(n, 1));
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; t++)
    pool.emplace_back([&, t] {
      const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
      f(lo, hi);
    });
  for (auto& th : pool) th.join();
}

struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

constexpr double kPi = 3.14159265358979323846;

}  // namespace

void rosenthal_frames(int64_t n_nodes, const double* xyz, int n_frames, double t0, double dt,
                      const RosenthalParams& p, double* out, int threads) {
  const double amp = p.eta * p.P / (2.0 * kPi * p.k);
  parallel_for(n_nodes, threads, [&](int64_t lo, int64_t hi) {
    for (int f = 0; f < n_frames; f++) {
      const double xs = p.x_start + p.v * (t0 + dt * f);
      double* o = out + (int64_t)f * n_nodes;
      for (int64_t n = lo; n < hi; n++) {
        const double xi = xyz[3 * n] - xs;
        const double dy = xyz[3 * n + 1] - p.y_c;
        const double dz = xyz[3 * n + 2] - p.z_top;
        double R = std::sqrt(xi * xi + dy * dy + dz * dz);
        if (R < p.r_min) R = p.r_min;
        o[n] = p.T0 + amp / R * std::exp(-p.v * (R + xi) / (2.0 * p.alpha));
      }
    }
  });
}

void voronoi_labels(int nx, int ny, int nz, double h, double x0, double y0, double z0,
                    int n_seeds, uint64_t seed, int32_t* label, double* euler3, int threads) {
  SplitMix64 rng{seed};
  const double L[3] = {nx * h, ny * h, nz * h};
  const double O[3] = {x0, y0, z0};
  std::vector<double> pts(3 * (size_t)n_seeds);
  for (int s = 0; s < n_seeds; s++)
    for (int q = 0; q < 3; q++) pts[3 * s + q] = O[q] + rng.uniform() * L[q];
  for (int s = 0; s < n_seeds; s++) {
    const double U = rng.uniform(), V = rng.uniform(), W = rng.uniform();
    euler3[3 * s] = U * 2 * kPi;
    euler3[3 * s + 1] = std::acos(2 * V - 1);
    euler3[3 * s + 2] = W * 0.5 * kPi;
  }
  // seeds bucketed on a coarse grid; a cell searches rings of buckets until
  // the ring distance exceeds the best distance found
  const int nb = std::max(1, (int)std::cbrt((double)n_seeds / 2.0));
  int bn[3];
  double bs[3];
  for (int q = 0; q < 3; q++) {
    bn[q] = nb;
    bs[q] = L[q] / nb;
  }
  std::vector<std::vector<int>> bucket((size_t)bn[0] * bn[1] * bn[2]);
  auto bidx = [&](int q, double x) {
    int b = (int)std::floor((x - O[q]) / bs[q]);
    return std::min(std::max(b, 0), bn[q] - 1);
  };
  for (int s = 0; s < n_seeds; s++) {
    const int b = (bidx(2, pts[3 * s + 2]) * bn[1] + bidx(1, pts[3 * s + 1])) * bn[0] +
                  bidx(0, pts[3 * s]);
    bucket[b].push_back(s);
  }
  const int nxG = nx + 2, nyG = ny + 2, nzG = nz + 2;
  const int64_t NG = (int64_t)nxG * nyG * nzG;
  parallel_for(NG, threads, [&](int64_t lo, int64_t hi) {
    for (int64_t c = lo; c < hi; c++) {
      const int k = (int)(c / ((int64_t)nxG * nyG));
      const int j = (int)((c / nxG) % nyG);
      const int i = (int)(c % nxG);
      if (i == 0 || j == 0 || k == 0 || i == nxG - 1 || j == nyG - 1 || k == nzG - 1) {
        label[c] = -1;
        continue;
      }
      const double p[3] = {x0 + h * (i - 0.5), y0 + h * (j - 0.5), z0 + h * (k - 0.5)};
      int b[3];
      for (int q = 0; q < 3; q++) b[q] = bidx(q, p[q]);
      double best = 1e300;
      int arg = -1;
      for (int r = 0;; r++) {
        // ring r: buckets at Chebyshev distance r from the cell's bucket
        bool any = false;
        for (int bz = b[2] - r; bz <= b[2] + r; bz++) {
          if (bz < 0 || bz >= bn[2]) continue;
          for (int by = b[1] - r; by <= b[1] + r; by++) {
            if (by < 0 || by >= bn[1]) continue;
            for (int bx = b[0] - r; bx <= b[0] + r; bx++) {
              if (bx < 0 || bx >= bn[0]) continue;
              if (std::max(std::abs(bx - b[0]), std::max(std::abs(by - b[1]), std::abs(bz - b[2]))) != r)
                continue;
              any = true;
              for (int s : bucket[((size_t)bz * bn[1] + by) * bn[0] + bx]) {
                const double dx = p[0] - pts[3 * s], dy = p[1] - pts[3 * s + 1],
                             dz = p[2] - pts[3 * s + 2];
                const double d = dx * dx + dy * dy + dz * dz;
                if (d < best || (d == best && s < arg)) {
                  best = d;
                  arg = s;
                }
              }
            }
          }
        }
        // every bucket beyond ring r is at least r * (smallest bucket side) away
        const double reach = r * std::min(bs[0], std::min(bs[1], bs[2]));
        if ((arg >= 0 && reach * reach > best) || !any) break;
      }
      label[c] = arg;
    }
  });
}

}  // namespace amca_host
