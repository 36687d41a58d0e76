// Host check of clexmc::accept_prob (the kernels' acceptance test).  For
// probabilities p from 1e-37 to 1 it finds, from the function's own answers,
// how many first words x accept outright and, for the one word that defers
// to a second word y, how many y accept; it prints p and the implied
// acceptance probability (n_x + n_y / 2^32) / 2^32, and counts words above
// the deferring one that accept (there must be none).  Built for the host.
#include <cmath>
#include <cstdio>

#include "../../casmcode_clexmonte_amd/csrc/clexmc_device.h"

int main() {
  const unsigned long long W = 4294967296ull;
  int bad_shape = 0;
  for (double lp = -37.0; lp <= 0.0; lp += 0.25) {
    const float p = (float)std::pow(10.0, lp);
    auto acc = [&](unsigned long long x, unsigned long long y) {
      return clexmc::accept_prob(p, (uint32_t)x, [&]() { return (uint32_t)y; });
    };
    // acceptance is monotone in x and in y (small words accept): bisect for
    // the first word x that rejects even with the smallest second word
    unsigned long long lo = 0, hi = W;
    while (lo < hi) {
      const unsigned long long mid = (lo + hi) / 2;
      if (acc(mid, 0)) lo = mid + 1; else hi = mid;
    }
    unsigned long long nx = lo, ny = 0;
    // the last accepting word defers to y if the largest y rejects there
    if (nx > 0 && !acc(nx - 1, W - 1)) {
      --nx;
      unsigned long long a = 0, b = W;
      while (a < b) {
        const unsigned long long m = (a + b) / 2;
        if (acc(nx, m)) a = m + 1; else b = m;
      }
      ny = a;
    }
    for (unsigned long long x = nx + 1; x < nx + 64 && x < W; ++x) bad_shape += acc(x, 0) ? 1 : 0;
    const double implied = ((double)nx + (double)ny / (double)W) / (double)W;
    std::printf("%.9e %.9e\n", (double)p, implied);
  }
  std::printf("shape_errors %d\n", bad_shape);
  return 0;
}
