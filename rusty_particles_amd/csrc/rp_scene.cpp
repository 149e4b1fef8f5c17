// rp_scene.cpp — seeded synthetic inputs for the benchmark workloads
// (host only; librp_scene.so).  Not part of the time step.
//
// rps_rsa: random sequential addition of non-overlapping spheres in a box:
// draw a radius uniformly in [rmin, rmax], then draw centres uniformly in the
// box (shrunk by the radius) until one overlaps no placed sphere; a uniform
// grid of cell 2*rmax makes each test O(1).  RNG: xoshiro256** seeded by
// splitmix64.
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int k = 0; k < 4; ++k) {
      seed += 0x9E3779B97F4A7C15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[k] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace

extern "C" {

// Places up to n spheres; returns the number placed.  x3: n*3, r: n.
long long rps_rsa(long long n, const double* lo, const double* hi, double rmin, double rmax, unsigned long long seed,
                  long long max_attempts, double* x3, double* r) {
  Rng rng(seed);
  double cs = 2.0 * rmax;
  int dims[3];
  for (int k = 0; k < 3; ++k) {
    dims[k] = (int)std::ceil((hi[k] - lo[k]) / cs);
    if (dims[k] < 1) dims[k] = 1;
  }
  size_t ncell = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<long long> head(ncell, -1), next((size_t)n, -1);
  auto cell_of = [&](const double* p, int c[3]) {
    for (int k = 0; k < 3; ++k) {
      int q = (int)std::floor((p[k] - lo[k]) / cs);
      c[k] = q < 0 ? 0 : (q >= dims[k] ? dims[k] - 1 : q);
    }
  };
  long long placed = 0;
  for (long long i = 0; i < n; ++i) {
    double ri = rmin + (rmax - rmin) * rng.uniform();
    bool ok = false;
    double p[3];
    for (long long att = 0; att < max_attempts && !ok; ++att) {
      for (int k = 0; k < 3; ++k) p[k] = lo[k] + ri + (hi[k] - lo[k] - 2.0 * ri) * rng.uniform();
      int c[3];
      cell_of(p, c);
      ok = true;
      for (int dz = -1; dz <= 1 && ok; ++dz)
        for (int dy = -1; dy <= 1 && ok; ++dy)
          for (int dx = -1; dx <= 1 && ok; ++dx) {
            int q[3] = {c[0] + dx, c[1] + dy, c[2] + dz};
            if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= dims[0] || q[1] >= dims[1] || q[2] >= dims[2]) continue;
            size_t id = (size_t)q[0] + (size_t)dims[0] * ((size_t)q[1] + (size_t)dims[1] * q[2]);
            for (long long j = head[id]; j >= 0; j = next[j]) {
              double d0 = p[0] - x3[3 * j], d1 = p[1] - x3[3 * j + 1], d2 = p[2] - x3[3 * j + 2];
              double rs = ri + r[j];
              if (d0 * d0 + d1 * d1 + d2 * d2 < rs * rs) {
                ok = false;
                break;
              }
            }
          }
    }
    if (!ok) break;
    x3[3 * placed] = p[0];
    x3[3 * placed + 1] = p[1];
    x3[3 * placed + 2] = p[2];
    r[placed] = ri;
    int c[3];
    cell_of(p, c);
    size_t id = (size_t)c[0] + (size_t)dims[0] * ((size_t)c[1] + (size_t)dims[1] * c[2]);
    next[placed] = head[id];
    head[id] = placed;
    ++placed;
  }
  return placed;
}

}  // extern "C"
