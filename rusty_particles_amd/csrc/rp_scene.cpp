// rp_scene.cpp — seeded synthetic inputs for the benchmark workloads
// (host only; librp_scene.so).  Not part of the time step.
//
// rps_rsa: random sequential addition of non-overlapping spheres in a box:
// draw a radius uniformly in [rmin, rmax], then draw centres uniformly in the
// box (shrunk by the radius) until one overlaps no placed sphere; a uniform
// grid of cell 2*rmax makes each test O(1).  RNG: xoshiro256** seeded by
// splitmix64.
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstdint>
#include <algorithm>
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

// Places up to n spheres with the given radii (radii[i], i < n; spheres are
// placed in the order given) or, when radii is NULL, radii drawn uniformly in
// [rmin, rmax]; periodic[k] != 0 wraps axis k (overlaps tested with the
// minimum image).  Returns the number placed.  x3: n*3, r: n.
//
// Two uniform grids keep each test O(1) for wide radius distributions: spheres
// with r > rthr go to a coarse grid (cell >= 2 rmax), the others to a fine
// grid (cell >= 2 rthr); a candidate checks the cells of both grids that can
// hold an overlapping sphere.
namespace {
struct SGrid {
  int dims[3];
  double w[3];
  double lo[3];
  int per[3];
  std::vector<long long> head;
  void init(const double* lo_, const double* hi_, const int* per_, double cell) {
    for (int k = 0; k < 3; ++k) {
      lo[k] = lo_[k];
      per[k] = per_[k];
      double L = hi_[k] - lo_[k];
      dims[k] = (int)std::floor(L / cell);
      if (dims[k] < 1) dims[k] = 1;
      w[k] = L / dims[k];
    }
    head.assign((size_t)dims[0] * dims[1] * dims[2], -1);
  }
  int coord(double x, int k) const {
    int q = (int)std::floor((x - lo[k]) / w[k]);
    if (per[k]) return ((q % dims[k]) + dims[k]) % dims[k];
    return q < 0 ? 0 : (q >= dims[k] ? dims[k] - 1 : q);
  }
  size_t id(int a, int b, int c) const { return (size_t)a + (size_t)dims[0] * ((size_t)b + (size_t)dims[1] * c); }
};
}  // namespace

static int g_progress = 0;
// Progress lines on stderr (every tenth of the particles placed) from the
// placement loop itself, for callers that run long placements unattended.
void rps_set_progress(int on) { g_progress = on; }

// Region of the optional cone constraint of rsa_impl: centres with
// y0 <= y <= y1 and a distance from the cone's surface (the axis is +Y
// through (0, y0, 0), radius r0 at y0, growing by tan(half-angle) per metre)
// of at least `clear` times their radius.
struct Cone {
  double y0, y1, r0, tan_a, cos_a, clear;
  bool inside(const double* p, double ri) const {
    if (p[1] < y0 || p[1] > y1) return false;
    const double rho = std::sqrt(p[0] * p[0] + p[2] * p[2]);
    return (r0 + (p[1] - y0) * tan_a - rho) * cos_a >= clear * ri;
  }
};

static long long rsa_impl(long long n, const double* lo, const double* hi, double rmin, double rmax,
                          const double* radii, const int* periodic, unsigned long long seed, long long max_attempts,
                          double* x3, double* r, const Cone* cone);

long long rps_rsa2(long long n, const double* lo, const double* hi, double rmin, double rmax, const double* radii,
                   const int* periodic, unsigned long long seed, long long max_attempts, double* x3, double* r) {
  return rsa_impl(n, lo, hi, rmin, rmax, radii, periodic, seed, max_attempts, x3, r, nullptr);
}

// RSA inside a conical hopper: radii uniform in [rmin, rmax], centres in the
// box [lo, hi] and inside the cone (Cone above) with clearance clear * r from
// its surface (clear < 1: the sphere overlaps the wall by (1 - clear) r).
long long rps_rsa_cone(long long n, const double* lo, const double* hi, double rmin, double rmax, double y0,
                       double y1, double r0, double half_angle_deg, double clear, unsigned long long seed,
                       long long max_attempts, double* x3, double* r) {
  const double a = half_angle_deg * 3.14159265358979323846 / 180.0;
  const Cone c{y0, y1, r0, std::tan(a), std::cos(a), clear};
  return rsa_impl(n, lo, hi, rmin, rmax, nullptr, nullptr, seed, max_attempts, x3, r, &c);
}

}  // extern "C"

static long long rsa_impl(long long n, const double* lo, const double* hi, double rmin, double rmax,
                          const double* radii, const int* periodic, unsigned long long seed, long long max_attempts,
                          double* x3, double* r, const Cone* cone) {
  Rng rng(seed);
  const int per[3] = {periodic ? periodic[0] : 0, periodic ? periodic[1] : 0, periodic ? periodic[2] : 0};
  double L[3];
  for (int k = 0; k < 3; ++k) L[k] = hi[k] - lo[k];
  // threshold radius: the 99th percentile of the given radii (or rmax)
  double rthr = rmax;
  if (radii && n > 100) {
    std::vector<double> tmp(radii, radii + n);
    std::nth_element(tmp.begin(), tmp.begin() + (size_t)(0.99 * n), tmp.end());
    rthr = std::max(tmp[(size_t)(0.99 * n)], 1e-300);
  }
  SGrid big, small;
  big.init(lo, hi, per, 2.0 * rmax);
  small.init(lo, hi, per, 2.0 * rthr);
  std::vector<long long> next((size_t)n, -1);
  auto overlaps = [&](const SGrid& g, const double* p, double ri, double reach) {
    int c[3], span[3];
    for (int k = 0; k < 3; ++k) {
      c[k] = g.coord(p[k], k);
      span[k] = (int)std::ceil(reach / g.w[k]);
      if (per[k] && 2 * span[k] + 1 > g.dims[k]) span[k] = -1;  // visit every cell once
    }
    int a0[3], a1[3];
    for (int k = 0; k < 3; ++k) {
      if (span[k] < 0) { a0[k] = 0; a1[k] = g.dims[k] - 1; }
      else if (per[k]) { a0[k] = c[k] - span[k]; a1[k] = c[k] + span[k]; }
      else { a0[k] = std::max(0, c[k] - span[k]); a1[k] = std::min(g.dims[k] - 1, c[k] + span[k]); }
    }
    for (int qz = a0[2]; qz <= a1[2]; ++qz)
      for (int qy = a0[1]; qy <= a1[1]; ++qy)
        for (int qx = a0[0]; qx <= a1[0]; ++qx) {
          int q[3] = {qx, qy, qz};
          for (int k = 0; k < 3; ++k)
            if (per[k]) q[k] = ((q[k] % g.dims[k]) + g.dims[k]) % g.dims[k];
          for (long long j = g.head[g.id(q[0], q[1], q[2])]; j >= 0; j = next[j]) {
            double d[3] = {p[0] - x3[3 * j], p[1] - x3[3 * j + 1], p[2] - x3[3 * j + 2]};
            for (int k = 0; k < 3; ++k)
              if (per[k]) {
                if (d[k] > 0.5 * L[k]) d[k] -= L[k];
                else if (d[k] < -0.5 * L[k]) d[k] += L[k];
              }
            double rs = ri + r[j];
            if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rs * rs) return true;
          }
        }
    return false;
  };
  long long placed = 0;
  long long attempts = 0;
  const auto t_start = std::chrono::steady_clock::now();
  for (long long i = 0; i < n; ++i) {
    if (g_progress && n >= 100000 && i > 0 && i % (n / 10) == 0) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
      std::fprintf(stderr, "[scene] RSA placed %lld of %lld (%lld attempts, %.1f s)\n", i, n, attempts, el);
      std::fflush(stderr);
    }
    double ri = radii ? radii[i] : rmin + (rmax - rmin) * rng.uniform();
    bool ok = false;
    double p[3];
    for (long long att = 0; att < max_attempts && !ok; ++att) {
      ++attempts;
      for (int k = 0; k < 3; ++k)
        p[k] = per[k] ? lo[k] + L[k] * rng.uniform() : lo[k] + ri + (hi[k] - lo[k] - 2.0 * ri) * rng.uniform();
      ok = (!cone || cone->inside(p, ri)) && !overlaps(big, p, ri, ri + rmax) && !overlaps(small, p, ri, ri + rthr);
    }
    if (!ok) break;
    x3[3 * placed] = p[0];
    x3[3 * placed + 1] = p[1];
    x3[3 * placed + 2] = p[2];
    r[placed] = ri;
    SGrid& g = ri > rthr ? big : small;
    size_t id = g.id(g.coord(p[0], 0), g.coord(p[1], 1), g.coord(p[2], 2));
    next[placed] = g.head[id];
    g.head[id] = placed;
    ++placed;
  }
  return placed;
}


extern "C" long long rps_rsa(long long n, const double* lo, const double* hi, double rmin, double rmax,
                             unsigned long long seed, long long max_attempts, double* x3, double* r) {
  return rps_rsa2(n, lo, hi, rmin, rmax, nullptr, nullptr, seed, max_attempts, x3, r);
}
