// rp_surrogate.hpp — BF16-MFMA MLP surrogate of the JKR contact-radius solve.
//
// The reference solves f(a) = a^2/R* - sqrt(C a) - delta = 0, C = 8 pi g / E*,
// by Newton iteration for every adhesive contact (physics.rs:132-219).  With
// s = (R*^2 C)^(1/3) and t = s^2 / R* the equation is universal:
//   alpha^2 - sqrt(alpha) = d,   alpha = a / s,   d = delta / t,
// and the force is F / (pi g R*) = (32/3) alpha^3 - 8 alpha^(3/2).
// The network maps the contact's (delta, R*, E*, gamma) -- as
// u = sqrt(d - d_min) (alpha is close to linear in u) and the logs of R*, E*,
// gamma, each scaled to [-1, 1] over the training envelope -- to
// (alpha, F / (pi g R*)), also scaled.  Layers 4 -> 64 -> 64 -> 2, tanh
// hidden activations; BF16 operands, FP32 accumulation on MFMA.
//
// Which contacts the network serves.  The reference's Newton iteration
// starts at a0 = sqrt(R* delta) for delta > 0, i.e. alpha0 = sqrt(d), and at
// a0 = R*/2 otherwise.  For 0 < d < 0.1575 the start lies left of the minimum
// of alpha^2 - sqrt(alpha) (alpha = 4^(-2/3) = 0.397), the iteration walks
// towards a = 0 and ends on a near-zero or rejected radius; those contacts
// keep the exact solve (the gap 0 < d < kSurrGapHi).  Everywhere else in the
// envelope the iteration converges to the larger root, which is what the
// network learns and the polish (FP64 Newton steps from the prediction)
// reaches.  The envelope also bounds s / R* to [kSurrSLo, kSurrSHi]: below
// 1e-3 the reference's 20 iterations may end before converging, above 0.9 its
// start R*/2 lies left of the minimum.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

#include "../../include/rp_abi.h"

namespace rp {

constexpr int kMlpIn = 4, kMlpHid = 64, kMlpOut = 2;
// packed FP32 parameter vector: W1[4x64] b1[64] W2[64x64] b2[64] W3[64x2] b3[2]
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kOffB3 + kMlpOut;  // 4610

constexpr double kJkrDMinS = -0.47247039371057744;  // min of alpha^2 - sqrt(alpha)
constexpr double kSurrDLoMin = -0.40;  // lowest d served (the slope 2 alpha - 1/(2 sqrt(alpha)) >= 0.6)
constexpr double kSurrGapHi = 0.20;    // 0 < d < kSurrGapHi keeps the exact solve
constexpr double kSurrSLo = 1e-3, kSurrSHi = 0.9;  // s / R* range served
// Feature rows are claimed by workgroups from kSurrRanges separate ranges
// (workgroup b from range b % kSurrRanges), so the claiming atomics spread
// over as many addresses.
constexpr int kSurrRanges = 16;
// The range counters sit 256 bytes apart (different cache lines and memory
// channels), so claims on different ranges do not contend.
constexpr int kSurrCounterStride = 64;

// Scaling of inputs and outputs (the training envelope).
struct MlpScale {
  float in_lo[4], in_hi[4];    // u = sqrt(d - d_min), ln R*, ln E*, ln gamma
  float out_lo[2], out_hi[2];  // alpha, F / (pi g R*)
};

// Device copy of the network: weights as BF16 (padded: W1 16 rows, W3 16
// columns, zero filled), biases FP32.
struct MlpDev {
  const uint16_t* w1;  // [16 x 64]
  const uint16_t* w2;  // [64 x 64]
  const uint16_t* w3;  // [64 x 16]
  const float* b1;     // [64]
  const float* b2;     // [64]
  const float* b3;     // [16]
  MlpScale sc;
};

__host__ __device__ inline float scale_in(float v, float lo, float hi) { return 2.0f * (v - lo) / (hi - lo) - 1.0f; }
__host__ __device__ inline float unscale_out(float v, float lo, float hi) { return lo + 0.5f * (v + 1.0f) * (hi - lo); }

// Engine-side surrogate state passed to the collect and force kernels.
struct SurrDev {
  double d_lo, d_hi;  // served d range (the gap (0, kSurrGapHi) excluded)
  double r_lo, r_hi;  // served R* range
  int pp_ok;          // the particle pair's E* and gamma lie in the envelope
  int check;          // also run the exact solve and record the force errors
  MlpScale sc;
  float f_lne, f_lng; // scaled ln E*, ln gamma of particle-particle pairs
  const float4* inst; // batched engines: per instance (served 1/0, scaled ln E*, scaled ln gamma, -)
  int* base;          // [slots] first feature row of each particle, -1 = none
  unsigned long long* rowmask;  // [slots] neighbour-list entries that have a row (in row order)
  unsigned long long* nmask;    // [slots] entries near the served set: the next collect visits only these
  double md_lo, md_hi;  // d range of nmask (the served range widened by a margin)
  float* x;           // [rows_cap x 4] scaled features
  float* y;           // [rows_cap x 2] scaled predictions
  unsigned* counter;  // rows claimed in this step: range r at counter[r * kSurrCounterStride]
  int rows_cap;       // rows in all; range r holds rows [r R, r R + R), R = rows_cap / kSurrRanges
};

// Per-workgroup surrogate counters of one force launch.
struct SurrStats {
  unsigned served;    // contacts whose radius came from the network + polish
  unsigned fallback;  // routed contacts that used the exact solve (no row, or the polish check failed)
  float err_max;      // check mode: max |F - F_exact| / max(|F_exact|, (2/3) pi gamma R*)
  float raw_max;      // the same for the network's own force output
  double err_sum, raw_sum;
};

// True when the network serves a contact: d = delta / t in the served range
// and s / R* inside [kSurrSLo, kSurrSHi]; `k` is the contact's position in
// the particle's neighbour list (a 64-bit mask records the served ones), `ok`
// says whether the pair's E* and gamma lie in the envelope.
__host__ __device__ inline bool surr_route(double ov, double rstar, double s, double t, const SurrDev& sd, int k,
                                           double& d, int ok) {
  if (k >= 64 || !ok || !(rstar >= sd.r_lo && rstar <= sd.r_hi)) return false;
  if (!(s >= kSurrSLo * rstar && s <= kSurrSHi * rstar)) return false;
  d = ov / t;
  return (d >= sd.d_lo && d <= 0.0) || (d >= kSurrGapHi && d <= sd.d_hi);
}

// FP64 Newton steps on alpha^2 - sqrt(alpha) = d from the network's alpha,
// kept right of the minimum at 4^(-2/3) (convexity then brings every start
// to the larger root).  False when the residual is not at rounding level.
__host__ __device__ inline bool polish_alpha(double d, double& al) {
  al = fmin(fmax(al, 0.45), 1.0e4);
  for (int it = 0; it < 4; ++it) {
    const double sq = sqrt(al);
    al -= (al * al - sq - d) / (2.0 * al - 0.5 / sq);
  }
  const double sq = sqrt(al);
  return al > 0.45 && fabs(al * al - sq - d) <= 1e-10 * (1.0 + fabs(d));
}

// Validates an envelope (RP_OK or an error message in *why).
int envelope_check(const rp_jkr_envelope& e, const char** why);

// Input/output scaling for a training envelope (alpha and F/(pi g R*) ranges
// follow from the d range).
MlpScale make_scale(const rp_jkr_envelope& e);

// Owns device buffers for one network (host side).
struct MlpHost {
  uint16_t* w = nullptr;  // 16*64 + 64*64 + 64*16 bf16
  float* b = nullptr;     // 64 + 64 + 16
  MlpDev dev() const;
};

// Uploads a packed FP32 parameter vector (kMlpParams floats) into m.
int mlp_upload(MlpHost& m, const float* params, hipStream_t stream);
void mlp_free(MlpHost& m);

// Runs the network on `rows` feature rows (FP32, 4 per row, already
// scaled) and writes 2 scaled outputs per row.  x and y are device pointers.
void mlp_infer(const MlpDev& net, const float* x, float* y, int rows, hipStream_t stream);

// The same on the claimed rows of each of the kSurrRanges ranges of
// [0, cap): counter[r * kSurrCounterStride] rows from r cap / kSurrRanges
// (counts read on the device).  With an event pair, the dispatch records
// its start and end in them (hipExtLaunchKernel).
void mlp_infer_counted(const MlpDev& net, const float* x, float* y, const unsigned* counter, int cap,
                       hipStream_t stream, hipEvent_t start = nullptr, hipEvent_t stop = nullptr);

}  // namespace rp
