// rp_surrogate.hip — BF16 MFMA training and inference of the JKR surrogate
// (see rp_surrogate.hpp for the model).
//
// Matrix products: one wave computes a 16x16 FP32 tile of op(A) * op(B)
// from BF16 operands.  v_mfma_f32_16x16x16_bf16 (training, and inference's
// first layer, whose 4 inputs are padded to K = 16): per lane l (64 lanes), A
// supplies A[l % 16][k0 + 4 (l / 16) + q], B supplies B[k0 + 4 (l / 16) + q][l % 16],
// q = 0..3.  gfx950's v_mfma_f32_16x16x32_bf16 (inference's 64-wide layers)
// takes K = 32 per instruction: A[l % 16][k0 + 8 (l / 16) + q] and
// B[k0 + 8 (l / 16) + q][l % 16], q = 0..7.  Both accumulate
// D[4 (l / 16) + q][l % 16], q = 0..3.  rp_mfma_probe exposes one 16x16x16
// tile so a test can check that layout against a host product; the 16x16x32
// layout is checked by the inference tests against a host forward pass.
//
// Training (all on the GPU): samples (d, R*, E*, gamma) drawn over the served
// envelope (sqrt(d - d_min) uniform, the others log-uniform) are labelled by the FP64
// Newton solve of the reference (rp::jkr, physics.rs:132-219); mini-batch
// Adam on the mean squared error of the scaled outputs; forward and backward
// products on MFMA with FP32 accumulation, FP32 master weights.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/rp_abi.h"
#include "rp_physics.hpp"
#include "rp_surrogate.hpp"

namespace rp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even
__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

namespace {

thread_local std::string s_err;
rp_status sfail(rp_status c, const std::string& m) {
  s_err = m;
  return c;
}
#define SHIP(expr)                                                                                   \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess) return sfail(RP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr int kWaveTiles = 4;  // 16x16 tiles per 256-thread workgroup

// C[M x N] = op(A)[M x K] * op(B)[K x N] (+ C if accumulate) (+ bias[col])
// (tanh if act), FP32 out and optionally a BF16 copy.  M, N, K multiples of 16.
// TA: A is stored K x M (row-major) and used transposed; TB likewise for B.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) k_gemm(const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, float* C, int ldc,
                                              uint16_t* Cb, int M, int N, int K, const float* bias, int act) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int tiles_n = N / 16;
  const int tm = wave / tiles_n, tn = wave % tiles_n;
  if (tm * 16 >= M) return;  // wave-uniform
  const int r0 = tm * 16, c0 = tn * 16;
  const int li = lane & 15, kq = 4 * (lane >> 4);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = k0 + kq + q;
      const uint16_t av = TA ? A[(size_t)k * lda + (r0 + li)] : A[(size_t)(r0 + li) * lda + k];
      const uint16_t bv = TB ? B[(size_t)(c0 + li) * ldb + k] : B[(size_t)k * ldb + (c0 + li)];
      a[q] = (short)av;
      b[q] = (short)bv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int rr = r0 + kq + q, cc = c0 + li;
    float v = acc[q];
    if (bias) v += bias[cc];
    if (act) v = tanhf(v);
    C[(size_t)rr * ldc + cc] = v;
    if (Cb) Cb[(size_t)rr * ldc + cc] = f2bf(v);
  }
}

template <bool TA, bool TB>
void gemm(hipStream_t s, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C, int ldc, uint16_t* Cb,
          int M, int N, int K, const float* bias, int act) {
  const int tiles = (M / 16) * (N / 16);
  const int blocks = (tiles + kWaveTiles - 1) / kWaveTiles;
  hipLaunchKernelGGL((k_gemm<TA, TB>), dim3(blocks), dim3(256), 0, s, A, lda, B, ldb, C, ldc, Cb, M, N, K, bias,
                     act);
}

// ---- fused inference: 16 rows per wave ---------------------------------------
// tanh x = 1 - 2 / (e^(2x) + 1) on the transcendental unit (v_exp_f32,
// v_rcp_f32): absolute error below 1e-6, far under the BF16 rounding the
// next layer applies to it.
__device__ inline float tanh_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(fminf(x * 2.8853900817779268f, 64.0f));  // 2^(2x log2 e)
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

// One lane's B operands (weights in fragment order, see mlp_upload) and
// biases, loaded once per workgroup and reused by every tile it runs.
struct MlpLane {
  bf16x4 w1[4];          // layer 1, K = 16: one 16x16x16 tile per output tile
  bf16x8 w2[8], w3[2];   // layers 2 and 3, K = 64: two 16x16x32 tiles along K
  float b1[4], b2[4], b3;
};

__device__ inline void mlp_lane_load(const MlpDev& net, MlpLane& L) {
  const int lane = threadIdx.x & 63, li = lane & 15;
  const uint2* w1 = reinterpret_cast<const uint2*>(net.w1);
  const uint4* w2 = reinterpret_cast<const uint4*>(net.w2);
  const uint4* w3 = reinterpret_cast<const uint4*>(net.w3);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    L.w1[t] = __builtin_bit_cast(bf16x4, w1[t * 64 + lane]);
    L.b1[t] = net.b1[t * 16 + li];
    L.b2[t] = net.b2[t * 16 + li];
  }
#pragma unroll
  for (int t = 0; t < 8; ++t) L.w2[t] = __builtin_bit_cast(bf16x8, w2[t * 64 + lane]);
#pragma unroll
  for (int t = 0; t < 2; ++t) L.w3[t] = __builtin_bit_cast(bf16x8, w3[t * 64 + lane]);
  L.b3 = net.b3[li];
}

// Hidden activations of one wave's 16 rows, BF16 (the precision the next
// layer's operands have anyway), 8 elements of padding per row so that the
// 16-byte operand reads of the 16 rows spread over the LDS banks.
constexpr int kHidStride = kMlpHid + 8;

// Layer outputs go through LDS to turn the accumulator layout (row = 4(l/16)+q)
// into the A-operand layout (row = l % 16) of the next layer; each value is
// rounded to BF16 once, when it is stored (v_cvt_pk_bf16_f32, round to
// nearest even).
// Rows [16 tile, 16 tile + 16) of one wave; every wave of the workgroup
// calls it together (it contains workgroup barriers).  Waves past the last
// row compute on zeros and store nothing.
__device__ inline void mlp_tile(const MlpLane& L, const float* __restrict__ x, float* __restrict__ y, int rows,
                                int tile, __bf16 (*h)[kHidStride]) {
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, kq = 4 * (lane >> 4), k8 = 8 * (lane >> 4);
  const int r0 = tile * 16;
  // layer 1: [16 x 16(4)] x W1 [16 x 64]
  bf16x4 a;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = kq + q, r = r0 + li;
    const float v = (k < kMlpIn && r < rows) ? x[(size_t)r * kMlpIn + k] : 0.0f;
    a[q] = __builtin_bit_cast(short, (__bf16)v);
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, L.w1[nt], acc, 0, 0, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) h[kq + q][nt * 16 + li] = (__bf16)tanh_fast(acc[q] + L.b1[nt]);
  }
  __syncthreads();
  // layer 2: [16 x 64] x W2 [64 x 64], two K = 32 steps
  f32x4 acc2[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) acc2[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < 2; ++kt) {
    const bf16x8 a2 = *reinterpret_cast<const bf16x8*>(&h[li][kt * 32 + k8]);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
      acc2[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, L.w2[kt * 4 + nt], acc2[nt], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int q = 0; q < 4; ++q) h[kq + q][nt * 16 + li] = (__bf16)tanh_fast(acc2[nt][q] + L.b2[nt]);
  __syncthreads();
  // layer 3: [16 x 64] x W3 [64 x 16(2)], two K = 32 steps
  f32x4 acc3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kt = 0; kt < 2; ++kt) {
    const bf16x8 a3 = *reinterpret_cast<const bf16x8*>(&h[li][kt * 32 + k8]);
    acc3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, L.w3[kt], acc3, 0, 0, 0);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int r = r0 + kq + q;
    if (li < kMlpOut && r < rows) y[(size_t)r * kMlpOut + li] = acc3[q] + L.b3;
  }
  __syncthreads();  // h is reused by the next tile
}

__global__ void __launch_bounds__(256) k_mlp_infer(MlpDev net, const float* __restrict__ x,
                                                   float* __restrict__ y, int rows) {
  __shared__ __align__(16) __bf16 hs[4][16][kHidStride];
  MlpLane L;
  mlp_lane_load(net, L);
  mlp_tile(L, x, y, rows, blockIdx.x * 4 + (threadIdx.x >> 6), hs[threadIdx.x >> 6]);
}

// Row counts read on the device: range r of [0, cap) holds
// counter[r * kSurrCounterStride] rows.
// A fixed grid walks 64-row groups (4 tiles, one per wave); a group lies in
// one range (ranges are multiples of 64 rows), so the skip test is the same
// for every wave of a workgroup and all meet the barriers.
__global__ void __launch_bounds__(256) k_mlp_counted(MlpDev net, const float* __restrict__ x,
                                                     float* __restrict__ y, const unsigned* counter, int cap) {
  __shared__ __align__(16) __bf16 hs[4][16][kHidStride];
  MlpLane L;
  mlp_lane_load(net, L);
  const int range_rows = cap / kSurrRanges;
  for (int g = blockIdx.x; g * 64 < cap; g += gridDim.x) {
    const int r = (g * 64) / range_rows;
    const int first = r * range_rows;
    const int rows = (int)min(counter[r * kSurrCounterStride], (unsigned)range_rows);
    if (g * 64 - first >= rows) continue;
    mlp_tile(L, x + (size_t)first * kMlpIn, y + (size_t)first * kMlpOut, rows, (g * 64 - first) / 16 + (threadIdx.x >> 6),
             hs[threadIdx.x >> 6]);
  }
}

// ---- training kernels ----------------------------------------------------------

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double u01(uint64_t seed, uint64_t i, int k) {
  return (double)(mix64(seed + 0x9E3779B97F4A7C15ull * (4 * i + k + 1)) >> 11) * (1.0 / 9007199254740992.0);
}

struct EnvDev {
  double d_lo, d_hi, lr_lo, lr_hi, le_lo, le_hi, lg_lo, lg_hi;
};

// One labelled sample: scaled features (4) and scaled targets (2); raw[8] =
// (delta, R*, E*, gamma, a, F_unclipped, d, ok).  u = sqrt(d - d_min) is
// drawn uniformly over the served set (the gap (0, kSurrGapHi) skipped);
// R*, E*, gamma log-uniformly, redrawn (up to 64 times) until s / R* lies in
// the served range.  The label is the reference's own Newton solve
// (rp::jkr, physics.rs:132-219) in FP64; ok = 1 when it converged and the
// sample lies in the served set.
__global__ void k_jkr_samples(EnvDev env, MlpScale sc, uint64_t seed, int n, float* x, float* t, double* raw) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u_lo = sqrt(env.d_lo - kJkrDMinS), u_0 = sqrt(-kJkrDMinS), u_g = sqrt(kSurrGapHi - kJkrDMinS),
               u_hi = sqrt(env.d_hi - kJkrDMinS);
  const double len = (u_0 - u_lo) + (u_hi - u_g);
  const double v = len * u01(seed, i, 0);
  const double u = v < u_0 - u_lo ? u_lo + v : u_g + (v - (u_0 - u_lo));
  const double d = u <= u_0 ? fmin(u * u + kJkrDMinS, 0.0) : fmax(u * u + kJkrDMinS, kSurrGapHi);
  const double pi = 3.14159265358979323846;
  double rs = 0.0, es = 0.0, g = 0.0, s = 0.0;
  bool in = false;
  for (int tr = 0; tr < 64 && !in; ++tr) {
    rs = exp(env.lr_lo + (env.lr_hi - env.lr_lo) * u01(seed, i, 1 + 3 * tr));
    es = exp(env.le_lo + (env.le_hi - env.le_lo) * u01(seed, i, 2 + 3 * tr));
    g = exp(env.lg_lo + (env.lg_hi - env.lg_lo) * u01(seed, i, 3 + 3 * tr));
    s = cbrt(rs * rs * (8.0 * pi * g / es));
    in = s >= kSurrSLo * rs && s <= kSurrSHi * rs;
  }
  const double tt = s * s / rs;
  const double delta = d * tt;
  const PairConsts pc = pair_consts(es, g);
  int failed = 0;
  double a = 0.0;
  (void)jkr(delta, mk(0.0, 1.0, 0.0), pc, rs, &failed, &a);
  const double alpha = a / s;
  const double fh = (32.0 / 3.0) * alpha * alpha * alpha - 8.0 * alpha * sqrt(alpha);
  x[4 * i + 0] = scale_in((float)u, sc.in_lo[0], sc.in_hi[0]);
  x[4 * i + 1] = scale_in((float)log(rs), sc.in_lo[1], sc.in_hi[1]);
  x[4 * i + 2] = scale_in((float)log(es), sc.in_lo[2], sc.in_hi[2]);
  x[4 * i + 3] = scale_in((float)log(g), sc.in_lo[3], sc.in_hi[3]);
  t[2 * i + 0] = scale_in((float)alpha, sc.out_lo[0], sc.out_hi[0]);
  t[2 * i + 1] = scale_in((float)fh, sc.out_lo[1], sc.out_hi[1]);
  if (raw) {
    double* r = raw + 8 * (size_t)i;
    r[0] = delta;
    r[1] = rs;
    r[2] = es;
    r[3] = g;
    r[4] = a;
    r[5] = fh * pi * g * rs;
    r[6] = d;
    r[7] = (failed || !in) ? 0.0 : 1.0;
  }
}

// rows [r0, r0 + m) of x (4 per row) -> BF16 [m x 16], zero padded
__global__ void k_pack_x(const float* x, const int* order, int r0, int m, uint16_t* xb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m * 16) return;
  const int r = i / 16, k = i % 16;
  const int src = order ? order[r0 + r] : r0 + r;
  xb[i] = k < kMlpIn ? f2bf(x[(size_t)src * kMlpIn + k]) : (uint16_t)0;
}

// dY = 2 (Y - T) / (m * 2) for the 2 real outputs, 0 in the padding; BF16.
// Also accumulates the loss sum in FP64 (one value per workgroup).
__global__ void k_loss_grad(const float* Y, const float* t, const int* order, int r0, int m, uint16_t* dYb,
                            float* dY, double* loss_part) {
  __shared__ double red[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < m * 16) {
    const int r = i / 16, k = i % 16;
    float g = 0.0f;
    if (k < kMlpOut) {
      const int src = order ? order[r0 + r] : r0 + r;
      const float e = Y[(size_t)r * 16 + k] - t[(size_t)src * kMlpOut + k];
      g = 2.0f * e / (float)(m * kMlpOut);
      l = (double)e * (double)e;
    }
    dY[i] = g;
    dYb[i] = f2bf(g);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ = dH * (1 - H^2) (tanh'), FP32 and BF16
__global__ void k_tanh_back(const float* dH, const float* H, int n, float* dZ, uint16_t* dZb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float h = H[i];
  const float v = dH[i] * (1.0f - h * h);
  dZ[i] = v;
  dZb[i] = f2bf(v);
}

// column sums of a [m x n] FP32 matrix (bias gradients); one workgroup per column
__global__ void k_colsum(const float* A, int m, int n, float* out) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (int r = threadIdx.x; r < m; r += blockDim.x) s += A[(size_t)r * n + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

__global__ void k_adam(float* p, const float* g, float* m1, float* m2, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

__global__ void k_to_bf16(const float* a, uint16_t* b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

// ---- padded parameter layout used on the device --------------------------------
// FP32 master (padded): W1 [16 x 64] | b1 [64] | W2 [64 x 64] | b2 [64] | W3 [64 x 16] | b3 [16]
constexpr int kPW1 = 0, kPB1 = kPW1 + 16 * 64, kPW2 = kPB1 + 64, kPB2 = kPW2 + 64 * 64, kPW3 = kPB2 + 64,
              kPB3 = kPW3 + 64 * 16, kPadded = kPB3 + 16;

void pad_params(const float* p, std::vector<float>& out) {
  out.assign(kPadded, 0.0f);
  for (int k = 0; k < kMlpIn; ++k)
    for (int j = 0; j < kMlpHid; ++j) out[kPW1 + k * 64 + j] = p[kOffW1 + k * kMlpHid + j];
  for (int j = 0; j < 64; ++j) out[kPB1 + j] = p[kOffB1 + j];
  for (int i = 0; i < 64 * 64; ++i) out[kPW2 + i] = p[kOffW2 + i];
  for (int j = 0; j < 64; ++j) out[kPB2 + j] = p[kOffB2 + j];
  for (int k = 0; k < 64; ++k)
    for (int j = 0; j < kMlpOut; ++j) out[kPW3 + k * 16 + j] = p[kOffW3 + k * kMlpOut + j];
  for (int j = 0; j < kMlpOut; ++j) out[kPB3 + j] = p[kOffB3 + j];
}

void unpad_params(const std::vector<float>& pp, float* p) {
  for (int k = 0; k < kMlpIn; ++k)
    for (int j = 0; j < kMlpHid; ++j) p[kOffW1 + k * kMlpHid + j] = pp[kPW1 + k * 64 + j];
  for (int j = 0; j < 64; ++j) p[kOffB1 + j] = pp[kPB1 + j];
  for (int i = 0; i < 64 * 64; ++i) p[kOffW2 + i] = pp[kPW2 + i];
  for (int j = 0; j < 64; ++j) p[kOffB2 + j] = pp[kPB2 + j];
  for (int k = 0; k < 64; ++k)
    for (int j = 0; j < kMlpOut; ++j) p[kOffW3 + k * kMlpOut + j] = pp[kPW3 + k * 16 + j];
  for (int j = 0; j < kMlpOut; ++j) p[kOffB3 + j] = pp[kPB3 + j];
}

int blocks_for(int n, int b = 256) { return (n + b - 1) / b; }

// Device workspace for forward + backward on a batch of m rows (m % 16 == 0).
struct Work {
  int m = 0;
  float *P = nullptr, *G = nullptr, *M1 = nullptr, *M2 = nullptr;  // padded params, grads, Adam moments
  uint16_t* Wb = nullptr;                                          // bf16 copy of padded params
  uint16_t *Xb = nullptr, *H1b = nullptr, *H2b = nullptr, *dYb = nullptr, *dZ2b = nullptr, *dZ1b = nullptr;
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH2 = nullptr, *dZ2 = nullptr, *dH1 = nullptr,
        *dZ1 = nullptr;
  double* lossp = nullptr;
  rp_status alloc(int rows) {
    m = rows;
    SHIP(hipMalloc(&P, kPadded * sizeof(float)));
    SHIP(hipMalloc(&G, kPadded * sizeof(float)));
    SHIP(hipMalloc(&M1, kPadded * sizeof(float)));
    SHIP(hipMalloc(&M2, kPadded * sizeof(float)));
    SHIP(hipMalloc(&Wb, kPadded * sizeof(uint16_t)));
    SHIP(hipMemset(M1, 0, kPadded * sizeof(float)));
    SHIP(hipMemset(M2, 0, kPadded * sizeof(float)));
    SHIP(hipMemset(G, 0, kPadded * sizeof(float)));
    size_t r = (size_t)rows;
    SHIP(hipMalloc(&Xb, r * 16 * 2));
    SHIP(hipMalloc(&H1b, r * 64 * 2));
    SHIP(hipMalloc(&H2b, r * 64 * 2));
    SHIP(hipMalloc(&dYb, r * 16 * 2));
    SHIP(hipMalloc(&dZ2b, r * 64 * 2));
    SHIP(hipMalloc(&dZ1b, r * 64 * 2));
    SHIP(hipMalloc(&H1, r * 64 * 4));
    SHIP(hipMalloc(&H2, r * 64 * 4));
    SHIP(hipMalloc(&Y, r * 16 * 4));
    SHIP(hipMalloc(&dY, r * 16 * 4));
    SHIP(hipMalloc(&dH2, r * 64 * 4));
    SHIP(hipMalloc(&dZ2, r * 64 * 4));
    SHIP(hipMalloc(&dH1, r * 64 * 4));
    SHIP(hipMalloc(&dZ1, r * 64 * 4));
    SHIP(hipMalloc(&lossp, (size_t)blocks_for((int)r * 16) * sizeof(double)));
    return RP_OK;
  }
  void release() {
    for (void* q : {(void*)P, (void*)G, (void*)M1, (void*)M2, (void*)Wb, (void*)Xb, (void*)H1b, (void*)H2b,
                    (void*)dYb, (void*)dZ2b, (void*)dZ1b, (void*)H1, (void*)H2, (void*)Y, (void*)dY, (void*)dH2,
                    (void*)dZ2, (void*)dH1, (void*)dZ1, (void*)lossp})
      if (q) (void)hipFree(q);
  }
};

// forward + backward on rows [r0, r0 + m) of (x, t); leaves the loss parts in
// w.lossp and the gradients (padded) in w.G.
void fwd_bwd(Work& w, hipStream_t s, const float* x, const float* t, const int* order, int r0) {
  const int m = w.m;
  const uint16_t *W1 = w.Wb + kPW1, *W2 = w.Wb + kPW2, *W3 = w.Wb + kPW3;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(kPadded)), dim3(256), 0, s, w.P, w.Wb, kPadded);
  hipLaunchKernelGGL(k_pack_x, dim3(blocks_for(m * 16)), dim3(256), 0, s, x, order, r0, m, w.Xb);
  gemm<false, false>(s, w.Xb, 16, W1, 64, w.H1, 64, w.H1b, m, 64, 16, w.P + kPB1, 1);
  gemm<false, false>(s, w.H1b, 64, W2, 64, w.H2, 64, w.H2b, m, 64, 64, w.P + kPB2, 1);
  gemm<false, false>(s, w.H2b, 64, W3, 16, w.Y, 16, nullptr, m, 16, 64, w.P + kPB3, 0);
  hipLaunchKernelGGL(k_loss_grad, dim3(blocks_for(m * 16)), dim3(256), 0, s, w.Y, t, order, r0, m, w.dYb, w.dY,
                     w.lossp);
  gemm<true, false>(s, w.H2b, 64, w.dYb, 16, w.G + kPW3, 16, nullptr, 64, 16, m, nullptr, 0);
  hipLaunchKernelGGL(k_colsum, dim3(16), dim3(256), 0, s, w.dY, m, 16, w.G + kPB3);
  gemm<false, true>(s, w.dYb, 16, W3, 16, w.dH2, 64, nullptr, m, 64, 16, nullptr, 0);
  hipLaunchKernelGGL(k_tanh_back, dim3(blocks_for(m * 64)), dim3(256), 0, s, w.dH2, w.H2, m * 64, w.dZ2, w.dZ2b);
  hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, s, w.dZ2, m, 64, w.G + kPB2);
  gemm<true, false>(s, w.H1b, 64, w.dZ2b, 64, w.G + kPW2, 64, nullptr, 64, 64, m, nullptr, 0);
  gemm<false, true>(s, w.dZ2b, 64, W2, 64, w.dH1, 64, nullptr, m, 64, 64, nullptr, 0);
  hipLaunchKernelGGL(k_tanh_back, dim3(blocks_for(m * 64)), dim3(256), 0, s, w.dH1, w.H1, m * 64, w.dZ1, w.dZ1b);
  hipLaunchKernelGGL(k_colsum, dim3(64), dim3(256), 0, s, w.dZ1, m, 64, w.G + kPB1);
  gemm<true, false>(s, w.Xb, 16, w.dZ1b, 64, w.G + kPW1, 64, nullptr, 16, 64, m, nullptr, 0);
}

double read_loss(Work& w, hipStream_t s) {
  int nb = blocks_for(w.m * 16);
  std::vector<double> lp((size_t)nb);
  (void)hipMemcpyAsync(lp.data(), w.lossp, nb * sizeof(double), hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  double l = 0.0;
  for (double v : lp) l += v;
  return l / (double)(w.m * kMlpOut);
}

// host FP64 solve of alpha^2 - sqrt(alpha) = d on the stable branch
double alpha_of(double d) {
  const double ac = std::pow(0.25, 2.0 / 3.0);
  double lo = ac, hi = std::max(2.0, std::sqrt(std::fabs(d)) + 2.0);
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    double f = mid * mid - std::sqrt(mid) - d;
    if (f > 0) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}
double fhat_of(double a) { return (32.0 / 3.0) * a * a * a - 8.0 * a * std::sqrt(a); }

}  // namespace

int envelope_check(const rp_jkr_envelope& e, const char** why) {
  auto bad = [&](const char* m) {
    *why = m;
    return (int)RP_ERR_INVALID;
  };
  if (!(e.d_lo >= kSurrDLoMin && e.d_lo <= 0.0)) return bad("envelope d_lo must lie in [-0.40, 0]");
  if (!(e.d_hi > kSurrGapHi && e.d_hi <= 1.0e4)) return bad("envelope d_hi must lie in (0.2, 1e4]");
  const double lo[3] = {e.rstar_lo, e.estar_lo, e.gamma_lo}, hi[3] = {e.rstar_hi, e.estar_hi, e.gamma_hi};
  for (int k = 0; k < 3; ++k)
    if (!(lo[k] > 0.0 && hi[k] >= lo[k] && std::isfinite(hi[k])))
      return bad("envelope R*, E*, gamma ranges must be positive, finite and ordered");
  *why = "";
  return RP_OK;
}

MlpScale make_scale(const rp_jkr_envelope& e) {
  MlpScale s{};
  s.in_lo[0] = (float)std::sqrt(e.d_lo - kJkrDMinS);
  s.in_hi[0] = (float)std::sqrt(e.d_hi - kJkrDMinS);
  s.in_lo[1] = (float)std::log(e.rstar_lo);
  s.in_hi[1] = (float)std::log(e.rstar_hi);
  s.in_lo[2] = (float)std::log(e.estar_lo);
  s.in_hi[2] = (float)std::log(e.estar_hi);
  s.in_lo[3] = (float)std::log(e.gamma_lo);
  s.in_hi[3] = (float)std::log(e.gamma_hi);
  for (int k = 1; k < 4; ++k)
    if (!(s.in_hi[k] > s.in_lo[k])) {  // a fixed quantity: give the range a width
      s.in_lo[k] -= 0.5f;
      s.in_hi[k] += 0.5f;
    }
  const double a_lo = alpha_of(e.d_lo), a_hi = alpha_of(e.d_hi);
  double f_lo = 1e300, f_hi = -1e300;
  for (int i = 0; i <= 1000; ++i) {
    double a = a_lo + (a_hi - a_lo) * i / 1000.0;
    f_lo = std::min(f_lo, fhat_of(a));
    f_hi = std::max(f_hi, fhat_of(a));
  }
  s.out_lo[0] = (float)a_lo;
  s.out_hi[0] = (float)a_hi;
  s.out_lo[1] = (float)f_lo;
  s.out_hi[1] = (float)f_hi;
  return s;
}

MlpDev MlpHost::dev() const {
  MlpDev d{};
  d.w1 = w;
  d.w2 = w + 16 * 64;
  d.w3 = w + 16 * 64 + 64 * 64;
  d.b1 = b;
  d.b2 = b + 64;
  d.b3 = b + 128;
  return d;
}

int mlp_upload(MlpHost& m, const float* params, hipStream_t stream) {
  if (!m.w && hipMalloc(&m.w, (16 * 64 + 64 * 64 + 64 * 16) * sizeof(uint16_t)) != hipSuccess) return RP_ERR_NOMEM;
  if (!m.b && hipMalloc(&m.b, (64 + 64 + 16) * sizeof(float)) != hipSuccess) return RP_ERR_NOMEM;
  std::vector<float> pp;
  pad_params(params, pp);
  // Fragment order: lane l of W1's tile nt holds the 16x16x16 B operand
  // W1[4 (l / 16) + q][16 nt + l % 16], q = 0..3, at 4 (64 nt + l) + q; lane
  // l of W2's tile t = 4 kt + nt holds the 16x16x32 B operand
  // W2[32 kt + 8 (l / 16) + q][16 nt + l % 16], q = 0..7, at 8 (64 t + l) + q,
  // and of W3's tile kt W3[32 kt + 8 (l / 16) + q][l % 16]; each lane loads
  // its operand of a tile as one 8- or 16-byte word.
  std::vector<uint16_t> wb(16 * 64 + 64 * 64 + 64 * 16);
  uint16_t* f1 = wb.data();
  uint16_t* f2 = f1 + 16 * 64;
  uint16_t* f3 = f2 + 64 * 64;
  for (int l = 0; l < 64; ++l) {
    const int li = l % 16, kq = 4 * (l / 16), k8 = 8 * (l / 16);
    for (int q = 0; q < 4; ++q)
      for (int nt = 0; nt < 4; ++nt) f1[4 * (64 * nt + l) + q] = f2bf(pp[kPW1 + (kq + q) * 64 + nt * 16 + li]);
    for (int q = 0; q < 8; ++q)
      for (int kt = 0; kt < 2; ++kt) {
        for (int nt = 0; nt < 4; ++nt)
          f2[8 * (64 * (kt * 4 + nt) + l) + q] = f2bf(pp[kPW2 + (kt * 32 + k8 + q) * 64 + nt * 16 + li]);
        f3[8 * (64 * kt + l) + q] = f2bf(pp[kPW3 + (kt * 32 + k8 + q) * 16 + li]);
      }
  }
  std::vector<float> bb(64 + 64 + 16);
  for (int i = 0; i < 64; ++i) bb[i] = pp[kPB1 + i];
  for (int i = 0; i < 64; ++i) bb[64 + i] = pp[kPB2 + i];
  for (int i = 0; i < 16; ++i) bb[128 + i] = pp[kPB3 + i];
  if (hipMemcpyAsync(m.w, wb.data(), wb.size() * 2, hipMemcpyHostToDevice, stream) != hipSuccess) return RP_ERR_DEVICE;
  if (hipMemcpyAsync(m.b, bb.data(), bb.size() * 4, hipMemcpyHostToDevice, stream) != hipSuccess) return RP_ERR_DEVICE;
  if (hipStreamSynchronize(stream) != hipSuccess) return RP_ERR_DEVICE;
  return RP_OK;
}

void mlp_free(MlpHost& m) {
  if (m.w) (void)hipFree(m.w);
  if (m.b) (void)hipFree(m.b);
  m.w = nullptr;
  m.b = nullptr;
}

void mlp_infer(const MlpDev& net, const float* x, float* y, int rows, hipStream_t stream) {
  if (rows <= 0) return;
  const int tiles = (rows + 15) / 16;
  hipLaunchKernelGGL(k_mlp_infer, dim3((tiles + 3) / 4), dim3(256), 0, stream, net, x, y, rows);
}

void mlp_infer_counted(const MlpDev& net, const float* x, float* y, const unsigned* counter, int cap,
                       hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
  const int blocks = std::max(1, std::min(1024, (cap + 63) / 64));
  if (start && stop)
    hipExtLaunchKernelGGL(k_mlp_counted, dim3(blocks), dim3(256), 0, stream, start, stop, 0u, net, x, y, counter, cap);
  else
    hipLaunchKernelGGL(k_mlp_counted, dim3(blocks), dim3(256), 0, stream, net, x, y, counter, cap);
}

// ---- inference without allocation (rp_mlp_infer_noalloc) ---------------------
// The packed network -- BF16 weights in MFMA fragment order (the layout of
// mlp_upload) and FP32 biases -- lives in this file-scope device array
// instead of a hipMalloc'd buffer.  Each GPU holds its own copy at its own
// address, kept until the process ends.  Host code never passes `scratch`
// itself to a kernel (its host-side name is not a device address): it uses
// the per-device address from hipGetSymbolAddress (noalloc_scratch below).
constexpr int kPackedW = 16 * 64 + 64 * 64 + 64 * 16;          // BF16 weights
constexpr int kPackedB = 64 + 64 + 16;                          // FP32 biases
constexpr size_t kPackedBytes = (size_t)kPackedW * 2 + (size_t)kPackedB * 4;  // 12,864 B
static_assert(kPackedBytes <= RP_MLP_NOALLOC_SCRATCH, "the packed network must fit the scratch array");
static __device__ __align__(256) unsigned char scratch[RP_MLP_NOALLOC_SCRATCH];

// Packs the FP32 parameter vector (device memory, RP_MLP_PARAMS floats) into
// the scratch layout: element idx of the fragment-order weights as
// mlp_upload writes it (zero padding of W1's rows 4..15 and W3's columns
// 2..15), then the biases.
__global__ void __launch_bounds__(256) k_pack_net(const float* __restrict__ p, uint16_t* __restrict__ w,
                                                  float* __restrict__ b) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < kPackedW) {
    int j = idx, base = 0;
    const int part = j < 16 * 64 ? 0 : (j < 16 * 64 + 64 * 64 ? 1 : 2);
    if (part == 1) base = 16 * 64;
    if (part == 2) base = 16 * 64 + 64 * 64;
    j -= base;
    float v;
    if (part == 0) {  // W1 tile nt = t: row 4 (l / 16) + q, column 16 t + l % 16
      const int q = j & 3, rest = j >> 2, l = rest & 63, t = rest >> 6;
      const int row = 4 * (l >> 4) + q;
      v = row < kMlpIn ? p[kOffW1 + row * kMlpHid + 16 * t + (l & 15)] : 0.0f;
    } else {
      const int q = j & 7, rest = j >> 3, l = rest & 63, t = rest >> 6;
      const int li = l & 15, k8 = 8 * (l >> 4);
      if (part == 1) {  // W2 tile t = 4 kt + nt: row 32 kt + 8 (l / 16) + q, column 16 nt + l % 16
        const int row = (t >> 2) * 32 + k8 + q;
        v = p[kOffW2 + row * kMlpHid + (t & 3) * 16 + li];
      } else {  // W3 tile kt = t: row 32 t + 8 (l / 16) + q, column l % 16
        const int row = t * 32 + k8 + q;
        v = li < kMlpOut ? p[kOffW3 + row * kMlpOut + li] : 0.0f;
      }
    }
    w[idx] = f2bf(v);
  } else if (idx < kPackedW + kPackedB) {
    const int j = idx - kPackedW;
    b[j] = j < 64 ? p[kOffB1 + j] : (j < 128 ? p[kOffB2 + j - 64] : (j - 128 < kMlpOut ? p[kOffB3 + j - 128] : 0.0f));
  }
}

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_dev;  // device address of `scratch`, per device

// The device address of `scratch` on device `dev` (current): looked up once
// per device with hipGetSymbolAddress, then reused.  nullptr on failure.
void* noalloc_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((size_t)dev >= g_scratch_dev.size()) g_scratch_dev.resize((size_t)dev + 1, nullptr);
  if (!g_scratch_dev[(size_t)dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return nullptr;
    g_scratch_dev[(size_t)dev] = ptr;
  }
  return g_scratch_dev[(size_t)dev];
}

}  // namespace rp

using namespace rp;

extern "C" {

rp_status rp_mlp_infer_noalloc(const float* params_dev, const float* x_scaled_dev, size_t rows,
                               float* y_scaled_dev, void* stream) {
  if (!params_dev || (rows && (!x_scaled_dev || !y_scaled_dev))) return sfail(RP_ERR_INVALID, "NULL argument");
  if (rows > (size_t)RP_MLP_NOALLOC_MAX_ROWS)
    return sfail(RP_ERR_INVALID, "rp_mlp_infer_noalloc: rows exceeds RP_MLP_NOALLOC_MAX_ROWS (" +
                                     std::to_string((long long)RP_MLP_NOALLOC_MAX_ROWS) + ")");
  int dev = 0;
  SHIP(hipGetDevice(&dev));
  char* base = (char*)noalloc_scratch(dev);
  if (!base) return sfail(RP_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed");
  uint16_t* w = (uint16_t*)base;
  float* b = (float*)(base + (size_t)kPackedW * 2);
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_pack_net, dim3((kPackedW + kPackedB + 255) / 256), dim3(256), 0, s, params_dev, w, b);
  MlpDev net{};
  net.w1 = w;
  net.w2 = w + 16 * 64;
  net.w3 = w + 16 * 64 + 64 * 64;
  net.b1 = b;
  net.b2 = b + 64;
  net.b3 = b + 128;
  mlp_infer(net, x_scaled_dev, y_scaled_dev, (int)rows, s);
  SHIP(hipGetLastError());
  return RP_OK;
}

const char* rp_surrogate_last_error(void) { return s_err.c_str(); }

rp_status rp_mfma_probe(const float* a16x16, const float* b16x16, float* c16x16) {
  if (!a16x16 || !b16x16 || !c16x16) return sfail(RP_ERR_INVALID, "NULL argument");
  std::vector<uint16_t> ab(256), bb(256);
  for (int i = 0; i < 256; ++i) {
    ab[i] = f2bf(a16x16[i]);
    bb[i] = f2bf(b16x16[i]);
  }
  uint16_t *da = nullptr, *db = nullptr;
  float* dc = nullptr;
  SHIP(hipMalloc(&da, 512));
  SHIP(hipMalloc(&db, 512));
  SHIP(hipMalloc(&dc, 1024));
  SHIP(hipMemcpy(da, ab.data(), 512, hipMemcpyHostToDevice));
  SHIP(hipMemcpy(db, bb.data(), 512, hipMemcpyHostToDevice));
  gemm<false, false>(nullptr, da, 16, db, 16, dc, 16, nullptr, 16, 16, 16, nullptr, 0);
  SHIP(hipDeviceSynchronize());
  SHIP(hipMemcpy(c16x16, dc, 1024, hipMemcpyDeviceToHost));
  (void)hipFree(da);
  (void)hipFree(db);
  (void)hipFree(dc);
  return RP_OK;
}

rp_status rp_mlp_infer(const float* params, const float* x_scaled, size_t rows, float* y_scaled) {
  if (!params || (!x_scaled && rows) || (!y_scaled && rows)) return sfail(RP_ERR_INVALID, "NULL argument");
  MlpHost m;
  if (mlp_upload(m, params, nullptr) != RP_OK) return sfail(RP_ERR_DEVICE, "weight upload failed");
  float *dx = nullptr, *dy = nullptr;
  SHIP(hipMalloc(&dx, std::max<size_t>(rows, 1) * 4 * sizeof(float)));
  SHIP(hipMalloc(&dy, std::max<size_t>(rows, 1) * 2 * sizeof(float)));
  SHIP(hipMemcpy(dx, x_scaled, rows * 4 * sizeof(float), hipMemcpyHostToDevice));
  mlp_infer(m.dev(), dx, dy, (int)rows, nullptr);
  SHIP(hipDeviceSynchronize());
  SHIP(hipMemcpy(y_scaled, dy, rows * 2 * sizeof(float), hipMemcpyDeviceToHost));
  (void)hipFree(dx);
  (void)hipFree(dy);
  mlp_free(m);
  return RP_OK;
}

rp_status rp_mlp_gradient(const float* params, const float* x_scaled, const float* t_scaled, size_t rows,
                          float* grad, double* loss) {
  if (!params || !x_scaled || !t_scaled || !grad || !loss) return sfail(RP_ERR_INVALID, "NULL argument");
  if (rows == 0 || rows % 16) return sfail(RP_ERR_INVALID, "rows must be a positive multiple of 16");
  Work w;
  rp_status st = w.alloc((int)rows);
  if (st) return st;
  std::vector<float> pp;
  pad_params(params, pp);
  float *dx = nullptr, *dt = nullptr;
  SHIP(hipMalloc(&dx, rows * 4 * sizeof(float)));
  SHIP(hipMalloc(&dt, rows * 2 * sizeof(float)));
  SHIP(hipMemcpy(dx, x_scaled, rows * 4 * sizeof(float), hipMemcpyHostToDevice));
  SHIP(hipMemcpy(dt, t_scaled, rows * 2 * sizeof(float), hipMemcpyHostToDevice));
  SHIP(hipMemcpy(w.P, pp.data(), kPadded * sizeof(float), hipMemcpyHostToDevice));
  fwd_bwd(w, nullptr, dx, dt, nullptr, 0);
  SHIP(hipDeviceSynchronize());
  *loss = read_loss(w, nullptr);
  std::vector<float> g(kPadded);
  SHIP(hipMemcpy(g.data(), w.G, kPadded * sizeof(float), hipMemcpyDeviceToHost));
  unpad_params(g, grad);
  (void)hipFree(dx);
  (void)hipFree(dt);
  w.release();
  return RP_OK;
}

rp_status rp_jkr_samples(const rp_jkr_envelope* env, uint64_t seed, size_t n, float* x_scaled, float* t_scaled,
                         double* raw8) {
  if (!env || !n) return sfail(RP_ERR_INVALID, "bad arguments");
  const char* why = "";
  if (envelope_check(*env, &why) != RP_OK) return sfail(RP_ERR_INVALID, why);
  MlpScale sc = make_scale(*env);
  EnvDev ed{env->d_lo, env->d_hi, std::log(env->rstar_lo), std::log(env->rstar_hi), std::log(env->estar_lo),
            std::log(env->estar_hi), std::log(env->gamma_lo), std::log(env->gamma_hi)};
  float *dx = nullptr, *dt = nullptr;
  double* dr = nullptr;
  SHIP(hipMalloc(&dx, n * 4 * sizeof(float)));
  SHIP(hipMalloc(&dt, n * 2 * sizeof(float)));
  if (raw8) SHIP(hipMalloc(&dr, n * 8 * sizeof(double)));
  hipLaunchKernelGGL(k_jkr_samples, dim3(blocks_for((int)n)), dim3(256), 0, nullptr, ed, sc, (uint64_t)seed, (int)n,
                     dx, dt, dr);
  SHIP(hipDeviceSynchronize());
  if (x_scaled) SHIP(hipMemcpy(x_scaled, dx, n * 4 * sizeof(float), hipMemcpyDeviceToHost));
  if (t_scaled) SHIP(hipMemcpy(t_scaled, dt, n * 2 * sizeof(float), hipMemcpyDeviceToHost));
  if (raw8) SHIP(hipMemcpy(raw8, dr, n * 8 * sizeof(double), hipMemcpyDeviceToHost));
  (void)hipFree(dx);
  (void)hipFree(dt);
  if (dr) (void)hipFree(dr);
  return RP_OK;
}

rp_status rp_jkr_surrogate_train(const rp_jkr_envelope* env, uint64_t seed, size_t nsamples, int epochs, int batch,
                                 float lr, float* params_out, double* epoch_loss) {
  if (!env || !params_out || nsamples < 16 || epochs < 1) return sfail(RP_ERR_INVALID, "bad arguments");
  if (batch < 16 || batch % 16) return sfail(RP_ERR_INVALID, "batch must be a positive multiple of 16");
  hipStream_t s;
  SHIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  const int n = (int)(nsamples / batch) * batch;
  const char* why = "";
  if (envelope_check(*env, &why) != RP_OK) return sfail(RP_ERR_INVALID, why);
  MlpScale sc = make_scale(*env);
  EnvDev ed{env->d_lo, env->d_hi, std::log(env->rstar_lo), std::log(env->rstar_hi), std::log(env->estar_lo),
            std::log(env->estar_hi), std::log(env->gamma_lo), std::log(env->gamma_hi)};
  float *dx = nullptr, *dt = nullptr;
  int* dorder = nullptr;
  SHIP(hipMalloc(&dx, (size_t)n * 4 * sizeof(float)));
  SHIP(hipMalloc(&dt, (size_t)n * 2 * sizeof(float)));
  SHIP(hipMalloc(&dorder, (size_t)n * sizeof(int)));
  hipLaunchKernelGGL(k_jkr_samples, dim3(blocks_for(n)), dim3(256), 0, s, ed, sc, (uint64_t)seed, n, dx, dt,
                     (double*)nullptr);
  Work w;
  rp_status st = w.alloc(batch);
  if (st) return st;
  // Xavier-uniform initial weights, zero biases (seeded, host)
  std::vector<float> p0(kMlpParams, 0.0f);
  uint64_t z = seed ^ 0xA5A5A5A5ull;
  auto urand = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t q = z;
    q = (q ^ (q >> 30)) * 0xBF58476D1CE4E5B9ull;
    q = (q ^ (q >> 27)) * 0x94D049BB133111EBull;
    return (double)((q ^ (q >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int fin, int fout) {
    double lim = std::sqrt(6.0 / (fin + fout));
    for (int i = 0; i < fin * fout; ++i) p0[off + i] = (float)((2.0 * urand() - 1.0) * lim);
  };
  fill(kOffW1, kMlpIn, kMlpHid);
  fill(kOffW2, kMlpHid, kMlpHid);
  fill(kOffW3, kMlpHid, kMlpOut);
  std::vector<float> pp;
  pad_params(p0.data(), pp);
  SHIP(hipMemcpyAsync(w.P, pp.data(), kPadded * sizeof(float), hipMemcpyHostToDevice, s));
  std::vector<int> order((size_t)n);
  for (int i = 0; i < n; ++i) order[i] = i;
  const int nb = n / batch;
  int t_adam = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    // shuffle (seeded Fisher-Yates on the host)
    for (int i = n - 1; i > 0; --i) std::swap(order[i], order[(int)(urand() * (i + 1))]);
    SHIP(hipMemcpyAsync(dorder, order.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, s));
    // cosine learning-rate decay to 5% of lr
    const float lr_ep = lr * (float)(0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979 * ep / epochs)));
    double lsum = 0.0;
    for (int b = 0; b < nb; ++b) {
      fwd_bwd(w, s, dx, dt, dorder, b * batch);
      ++t_adam;
      const float b1 = 0.9f, b2 = 0.999f;
      const float c1 = 1.0f - std::pow(b1, (float)t_adam), c2 = 1.0f - std::pow(b2, (float)t_adam);
      hipLaunchKernelGGL(k_adam, dim3(blocks_for(kPadded)), dim3(256), 0, s, w.P, w.G, w.M1, w.M2, kPadded, lr_ep,
                         b1, b2, 1e-8f, c1, c2);
      if (epoch_loss && b == nb - 1) lsum = read_loss(w, s);
    }
    if (epoch_loss) epoch_loss[ep] = lsum;
  }
  SHIP(hipStreamSynchronize(s));
  SHIP(hipMemcpy(pp.data(), w.P, kPadded * sizeof(float), hipMemcpyDeviceToHost));
  unpad_params(pp, params_out);
  w.release();
  (void)hipFree(dx);
  (void)hipFree(dt);
  (void)hipFree(dorder);
  (void)hipStreamDestroy(s);
  return RP_OK;
}

rp_status rp_jkr_envelope_scale(const rp_jkr_envelope* env, float* in_lo4, float* in_hi4, float* out_lo2,
                                float* out_hi2) {
  if (!env) return sfail(RP_ERR_INVALID, "NULL argument");
  const char* why = "";
  if (envelope_check(*env, &why) != RP_OK) return sfail(RP_ERR_INVALID, why);
  MlpScale sc = make_scale(*env);
  for (int k = 0; k < 4; ++k) {
    if (in_lo4) in_lo4[k] = sc.in_lo[k];
    if (in_hi4) in_hi4[k] = sc.in_hi[k];
  }
  for (int k = 0; k < 2; ++k) {
    if (out_lo2) out_lo2[k] = sc.out_lo[k];
    if (out_hi2) out_hi2[k] = sc.out_hi[k];
  }
  return RP_OK;
}

}  // extern "C"
