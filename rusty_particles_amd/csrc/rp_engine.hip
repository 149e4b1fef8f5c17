// rp_engine.hip — MI355X (gfx950) DEM time step behind the rp_abi.h C ABI.
//
// One engine holds one simulation domain on one GPU.  Per time step it runs
// the reference's Simulation::step_cpu (src/simulation.rs:173-562) as:
//
//   rebuild chain (runs only when a device-side flag asks for it; otherwise
//   every kernel of the chain returns at once):
//     k_keys        fine-cell key + per-cell histogram (integer atomics)
//     k_scan_*      exclusive scan of the histogram -> cell start offsets
//     k_scatter     particles into their cell's range (order within the cell
//                   still arbitrary)
//     k_rank_permute rank inside the cell by stable id (deterministic), move
//                   the particle state to its new slot, drop removed ones
//     k_cell_rmax   largest radius per cell (cells well below the largest
//                   cutoff only: k_build then skips cells out of reach)
//     k_build       Verlet neighbour list per particle (cutoff + skin), the
//                   last workgroup to finish swaps the state buffers
//   k_force         gather over the neighbour list: exact pair predicate,
//                   reference-grid visibility, normal + tangential laws,
//                   triangles of the particle's cell, then a = g + F/m,
//                   tau = T and the second half kick; when another step
//                   follows in the same call, also the next step's periodic
//                   wrap, removal test, first half kick and drift, so the
//                   state makes one round trip through HBM per step.
//
// Memory layout (HBM): structure of arrays, FP64, each vec3 field stored as
// three planes of `cap` doubles, except position + radius, interleaved as one
// 32-byte record per slot.  Two copies of the per-step state (x, v, w, their
// Kahan residuals, a, tau): the force kernel reads one and writes the other.
// Radius, mass, flags and stable id are one copy plus a scratch copy used
// while a rebuild re-orders the particles.  Neighbour lists are column-major
// (entry k of slot i at nbr[k*cap + i]).  No float atomics anywhere: results
// are bitwise reproducible run to run.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/rp_abi.h"
#include "rp_physics.hpp"
#include "rp_stl.hpp"
#include "rp_surrogate.hpp"

using namespace rp;

namespace {

thread_local std::string g_last_error;

rp_status fail(rp_status code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess)                                                                     \
      return fail(RP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e));          \
  } while (0)

constexpr int kBlock = 256;
// Workgroup of the force kernel: 128 lanes (2 waves) measured 1-2.5 % faster
// than 256 on the hopper, JKR and batch workloads (more, smaller workgroups
// fill the CUs more evenly); 512 was slower.
#ifndef RP_FORCE_BLOCK
#define RP_FORCE_BLOCK 128
#endif
constexpr int kForceBlock = RP_FORCE_BLOCK;
// Four waves per SIMD for the force kernels without a tangential history:
// the JKR and sJKR variants need 140-150 VGPRs (three waves) unconstrained;
// capped at 128 they spill a few registers and still run 5-12 % faster (batch
// 592 -> 532 us, JKR shortcut 406 -> 361 us per launch); the hopper kernel
// fits in 128 anyway.  The physics-fix history kernels spill 60-200 B at 128
// and run faster at three waves with none (16M drum shortcut 2,849 -> 2,684
// us, surrogate 3,626 -> 3,003 us; JKR ensemble 206 -> 193 and 256 -> 209 us).
// The expression is evaluated per instantiation (HIST is the kernel's
// template argument).
#ifndef RP_FORCE_ATTR
#define RP_FORCE_ATTR __attribute__((amdgpu_waves_per_eu(HIST ? 3 : 4)))
#endif
constexpr uint8_t FLAG_FIXED = 1;
constexpr uint8_t FLAG_DEAD = 2;
constexpr int kScanItems = 16;                   // items per thread in the scan
constexpr int kScanChunk = kBlock * kScanItems;  // items per scan workgroup
constexpr int kMaxGrid = 2048;                   // grid cap of the grid-stride rebuild kernels
constexpr uint64_t kRebuildInterval = 10;        // steps between launches of the rebuild chain
constexpr long long kNoStall = 0x7fffffffffffffffLL;  // no stalled step recorded
// Launches per captured step loop: one rebuild chain and kRebuildInterval
// force launches (an even number, so the state buffers are where they were at
// the capture); the rebuild-request parity alternates from one replay to the
// next, and each parity has its own executable (graph_exec).  Ten launches
// let a 20-step call replay one loop.
constexpr uint64_t kGraphSteps = kRebuildInterval;
#ifndef RP_DEAL_BATCH
#define RP_DEAL_BATCH 4
#endif
constexpr int kDealBatch = RP_DEAL_BATCH;  // entries per lane and round of the dealt zero-force pass
constexpr int kNbrBatch = 4;                     // neighbours whose loads the force kernel issues together
constexpr double kPi = 3.14159265358979323846;
// Margin (in d = delta / t) of the surrogate's near masks: a contact whose d
// changes by more than this in one step, into the served set, takes the exact
// solve that step.  A step moves d by (relative speed) dt / t, about 2e-3 in
// the 2M-particle JKR packing.
constexpr double kSurrMaskMargin = 0.5;

// Per-step particle state of one buffer (device pointers, `cap` slots).
// The force kernel reads one buffer and writes the other, so no thread can
// see a neighbour's next-step values while gathering this step's forces.
struct Dyn {
  double* xr;  // position and radius, interleaved (x, y, z, r) per slot: a
               // neighbour test touches one 32-byte record
  double* v;
  double* w;
  double* rx;  // Kahan residuals of x, v, w
  double* rv;
  double* rw;
  double* a;   // acceleration and torque of the last step; written only by
  double* tau; // a step that ends at rest, read by the next call's first half
};

// Per-particle constants of one buffer; swapped only when a rebuild
// re-orders the particles.
struct Stat {
  double* r;
  double* m;
  uint8_t* flags;
  uint32_t* sid;  // stable id = insertion index (API order key)
  uint16_t* inst; // instance of a batched engine (0 for a single simulation)
};

// Per-instance constants of a batched engine (materials differ by instance).
struct InstConsts {
  MaterialConsts mc;
  double jkr_far_pp, jkr_far_pw;
  double sjkr_zero_pp;
};

// Device-resident control block: slot count and rare flags.  Per-step
// counters live in per-workgroup arrays written with plain stores, so no
// kernel performs same-address atomics on its common path.
struct Ctl {
  int n;                   // slots in use (alive + removed-not-yet-compacted)
  int alive;               // alive particles
  int rebuild[2];          // rebuild requests, indexed by step parity
  int overflow;            // neighbour list overflow seen
  int max_count;           // largest candidate count seen by k_build
  unsigned long long rebuilds;  // rebuilds done
  int pair_overflow;       // pair record overflow
  // First step that must not run on the current neighbour lists: a particle
  // moved more than skin/2 since the last build (its list is no longer a
  // superset).  Force launches of steps at or after it leave the state
  // unchanged; the next rebuild resets it (kNoStall) and the host re-issues
  // the steps.  Steps are labelled by their offset from a base the host
  // moves (k_shift_labels): plain launches carry offsets that grow, and a
  // captured step loop shifts the base by its length at its end, so each
  // replay labels its launches afresh.
  long long stall_off;
  unsigned long long steps_done;  // force launches that advanced the state
  int at_rest;             // 1 when the last advancing launch left the at-rest state
  // rebuilds + 1 when the force kernel last wrote the surrogate's near masks
  // (a rebuild re-orders the lists, so the next collect scans every entry)
  unsigned long long surr_mask_epoch;
  unsigned surr_rows_max;  // most rows any range was asked for in one step (sizes the row buffers)
  // Absolute launch number of the label base (k_shift_labels moves both):
  // label_base + off numbers every force launch uniquely.
  unsigned long long label_base;
  // Steps with excessive overlap: the last launch number that saw one and
  // the count (one increment per launch, by the first workgroup to see it).
  unsigned long long ovl_last, ovl_steps;
};

// Per-workgroup counters of one force launch.
struct BlockStats {
  unsigned int contacts;       // directed P-P contacts (with multiplicity)
  unsigned int wall_contacts;  // particle-triangle contacts
  unsigned int jkr_failures;   // JKR residual rejections
  unsigned int overlap;        // excessive-overlap flag
  unsigned int history;        // physics-fix: directed contacts stored with a history
};

// Tangential history of the physics-fix mode, one table per slot: the
// loaded directed contacts of the slot's particle, keyed by the partner's
// stable id and sorted by it, with the stored displacement.  Entry k of
// slot i: key[k*cap + i], xi[(c*hcap + k)*cap + i] (c = component), count in
// cnt[i].  The key is a stable id, so a table stays valid when a rebuild
// moves its owner to another slot: it moves with the owner.  The force
// kernel reads one copy and writes the other (like Dyn).
struct Hist {
  uint32_t* key;
  double* xi;
  int* cnt;
};

// Everything a kernel needs that is fixed between parameter changes.
// Meshes whose surface moves (rp_engine_set_mesh_motion) at one time.
constexpr int kMaxMoving = 4;

// Surface motion of triangles [lo, hi): the rigid rotation w (rad/s) about
// the point o.  A table of these lives in device memory, read only by wall
// contacts.
struct WallMotion {
  unsigned int lo, hi;
  double o[3], w[3];
};

struct KParams {
  double dt;
  double bmin[3], bmax[3], bsize[3];
  int periodic[3];
  int any_periodic;
  int normal, tangential;
  double g[3];
  MaterialConsts mc;
  double margin_factor;  // 1.5 for JKR/sJKR (r_sum + 0.5 r_sum), else 1.0
  // reference grid (grid.rs:12-22, cell 1.0 m at simulation.rs:66)
  double ref_origin[3];
  double ref_cs;
  double ref_inv_cs;  // 1 / ref_cs (rounded): fast path of ref_coord_fast
  int ref_dims[3];
  // fine grid
  double go[3];      // origin
  double ginv[3];    // 1 / cell size
  int gn[3];         // cells per axis
  int gstride[3];    // row-major strides (host)
  // cell number of fine coordinate c on axis k:
  // (c >> gsh[k]) * ghi[k] + (c & (2^gsh[k] - 1)) * glo[k]  (see cell_linear)
  int gsh[3], ghi[3], glo[3];
  unsigned int ncells;      // fine cells of one instance
  int cell_reach;           // list build tests each cell's reach (cells below 3/4 of the largest cutoff)
  unsigned int ncells_all;  // ncells * ninst: sort buckets (instance-major)
  int ninst;
  const InstConsts* itab;   // [ninst] per-instance materials (batched engines)
  double skin;
  double half_skin;
  double lookahead_steps;  // steps between rebuild checks (see moved_too_far)
  double safe_half_skin;   // skin/2 minus the float rounding of the stored build positions
  double r_max;
  // JKR: every contact with overlap below this (less 1e-8 |overlap|) has no
  // root (jkr_no_root holds for any R* up to r_max / 2, resp. r_max for walls)
  double jkr_far_pp, jkr_far_pw;
  // sJKR: every particle-particle overlap below this has zero force
  // (below the cut-off -sqrt(3 pi^2 gamma^2 R* / E*^2) of any R* up to r_max / 2;
  // 0 when gamma <= 0, where only overlap > 0 has force)
  double sjkr_zero_pp;
  // moving walls: nmov entries of mov (device memory); the geometry itself
  // stays where it is
  int nmov;
  const WallMotion* mov;
  int cap;           // slot capacity (plane stride)
  int nbr_cap;       // neighbour list capacity per particle
  int pair_cap;      // recorded pair entries per particle
  int fix;           // RP_MODE_PHYSICS_FIX
  int hcap;          // history entries per slot (= nbr_cap when a history is kept)
};

// ---- small device helpers ---------------------------------------------------

__device__ inline dv3 ld3(const double* p, int i, int cap) { return mk(p[i], p[i + cap], p[i + 2 * cap]); }
__device__ inline void ldxr(const double* xr, int i, dv3& x, double& r) {
  const double4 q = reinterpret_cast<const double4*>(xr)[i];
  x = mk(q.x, q.y, q.z);
  r = q.w;
}
__device__ inline dv3 ldx(const double* xr, int i) {
  const double4 q = reinterpret_cast<const double4*>(xr)[i];
  return mk(q.x, q.y, q.z);
}
__device__ inline void stxr(double* xr, int i, dv3 x, double r) {
  double4 q;
  q.x = x.x;
  q.y = x.y;
  q.z = x.z;
  q.w = r;
  reinterpret_cast<double4*>(xr)[i] = q;
}
#ifndef RP_STREAM_HINTS
#define RP_STREAM_HINTS 2
#endif
// Streaming hints for the force kernel's own-particle state: it is read once
// and written once per step, so non-temporal accesses keep the caches for
// the neighbours' position records, which many threads gather.
// 0 = none, 1 = non-temporal stores, 2 = also non-temporal loads of the planes.
__device__ inline dv3 ld3s(const double* p, int i, int cap) {
  if (RP_STREAM_HINTS >= 2)
    return mk(__builtin_nontemporal_load(p + i), __builtin_nontemporal_load(p + i + cap),
              __builtin_nontemporal_load(p + i + 2 * cap));
  return mk(p[i], p[i + cap], p[i + 2 * cap]);
}
__device__ inline void st3s(double* p, int i, int cap, dv3 v) {
  if (RP_STREAM_HINTS >= 1) {
    __builtin_nontemporal_store(v.x, p + i);
    __builtin_nontemporal_store(v.y, p + i + cap);
    __builtin_nontemporal_store(v.z, p + i + 2 * cap);
  } else {
    p[i] = v.x;
    p[i + cap] = v.y;
    p[i + 2 * cap] = v.z;
  }
}

__device__ inline void st3(double* p, int i, int cap, dv3 v) {
  p[i] = v.x;
  p[i + cap] = v.y;
  p[i + 2 * cap] = v.z;
}
__host__ __device__ inline double comp(dv3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// Fine-cell coordinate along one axis: periodic axes wrap, others clamp.
__device__ inline int fine_coord(double x, int k, const KParams& P, bool clamp_only) {
  double q = floor((x - P.go[k]) * P.ginv[k]);
  int n = P.gn[k];
  if (!(q == q)) return 0;
  if (q < -1.0e9) q = -1.0e9;
  if (q > 1.0e9) q = 1.0e9;
  long long c = (long long)q;
  if (P.periodic[k] && !clamp_only) {
    c %= n;
    if (c < 0) c += n;
    return (int)c;
  }
  if (c < 0) return 0;
  if (c >= n) return n - 1;
  return (int)c;
}

// Number of the fine cell (c0, c1, c2), coordinates inside the grid.  Row
// major (gsh = 0, ghi = stride), or in "pencils": 8 x 8 cells of the two
// fast axes, each pencil running the whole slow axis, pencils in row-major
// order -- so a contiguous range of cells is a compact block of space, not a
// full-width layer, and neighbour gathers stay in L2.  Any numbering gives
// the same lists (they are sorted by stable id).
template <int K>
__host__ __device__ inline uint32_t cell_term(int c, const KParams& P) {
  return (uint32_t)((c >> P.gsh[K]) * P.ghi[K] + (c & ((1 << P.gsh[K]) - 1)) * P.glo[K]);
}
__host__ __device__ inline uint32_t cell_linear(const KParams& P, int c0, int c1, int c2) {
  return cell_term<0>(c0, P) + cell_term<1>(c1, P) + cell_term<2>(c2, P);
}

// Reference-grid cell coordinate (grid.rs:84-86): floor((x - origin)/cs) as i32.
// The axis is a template argument so every parameter access has a constant
// index (a runtime index would copy the parameter block to scratch memory).
template <int K>
__device__ inline int ref_coord(double x, const KParams& P) {
  return sat_i32(floor((x - P.ref_origin[K]) / P.ref_cs));
}

// Bit mask of the offsets o in {-1,0,1} (bit o+1) whose wrapped reference
// neighbour cell of ci equals cj (grid.rs:108-129); popcount = how often the
// reference's 27-cell walk reaches cj along this axis.
template <int K>
__device__ inline int ref_mask(int ci, int cj, const KParams& P) {
  const int d = P.ref_dims[K];
  int msk = 0;
#pragma unroll
  for (int o = -1; o <= 1; ++o) {
    int nn = wrap_add_i32(ci, o);
    if (P.periodic[K]) {
      if (nn < 0)
        nn += d;
      else if (nn >= d)
        nn -= d;
    }
    if (nn == cj) msk |= 1 << (o + 1);
  }
  return msk;
}

// The same cell without a division where that is exact: q = (x - origin) *
// RN(1/cs) lies within 4e-16 |q| of the rounded quotient, so when q is more
// than 1e-9 from an integer (and |q| < 1e6) both have the same floor.
template <int K>
__device__ inline int ref_coord_fast(double x, const KParams& P) {
  const double a = x - P.ref_origin[K];
  const double q = a * P.ref_inv_cs;
  const double f = floor(q);
  const double fr = q - f;
  if (fabs(q) < 1.0e6 && fr > 1.0e-9 && fr < 1.0 - 1.0e-9) return (int)f;
  return sat_i32(floor(a / P.ref_cs));
}

// ref_mask by the difference cj - ci when ci lies inside the grid (cj always
// does): offset o reaches ci + o, wrapped once on a periodic axis.
template <int K>
__device__ inline int ref_mask_fast(int ci, int cj, const KParams& P) {
  const int d = P.ref_dims[K];
  if (ci < 0 || ci >= d) return ref_mask<K>(ci, cj, P);
  const int diff = cj - ci;
  int m = (diff == -1 ? 1 : 0) | (diff == 0 ? 2 : 0) | (diff == 1 ? 4 : 0);
  if (P.periodic[K]) m |= (diff == d - 1 ? 1 : 0) | (diff == 1 - d ? 4 : 0);
  return m;
}

// (x - origin) * RN(1/cs): the scaled coordinate ref_coord_fast floors.
template <int K>
__device__ inline double ref_scaled(double x, const KParams& P) {
  return (x - P.ref_origin[K]) * P.ref_inv_cs;
}

// True when reference cell c along axis K lies inside the grid, the axis has
// at least 2 cells and |c| is small enough for in_ref_cell's bound.
template <int K>
__device__ inline bool ref_window_ok(int c, const KParams& P) {
  return c >= 0 && c < P.ref_dims[K] && P.ref_dims[K] >= 2 && c < 999990;
}

// True when x certainly lies in reference cell c along axis K (c passed
// ref_window_ok): t = ref_scaled - c is within 2e-9 of neither end of
// [0, 1].  ref_scaled is within 4e-10 of the rounded quotient the reference
// floors (|q| < 1e6, see ref_coord_fast) and the subtraction adds ~1e-16,
// so that floor is c.  A neighbour in the owner's cell on all three axes is
// visible, and the 27-cell walk reaches it exactly once (offset 0; with at
// least 2 cells per axis the offsets -1 and +1 lead to other cells, also
// after a periodic wrap).
template <int K>
__device__ inline bool in_ref_cell(double x, int c, const KParams& P) {
  const double t = ref_scaled<K>(x, P) - (double)c;
  return t >= 2.0e-9 && t <= 1.0 - 2.0e-9;
}

// XCD-aware block index: consecutive logical blocks land on the same XCD
// (hardware dispatches block b to XCD b % 8), so each XCD's L2 sees a
// contiguous range of cell-sorted particles and their neighbours.
__device__ inline int xcd_block(int b, int nb) {
  int q = nb / 8, r = nb % 8;
  int x = b % 8, k = b / 8;
  return x * q + min(x, r) + k;
}

// JKR past its no-root bound (shortcut and surrogate solvers) and sJKR past
// its cut-off have zero normal force; with Coulomb or the spring law (G* = 0)
// such a contact adds nothing and is only counted, in a pass of its own.
// Mindlin propagates its NaN for negative overlaps, and recording needs every
// pair in order.
// In the physics-fix mode (hist) Mindlin has no NaN: a zero normal force at
// negative overlap gives zero stiffness, no force and no history.
constexpr bool split_pass(int normal, int tang, bool record, int js, bool hist = false) {
  return !record && (tang != RP_TANGENTIAL_MINDLIN || hist) &&
         ((normal == RP_NORMAL_JKR && js >= 1) || normal == RP_NORMAL_SJKR);
}

// Copies slot `from`'s history table of A to slot `to` of B.
__device__ inline void copy_hist(const Hist& A, const Hist& B, int from, int to, const KParams& P) {
  const int c = A.cnt[from];
  B.cnt[to] = c;
  const size_t cap = (size_t)P.cap, st = (size_t)P.hcap * cap;
  for (int k = 0; k < c; ++k) {
    const size_t a = (size_t)k * cap + (size_t)from, b = (size_t)k * cap + (size_t)to;
    B.key[b] = A.key[a];
    B.xi[b] = A.xi[a];
    B.xi[b + st] = A.xi[a + st];
    B.xi[b + 2 * st] = A.xi[a + 2 * st];
  }
}

// Orders the LDS accesses of one wavefront's lanes around this point (each
// wave works in its own LDS slice, so no workgroup barrier is needed).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline unsigned int wave_sum_u32(unsigned int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline int wave_max_i32(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_down(v, off, 64));
  return v;
}

// ---- rebuild chain ----------------------------------------------------------

// Keys + histogram.  Also clears the other parity's rebuild request (it was
// last read one step ago and is next written by this step's force kernel).
__global__ void __launch_bounds__(kBlock) k_keys(Ctl* ctl, int par, KParams P, Dyn S, Stat T, uint32_t* key,
                                                 uint32_t* cell_count) {
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rebuild[par ^ 1] = 0;
  if (!ctl->rebuild[par]) return;
  const int n = ctl->n;
  const int cap = P.cap;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t c;
    if (T.flags[i] & FLAG_DEAD) {
      c = P.ncells_all;
    } else {
      const dv3 x = ldx(S.xr, i);
      int cx = fine_coord(x.x, 0, P, false);
      int cy = fine_coord(x.y, 1, P, false);
      int cz = fine_coord(x.z, 2, P, false);
      c = (uint32_t)T.inst[i] * P.ncells + cell_linear(P, cx, cy, cz);
    }
    key[i] = c;
    atomicAdd(&cell_count[c], 1u);
  }
}

// The kScanItems counts from `base` (a multiple of kScanItems, so 64-byte
// aligned): four 16-byte loads when all lie below m, zeros past m.
__device__ inline void load_items(const uint32_t* in, int base, int m, uint32_t* vals) {
  if (base + kScanItems <= m) {
    const uint4* p4 = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
    for (int q = 0; q < kScanItems / 4; ++q) {
      const uint4 u = p4[q];
      vals[4 * q] = u.x;
      vals[4 * q + 1] = u.y;
      vals[4 * q + 2] = u.z;
      vals[4 * q + 3] = u.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) vals[k] = base + k < m ? in[base + k] : 0u;
  }
}

// Exclusive scan of cell_count[0..m) into cell_start[0..m], cell_start[m] = total.
__global__ void __launch_bounds__(kBlock) k_scan_sums(const Ctl* ctl, int par, const uint32_t* in, int m,
                                                      uint32_t* block_sums) {
  if (!ctl->rebuild[par]) return;
  __shared__ uint32_t red[kBlock];
  for (int b = blockIdx.x; b * kScanChunk < m; b += gridDim.x) {
    // thread t sums the kScanItems items from b*chunk + t*items (64 bytes,
    // four 16-byte loads), the same items k_scan_apply gives it
    const int base = b * kScanChunk + threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    load_items(in, base, m, vals);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += vals[k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
      if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[b] = red[0];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024) k_scan_top(const Ctl* ctl, int par, uint32_t* block_sums, int nb) {
  if (!ctl->rebuild[par]) return;
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += 1024) {
    int idx = base + threadIdx.x;
    uint32_t v = idx < nb ? block_sums[idx] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      uint32_t t = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    uint32_t incl = buf[threadIdx.x];
    if (idx < nb) block_sums[idx] = carry + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += incl;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_scan_apply(const Ctl* ctl, int par, const uint32_t* in, int m,
                                                       const uint32_t* block_offsets, uint32_t* out) {
  if (!ctl->rebuild[par]) return;
  __shared__ uint32_t tsum[kBlock];
  for (int b = blockIdx.x; b * kScanChunk < m; b += gridDim.x) {
    // thread t owns items [b*chunk + t*items, +items): contiguous per thread
    const int base = b * kScanChunk + threadIdx.x * kScanItems;
    uint32_t vals[kScanItems];
    load_items(in, base, m, vals);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) s += vals[k];
    tsum[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
      uint32_t t = threadIdx.x >= off ? tsum[threadIdx.x - off] : 0;
      __syncthreads();
      tsum[threadIdx.x] += t;
      __syncthreads();
    }
    uint32_t run = block_offsets[b] + tsum[threadIdx.x] - s;
    if (base + kScanItems <= m) {
      uint32_t ex[kScanItems];
#pragma unroll
      for (int k = 0; k < kScanItems; ++k) {
        ex[k] = run;
        run += vals[k];
      }
      uint4* o4 = reinterpret_cast<uint4*>(out + base);
#pragma unroll
      for (int q = 0; q < kScanItems / 4; ++q) o4[q] = make_uint4(ex[4 * q], ex[4 * q + 1], ex[4 * q + 2], ex[4 * q + 3]);
      if (base + kScanItems == m) out[m] = run;
    } else {
      for (int k = 0; k < kScanItems; ++k) {
        int idx = base + k;
        if (idx < m) out[idx] = run;
        if (idx == m - 1) out[m] = run + vals[k];
        run += vals[k];
      }
    }
    __syncthreads();
  }
}

// Scatter into cell ranges; the decrementing cursor leaves cell_count zeroed
// for the next rebuild.
__global__ void __launch_bounds__(kBlock) k_scatter(const Ctl* ctl, int par, const uint32_t* key,
                                                    const uint32_t* cell_start, uint32_t* cell_count,
                                                    uint32_t* tmp) {
  if (!ctl->rebuild[par]) return;
  const int n = ctl->n;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t c = key[i];
    uint32_t pos = cell_start[c] + atomicSub(&cell_count[c], 1u) - 1u;
    tmp[pos] = (uint32_t)i;
  }
}

// Deterministic slot = cell start + rank by stable id among the cell's
// particles; copies the state into the scratch buffers (D, U).  Removed
// particles (sentinel cell) are dropped.
__global__ void __launch_bounds__(kBlock) k_rank_permute(Ctl* ctl, int par, KParams P, Dyn S, Stat T, Dyn D,
                                                         Stat U, const uint32_t* key, const uint32_t* cell_start,
                                                         const uint32_t* tmp, Hist HS, Hist HD) {
  if (!ctl->rebuild[par]) return;
  const int n = ctl->n;
  const int cap = P.cap;
  // a and tau are read only by the first half of the next call (k_begin
  // clears at_rest once it has used them): moved only while they are live
  const bool keep_at = ctl->at_rest != 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t c = key[i];
    if (c >= P.ncells_all) continue;
    uint32_t s0 = cell_start[c], s1 = cell_start[c + 1];
    uint32_t my = T.sid[i];
    uint32_t rank = 0;
    for (uint32_t p = s0; p < s1; ++p) rank += (T.sid[tmp[p]] < my) ? 1u : 0u;
    int d = (int)(s0 + rank);
    for (int k = 0; k < 3; ++k) {
      int o = k * cap;
      D.v[d + o] = S.v[i + o];
      D.w[d + o] = S.w[i + o];
      D.rx[d + o] = S.rx[i + o];
      D.rv[d + o] = S.rv[i + o];
      D.rw[d + o] = S.rw[i + o];
      if (keep_at) {
        D.a[d + o] = S.a[i + o];
        D.tau[d + o] = S.tau[i + o];
      }
    }
    reinterpret_cast<double4*>(D.xr)[d] = reinterpret_cast<const double4*>(S.xr)[i];
    U.r[d] = T.r[i];
    U.m[d] = T.m[i];
    U.flags[d] = T.flags[i];
    U.sid[d] = my;
    U.inst[d] = T.inst[i];
    if (HS.cnt) copy_hist(HS, HD, i, d, P);  // the history table moves with its owner
  }
}

// Overlap below which a particle-particle contact of a particle of radius r
// has zero force in every law the zero-force pass handles, -inf for the
// other laws (and for JKR without surface energy).  Its contacts have
// R* = r r_j / (r + r_j) <= rs = r r_max / (r + r_max), and both thresholds
// fall as R* grows -- JKR's no-root bound t d_min - 1e-3 R* - 1e-8 (R* + t)
// with t = R*^(1/3) jc^(2/3) (jkr_no_root), sJKR's cut-off
// -sqrt(3 pi^2 gamma^2 R* / E*^2) (physics.rs:256-260) -- so the value at rs,
// pushed down by 1e-12 (JKR) or 1e-9 (sJKR) of itself for rounding, bounds
// every contact of the particle: a contact below it has zero force.
__device__ inline double zero_threshold_r(const KParams& P, const PairConsts& c, double r, int normal) {
  const double rs = r * P.r_max / (r + P.r_max);
  if (normal == RP_NORMAL_JKR) {
    if (!(c.gamma > 0.0)) return -HUGE_VAL;
    const double t = cbrt(rs) * c.jc23;
    return (kJkrDMin * t - 1e-3 * rs - 1e-8 * (rs + t)) * (1.0 + 1e-12);
  }
  if (normal == RP_NORMAL_SJKR) {
    if (!(c.gamma > 0.0)) return 0.0;
    return -sqrt(c.sterm * rs / c.estar_sq) * (1.0 + 1e-9);
  }
  return -HUGE_VAL;
}

// The threshold of zero_threshold_r for slot i of the re-ordered buffers.
__device__ inline double zero_threshold(const KParams& P, const Stat& U, int i, double r) {
  const InstConsts* ic = P.ninst > 1 ? P.itab + U.inst[i] : nullptr;
  return zero_threshold_r(P, ic ? ic->mc.pp : P.mc.pp, r, P.normal);
}

// Largest radius among each cell's particles (0 for an empty cell), rounded
// up to float, for the list build's per-cell reach test.
__global__ void __launch_bounds__(kBlock) k_cell_rmax(const Ctl* ctl, int par, KParams P, Dyn D,
                                                      const uint32_t* cell_start, float* cell_rmax) {
  if (!ctl->rebuild[par]) return;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < P.ncells_all; c += gridDim.x * blockDim.x) {
    double m = 0.0;
    for (uint32_t j = cell_start[c], j1 = cell_start[c + 1]; j < j1; ++j)
      m = fmax(m, reinterpret_cast<const double4*>(D.xr)[j].w);
    float f = (float)m;
    if ((double)f < m) f = nextafterf(f, HUGE_VALF);
    cell_rmax[c] = f;
  }
}

// Where slot i sits in its cell along each axis, for the distance from it to
// the cells of its search box: f = fraction of the cell below it, w = cell
// width; use = false on an axis where that distance is not bounded this way
// (the box covers the whole periodic axis, or the particle lies outside the
// grid and was clamped into an edge cell).
struct CellReach {
  double f[3], w[3];
  bool use[3];
};

// Lower bound, in cells, on the distance along one axis from the particle to
// any point of the cell `o` cells away (its fraction f), less a rounding margin.
__device__ inline double cell_gap(int o, double f) {
  const double g = o > 0 ? (double)o - f : (o < 0 ? (double)(-o - 1) + f : 0.0);
  return fmax(g - 1e-9, 0.0);
}

// Candidates of slot i in the slots [s0, s1) of the re-ordered scratch
// buffer D (one or more cells with consecutive numbers): appends every j with
// |x_i - x_j|_MIC < margin_factor*(r_i + r_j) + skin to i's list (the radius
// comes with the position record); loads issued four at a time.
__device__ inline int build_range(const KParams& P, const Dyn& D, uint32_t s0, uint32_t s1, int i, dv3 xi,
                                  double ri, double zthr, int* nbr, int cnt) {
  for (uint32_t j0 = s0; j0 < s1; j0 += 4) {
    dv3 xq[4];
    double rq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) ldxr(D.xr, (int)min(j0 + (uint32_t)q, s1 - 1u), xq[q], rq[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t j = j0 + (uint32_t)q;
      if (j >= s1 || (int)j == i) continue;
      dv3 d = sub(xi, xq[q]);
      if (P.periodic[0]) {
        if (d.x > P.bsize[0] * 0.5) d.x -= P.bsize[0];
        else if (d.x < -P.bsize[0] * 0.5) d.x += P.bsize[0];
      }
      if (P.periodic[1]) {
        if (d.y > P.bsize[1] * 0.5) d.y -= P.bsize[1];
        else if (d.y < -P.bsize[1] * 0.5) d.y += P.bsize[1];
      }
      if (P.periodic[2]) {
        if (d.z > P.bsize[2] * 0.5) d.z -= P.bsize[2];
        else if (d.z < -P.bsize[2] * 0.5) d.z += P.bsize[2];
      }
      const double lim = P.margin_factor * (ri + rq[q]) + P.skin;
      const double dsq = dot(d, d);
      if (dsq < lim * lim) {
        // far (bit 31): beyond the zero-force distance with rounding to
        // spare by more than the skin, so no step before the next rebuild
        // can give the pair a force (zero_far)
        const double lz = (ri + rq[q] - zthr) * (1.0 + 1e-6) + P.skin;
        const bool far = dsq > lz * lz * (1.0 + 1e-9);
        if (cnt < P.nbr_cap) nbr[cnt * P.cap + i] = (int)(j | (far ? 0x80000000u : 0u));
        ++cnt;
      }
    }
  }
  return cnt;
}

// Cell coordinate c of the search box along axis K, wrapped on a periodic
// axis the box does not cover whole.
template <int K>
__device__ inline int box_wrap(int c, const KParams& P, const bool* all) {
  return (P.periodic[K] && !all[K]) ? (c % P.gn[K] + P.gn[K]) % P.gn[K] : c;
}

// The search box lo..hi of slot i, axis A0 innermost: consecutive cell
// numbers along a row are merged into one slot range (two cell_start loads
// per range instead of per cell).  The set of candidates is the same in any
// walk order; the list is sorted by stable id afterwards.
// A cell is skipped when its nearest point is farther from the particle than
// margin_factor*(r_i + the cell's largest radius) + skin, with rounding to
// spare: no particle in it can pass build_range's test.  With a wide radius
// distribution (cells sized for the largest pairs) most of the box is out of
// reach of a small particle; empty cells cost no range loads either.
template <int A0, int A1, int A2, bool REACH>
__device__ inline int build_walk(const KParams& P, const Dyn& D, const uint32_t* cell_start, const float* cell_rmax,
                                 uint32_t ibase, int i, dv3 xi, double ri, double zthr, const int* ci, const int* lo,
                                 const int* hi, const bool* all, const CellReach& G, double lim2_max, int* nbr) {
  int cnt = 0;
  for (int c2 = lo[A2]; c2 <= hi[A2]; ++c2) {
    const uint32_t t2 = ibase + cell_term<A2>(box_wrap<A2>(c2, P, all), P);
    const double g2 = REACH && G.use[A2] ? cell_gap(c2 - ci[A2], G.f[A2]) * G.w[A2] : 0.0;
    for (int c1 = lo[A1]; c1 <= hi[A1]; ++c1) {
      const double g1 = REACH && G.use[A1] ? cell_gap(c1 - ci[A1], G.f[A1]) * G.w[A1] : 0.0;
      const double g12 = g2 * g2 + g1 * g1;
      if (REACH && g12 > lim2_max) continue;  // the whole row is out of reach
      const uint32_t t21 = t2 + cell_term<A1>(box_wrap<A1>(c1, P, all), P);
      uint32_t r0 = 0, r1 = 0;  // current run of cells [r0, r1)
      for (int c0 = lo[A0]; c0 <= hi[A0]; ++c0) {
        const uint32_t c = t21 + cell_term<A0>(box_wrap<A0>(c0, P, all), P);
        if constexpr (REACH) {
          const double g0 = G.use[A0] ? cell_gap(c0 - ci[A0], G.f[A0]) * G.w[A0] : 0.0;
          const double d2 = g12 + g0 * g0;
          bool take = d2 <= lim2_max;
          if (take && d2 > 0.0) {
            const double l = (P.margin_factor * (ri + (double)cell_rmax[c]) + P.skin) * (1.0 + 1e-6);
            take = d2 <= l * l;
          }
          if (!take) continue;
        }
        if (r1 > r0 && c == r1) {
          r1 = c + 1u;
          continue;
        }
        if (r1 > r0) cnt = build_range(P, D, cell_start[r0], cell_start[r1], i, xi, ri, zthr, nbr, cnt);
        r0 = c;
        r1 = c + 1u;
      }
      if (r1 > r0) cnt = build_range(P, D, cell_start[r0], cell_start[r1], i, xi, ri, zthr, nbr, cnt);
    }
  }
  return cnt;
}

// Verlet list of slot i (from the re-ordered scratch buffers D, U): every j
// with |x_i - x_j|_MIC < margin_factor*(r_i + r_j) + skin; then copies slot
// i back into the working buffers (W, V) and records the new slot count.
// Four waves per SIMD: the per-cell reach test takes the kernel to 131 VGPRs
// (three waves); capped at 128 it spills 12 B and builds the 2M JKR packing's
// lists in 5.8 ms instead of 6.9 ms.
template <bool REACH>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_build(Ctl* ctl, int par, KParams P, Dyn D, Stat U, Dyn W, Stat V,
                                                  const uint32_t* cell_start, const float* cell_rmax, int* nbr_count,
                                                  int* nbr, float* xb, Hist HD, Hist HW) {
  if (!ctl->rebuild[par]) return;
  const int n = (int)cell_start[P.ncells_all];  // alive particles after compaction
  const int cap = P.cap;
  const bool keep_at = ctl->at_rest != 0;
  int local_max = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    dv3 xi;
    double ri;
    ldxr(D.xr, i, xi, ri);
    const uint32_t ibase = (uint32_t)U.inst[i] * P.ncells;  // candidates come from i's own instance only
    int ci[3] = {fine_coord(xi.x, 0, P, false), fine_coord(xi.y, 1, P, false), fine_coord(xi.z, 2, P, false)};
    double lim_max = P.margin_factor * (ri + P.r_max) + P.skin;
    int lo[3], hi[3];
    bool all[3];
    for (int k = 0; k < 3; ++k) {
      int kk = (int)ceil(lim_max * P.ginv[k]);
      if (kk < 1) kk = 1;
      if (P.periodic[k]) {
        all[k] = (2 * kk + 1 >= P.gn[k]);
        lo[k] = all[k] ? 0 : ci[k] - kk;
        hi[k] = all[k] ? P.gn[k] - 1 : ci[k] + kk;
      } else {
        all[k] = false;
        lo[k] = max(0, ci[k] - kk);
        hi[k] = min(P.gn[k] - 1, ci[k] + kk);
      }
    }
    CellReach G{};
    if constexpr (REACH) {
      const double xk[3] = {xi.x, xi.y, xi.z};
      for (int k = 0; k < 3; ++k) {
        const double q = (xk[k] - P.go[k]) * P.ginv[k];  // fine_coord's scaled coordinate
        const double fl = floor(q);
        G.f[k] = q - fl;
        G.w[k] = 1.0 / P.ginv[k];
        G.use[k] = !all[k] && fabs(q) < 1.0e9 && (P.periodic[k] || (fl >= 0.0 && fl < (double)P.gn[k]));
      }
    }
    const double lm = lim_max * (1.0 + 1e-6);
    const double lim2_max = lm * lm;
    int cnt = 0;
    const double zthr = zero_threshold(P, U, i, ri);
    // the grid axis whose neighbouring cells have consecutive numbers walks
    // innermost, so each row of the search box is one or a few slot ranges
    if (P.gstride[0] == 1)
      cnt = build_walk<0, 1, 2, REACH>(P, D, cell_start, cell_rmax, ibase, i, xi, ri, zthr, ci, lo, hi, all, G, lim2_max,
                                nbr);
    else if (P.gstride[1] == 1)
      cnt = build_walk<1, 0, 2, REACH>(P, D, cell_start, cell_rmax, ibase, i, xi, ri, zthr, ci, lo, hi, all, G, lim2_max,
                                nbr);
    else
      cnt = build_walk<2, 0, 1, REACH>(P, D, cell_start, cell_rmax, ibase, i, xi, ri, zthr, ci, lo, hi, all, G, lim2_max,
                                nbr);
    const int kept = min(cnt, P.nbr_cap);
    // Near entries first, each part by ascending stable id = ascending API
    // index: the reference's summation order inside one reference cell
    // (grid.rs:42-46 inserts in index order) for every entry that can carry
    // a force; the far part only ever adds zeros, in any order.
    auto key = [&](int e) {
      return ((unsigned long long)((uint32_t)e >> 31) << 32) | (unsigned long long)U.sid[e & 0x7fffffff];
    };
    for (int a = 1; a < kept; ++a) {
      int ja = nbr[a * cap + i];
      const unsigned long long ka = key(ja);
      int b = a - 1;
      while (b >= 0) {
        int jb = nbr[b * cap + i];
        if (key(jb) <= ka) break;
        nbr[(b + 1) * cap + i] = jb;
        --b;
      }
      nbr[(b + 1) * cap + i] = ja;
    }
    int near = kept;
    for (int a = kept - 1; a >= 0 && nbr[a * cap + i] < 0; --a) {
      nbr[a * cap + i] &= 0x7fffffff;
      near = a;
    }
    nbr_count[i] = kept;
    nbr_count[cap + i] = near;
    local_max = max(local_max, cnt);
    xb[i] = (float)xi.x;
    xb[i + cap] = (float)xi.y;
    xb[i + 2 * cap] = (float)xi.z;
    for (int k = 0; k < 3; ++k) {
      const int o = i + k * cap;
      W.v[o] = D.v[o];
      W.w[o] = D.w[o];
      W.rx[o] = D.rx[o];
      W.rv[o] = D.rv[o];
      W.rw[o] = D.rw[o];
      if (keep_at) {  // live only in the at-rest state (see k_rank_permute)
        W.a[o] = D.a[o];
        W.tau[o] = D.tau[o];
      }
    }
    reinterpret_cast<double4*>(W.xr)[i] = reinterpret_cast<const double4*>(D.xr)[i];
    V.r[i] = U.r[i];
    V.m[i] = U.m[i];
    V.flags[i] = U.flags[i];
    V.sid[i] = U.sid[i];
    V.inst[i] = U.inst[i];
    if (HW.cnt) copy_hist(HD, HW, i, i, P);
  }
  const int wmax = wave_max_i32(local_max);
  if ((threadIdx.x & 63) == 0 && wmax > 0) {
    atomicMax(&ctl->max_count, wmax);
    if (wmax > P.nbr_cap) atomicOr(&ctl->overflow, 1);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->n = n;
    ctl->rebuilds += 1;
    ctl->stall_off = kNoStall;
  }
}

// ---- per-step kernels ---------------------------------------------------------

// Start of step s+1 for one particle, from its at-rest state after step s:
// periodic wrap (simulation.rs:182-197), removal test (:200-214), first half
// kick, drift and angular half kick (:235-263).  Returns false when removed.
// alpha = tau / I with I = 0.4 m r^2 (simulation.rs:253-254), computed by the
// caller (the second half kick of the previous step uses the same value).
__device__ inline bool begin_step(const KParams& P, dv3& x, dv3& v, dv3& w, dv3& rx, dv3& rv, dv3& rw, dv3 a,
                                  dv3 alpha, uint8_t flags) {
  if (P.any_periodic) {
    if (P.periodic[0]) {
      if (x.x < P.bmin[0]) x.x += P.bsize[0];
      else if (x.x >= P.bmax[0]) x.x -= P.bsize[0];
    }
    if (P.periodic[1]) {
      if (x.y < P.bmin[1]) x.y += P.bsize[1];
      else if (x.y >= P.bmax[1]) x.y -= P.bsize[1];
    }
    if (P.periodic[2]) {
      if (x.z < P.bmin[2]) x.z += P.bsize[2];
      else if (x.z >= P.bmax[2]) x.z -= P.bsize[2];
    }
  }
  bool keep = (P.periodic[0] || (x.x >= P.bmin[0] - 1.0 && x.x <= P.bmax[0] + 1.0)) &&
              (P.periodic[1] || (x.y >= P.bmin[1] - 1.0)) &&
              (P.periodic[2] || (x.z >= P.bmin[2] - 1.0 && x.z <= P.bmax[2] + 1.0));
  if (!keep) return false;
  if (!(flags & FLAG_FIXED)) {
    const double dt = P.dt;
    kahan1(v.x, rv.x, (0.5 * a.x) * dt);
    kahan1(v.y, rv.y, (0.5 * a.y) * dt);
    kahan1(v.z, rv.z, (0.5 * a.z) * dt);
    kahan1(x.x, rx.x, v.x * dt);
    kahan1(x.y, rx.y, v.y * dt);
    kahan1(x.z, rx.z, v.z * dt);
    kahan1(w.x, rw.x, (0.5 * alpha.x) * dt);
    kahan1(w.y, rw.y, (0.5 * alpha.y) * dt);
    kahan1(w.z, rw.z, (0.5 * alpha.z) * dt);
  }
  return true;
}

__device__ inline dv3 angular_acc(dv3 tau, double r, double m) {
  const double inertia = 0.4 * m * r * r;
  return mk(tau.x / inertia, tau.y / inertia, tau.z / inertia);
}

// Displacement since the last list build, minimum image on periodic axes.
template <int K>
__device__ inline double mic_disp(double d, const KParams& P) {
  if (P.periodic[K]) {
    if (d > P.bsize[K] * 0.5) d -= P.bsize[K];
    else if (d < -P.bsize[K] * 0.5) d += P.bsize[K];
  }
  return d;
}

// The Verlet list stays a superset of the contact pairs while every particle
// is within skin/2 of its position at the list build.  The host launches the
// rebuild chain only every `lookahead_steps` steps, so a particle asks for a
// rebuild once its displacement plus a bound on its motion over that many
// steps (2 K |v| dt + K^2 |a| dt^2) reaches skin/2.  A particle already past
// skin/2 at the position of step `next_step` (its label offset) stalls that
// step (stall_off); the step runs again after the rebuild.
__device__ inline bool moved_too_far(const KParams& P, dv3 x, dv3 v, dv3 a, const float* xb, int i, Ctl* ctl,
                                     unsigned long long next_step) {
  const double d0 = mic_disp<0>(x.x - (double)xb[i], P);
  const double d1 = mic_disp<1>(x.y - (double)xb[i + P.cap], P);
  const double d2 = mic_disp<2>(x.z - (double)xb[i + 2 * P.cap], P);
  const double dd = d0 * d0 + d1 * d1 + d2 * d2;
  if (!(dd < P.safe_half_skin * P.safe_half_skin)) {
    atomicMin(&ctl->stall_off, (long long)next_step);
    return true;
  }
  const double k = P.lookahead_steps, dt = P.dt;
  const double ahead = 2.0 * k * sqrt(dot(v, v)) * dt + k * k * sqrt(dot(a, a)) * dt * dt;
  return !(sqrt(dd) + ahead < P.safe_half_skin);
}

// A removed particle keeps slot i until the next rebuild drops it; its
// position becomes NaN in the buffer written now, so every later pair test
// against it is false.
__device__ inline void mark_removed(Ctl* ctl, const Stat& T, const Dyn& D, int i, int cap, uint8_t flags) {
  T.flags[i] = flags | FLAG_DEAD;
  const double qnan = __builtin_nan("");
  stxr(D.xr, i, mk(qnan, qnan, qnan), qnan);
  atomicSub(&ctl->alive, 1);
}

// First half of a step from the at-rest state (used once at the start of
// each rp_engine_step call).
// flag_slot: the rebuild-request slot read by the next rebuild chain.
__global__ void __launch_bounds__(kBlock) k_begin(Ctl* ctl, int flag_slot, KParams P, Dyn S, Stat T,
                                                  const float* xb, unsigned long long off) {
  const int n = ctl->n;
  const int cap = P.cap;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool far = false;
  if (i < n) {
    uint8_t flags = T.flags[i];
    if (!(flags & FLAG_DEAD)) {
      dv3 x;
      double r;
      ldxr(S.xr, i, x, r);
      dv3 v = ld3(S.v, i, cap), w = ld3(S.w, i, cap);
      dv3 rx = ld3(S.rx, i, cap), rv = ld3(S.rv, i, cap), rw = ld3(S.rw, i, cap);
      dv3 a = ld3(S.a, i, cap), tau = ld3(S.tau, i, cap);
      if (!begin_step(P, x, v, w, rx, rv, rw, a, angular_acc(tau, T.r[i], T.m[i]), flags)) {
        mark_removed(ctl, T, S, i, cap, flags);
      } else {
        stxr(S.xr, i, x, r);
        st3(S.v, i, cap, v);
        st3(S.w, i, cap, w);
        st3(S.rx, i, cap, rx);
        st3(S.rv, i, cap, rv);
        st3(S.rw, i, cap, rw);
        far = moved_too_far(P, x, v, a, xb, i, ctl, off);
      }
    }
  }
  if (__any(far) && (threadIdx.x & 63) == 0) atomicOr(&ctl->rebuild[flag_slot], 1);
  if (blockIdx.x == 0 && threadIdx.x == 0) ctl->at_rest = 0;
}

// Pair test of slot i against slot j (simulation.rs:323-357): minimum image
// per periodic axis, then dist^2 < (r_sum + margin)^2.
__device__ inline dv3 pair_delta(const KParams& P, dv3 xi, dv3 xj) {
  dv3 d = sub(xi, xj);
  if (P.periodic[0]) {
    if (d.x > P.bsize[0] * 0.5) d.x -= P.bsize[0];
    else if (d.x < -P.bsize[0] * 0.5) d.x += P.bsize[0];
  }
  if (P.periodic[1]) {
    if (d.y > P.bsize[1] * 0.5) d.y -= P.bsize[1];
    else if (d.y < -P.bsize[1] * 0.5) d.y += P.bsize[1];
  }
  if (P.periodic[2]) {
    if (d.z > P.bsize[2] * 0.5) d.z -= P.bsize[2];
    else if (d.z < -P.bsize[2] * 0.5) d.z += P.bsize[2];
  }
  return d;
}

__device__ inline bool in_contact(const KParams& P, double dsq, double rsum) {
  const bool adhesive = P.normal == RP_NORMAL_JKR || P.normal == RP_NORMAL_SJKR;
  const double margin = adhesive ? rsum * 0.5 : 0.0;
  const double cutoff = rsum + margin;
  return dsq < cutoff * cutoff;
}

struct Own {
  dv3 x, v, w;
  double r, m;
  int rc0, rc1, rc2;  // reference-grid cell of this particle
};

struct Acc {
  dv3 f, t;
  unsigned int ncont, njkr;
  int npair;
  bool overlap;
  // JKR surrogate: first feature row of this particle (-1 = none), the list
  // entries that have rows, the entries near the served set (next step's
  // collect visits only these), and counters
  int srow;
  int sok;  // the particle's material pair lies in the surrogate's envelope
  unsigned long long rmask, nmask;
  unsigned int sserved, sfall;
  float emax, rmax;
  double esum, rsum;
  // physics-fix history: cursor and size of the old table, entries written
  int hp, hold, hn;
};

// JKR normal force with the exact no-root shortcut (JS >= 1) and the network
// (JS >= 2; JS == 3 also runs the exact solve and records the force errors).
// JS == 0 is the reference's solve as written.  The shortcut and
// jkr() agree exactly (zero force, a rejected root); the network path polishes
// the predicted radius to the root the reference converges to.
template <int JS>
__device__ inline dv3 jkr_fast(double ov, dv3 n, const PairConsts& c, double rstar, int* failed, int k,
                               const SurrDev& sd, Acc& acc, double far) {
  if (JS == 0 || !(c.gamma > 0.0)) return jkr(ov, n, c, rstar, failed);
  if (ov < far - 1e-8 * fabs(ov)) {
    *failed = 1;
    return zero3();
  }
  // A compressed contact always has a root (the bound below is negative), so
  // the shortcut solver skips the cube root there.  Every path ends in the
  // one jkr() call below, so a wavefront runs the Newton loop once.
  if (JS >= 2 || ov < 0.0) {
    const double s = cbrt(rstar * rstar * c.jc);
    const double t = s * s / rstar;
    if (ov < kJkrDMin * t - 1e-3 * rstar - 1e-8 * (rstar + t + fabs(ov))) {
      *failed = 1;
      return zero3();
    }
    if (JS >= 2) {
      if (k < 64 && acc.sok && rstar >= sd.r_lo && rstar <= sd.r_hi && s >= kSurrSLo * rstar &&
          s <= kSurrSHi * rstar && ov >= sd.md_lo * t && ov <= sd.md_hi * t)
        acc.nmask |= 1ull << k;
      double d = 0.0;
      if (surr_route(ov, rstar, s, t, sd, k, d, acc.sok)) {
        const unsigned long long bit = 1ull << k;  // k < 64 here
        const int row = (acc.rmask & bit) ? acc.srow + __popcll(acc.rmask & (bit - 1ull)) : -1;
        if (row >= 0 && row < sd.rows_cap) {
          const float y0 = sd.y[2 * (size_t)row], y1 = sd.y[2 * (size_t)row + 1];
          double al = (double)unscale_out(y0, sd.sc.out_lo[0], sd.sc.out_hi[0]);
          if (polish_alpha(d, al)) {
            const dv3 f = jkr_finish(al * s, ov, n, c, rstar, failed);
            ++acc.sserved;
            if (JS == 3) {
              int ef = 0;
              const dv3 fe = jkr(ov, n, c, rstar, &ef);
              const double fpull = c.jpull * rstar;  // -(2/3) pi gamma R*
              const double scale = fmax(len(fe), -fpull);
              const double err = len(sub(f, fe)) / scale;
              double fraw = (double)unscale_out(y1, sd.sc.out_lo[1], sd.sc.out_hi[1]) * (kPi * c.gamma * rstar);
              if (fraw < fpull) fraw = 0.0;
              const double raw = fabs(fraw - dot(fe, n)) / scale;
              acc.emax = fmaxf(acc.emax, (float)err);
              acc.rmax = fmaxf(acc.rmax, (float)raw);
              acc.esum += err;
              acc.rsum += raw;
            }
            return f;
          }
        }
        ++acc.sfall;
      }
    }
  }
  return jkr(ov, n, c, rstar, failed);
}

// Pair entries of one contact in the reference's 27-cell walk order (rank
// oz*9 + oy*3 + ox of each reaching offset) for rp_engine_pairs.
__device__ inline void record_pairs(const KParams& P, const Stat& T, int j, int mask0, int mask1, int mask2, Acc& acc,
                                    Ctl* ctl, unsigned long long* prow) {
  const unsigned long long sj = T.sid[j];
  for (int oz = 0; oz < 3; ++oz) {
    if (!(mask2 >> oz & 1)) continue;
    for (int oy = 0; oy < 3; ++oy) {
      if (!(mask1 >> oy & 1)) continue;
      for (int ox = 0; ox < 3; ++ox) {
        if (!(mask0 >> ox & 1)) continue;
        const unsigned long long rank = (unsigned long long)(oz * 9 + oy * 3 + ox);
        if (acc.npair < P.pair_cap) prow[acc.npair] = (rank << 32) | sj;
        else atomicOr(&ctl->pair_overflow, 1);
        ++acc.npair;
      }
    }
  }
}

// One particle-particle contact of the gather (simulation.rs:357-460), added
// `mult` times when the reference's grid walk reaches j more than once.
// HIST: physics-fix tangential history (tables HR -> HW of slot i).
template <int NORMAL, int TANG, bool RECORD, int JS, bool HIST>
__device__ inline void contact_pp(const KParams& P, const MaterialConsts& mc, double far_pp, const Dyn& S,
                                  const Stat& T, const Own& o, int j, int k, Acc& acc, Ctl* ctl,
                                  unsigned long long* prow, const SurrDev& sd, const Hist& HR, const Hist& HW,
                                  int i) {
  dv3 xj;
  double rj;
  ldxr(S.xr, j, xj, rj);
  const dv3 d = pair_delta(P, o.x, xj);
  const double dsq = dot(d, d);
  const double rsum = o.r + rj;
  // visible to the reference grid, and reached how often by its 27-cell walk
  // (physics-fix mode: every pair once, as the walk's own cell)
  int mask0 = 2, mask1 = 2, mask2 = 2;
  if (!P.fix) {
    const int cj0 = ref_coord_fast<0>(xj.x, P), cj1 = ref_coord_fast<1>(xj.y, P), cj2 = ref_coord_fast<2>(xj.z, P);
    if (cj0 < 0 || cj0 >= P.ref_dims[0] || cj1 < 0 || cj1 >= P.ref_dims[1] || cj2 < 0 || cj2 >= P.ref_dims[2])
      return;
    mask0 = ref_mask_fast<0>(o.rc0, cj0, P);
    mask1 = ref_mask_fast<1>(o.rc1, cj1, P);
    mask2 = ref_mask_fast<2>(o.rc2, cj2, P);
  }
  const int mult = __popc(mask0) * __popc(mask1) * __popc(mask2);
  if (mult == 0) return;
  const double dist = sqrt(dsq);
  const double overlap = rsum - dist;
  if (overlap > 0.2 * fmin(o.r, rj)) acc.overlap = true;
  if (NORMAL == RP_NORMAL_JKR && JS >= 1 && (TANG != RP_TANGENTIAL_MINDLIN || HIST) && mc.pp.gamma > 0.0 &&
      overlap < far_pp - 1e-8 * fabs(overlap)) {
    // no root: zero normal force and a rejected root; Coulomb and the spring
    // law with G* = 0 then give a zero tangential force, so the contact adds
    // nothing (zeros of either sign do not change a sum that starts at +0 and
    // is re-based with 0.0 + in the epilogue) and j's velocities are not read.
    // Mindlin's sqrt(R* delta) is NaN for delta < 0 and must propagate.
    acc.njkr += (unsigned int)mult;
    acc.ncont += (unsigned int)mult;
    if (RECORD) record_pairs(P, T, j, mask0, mask1, mask2, acc, ctl, prow);
    return;
  }
  const dv3 nrm = dist > 1e-6 ? divs(d, dist) : mk(0.0, 1.0, 0.0);
  const dv3 ra = mul(neg(nrm), o.r);
  const dv3 rb = mul(nrm, rj);
  const dv3 vj = ld3(S.v, j, P.cap);
  const dv3 wj = ld3(S.w, j, P.cap);
  const double mj = T.m[j];
  const dv3 vrel = sub(add(o.v, cross(o.w, ra)), add(vj, cross(wj, rb)));
  const double rstar = (o.r * rj) / (o.r + rj);
  const double mstar = (o.m * mj) / (o.m + mj);
  dv3 fn;
  int jfail = 0;
  if (NORMAL == RP_NORMAL_HERTZ) {
    fn = hertz(overlap, nrm, vrel, mc.pp, rstar, mstar, mc.beta_abs);
  } else if (NORMAL == RP_NORMAL_LINEAR) {
    const double kn = mc.pp.estar * rstar;
    const double gn = 0.5 * sqrt(mstar * kn);
    fn = linear_sd(overlap, nrm, vrel, kn, gn);
  } else if (NORMAL == RP_NORMAL_HYSTERETIC) {
    fn = linear_sd(overlap, nrm, zero3(), mc.pp.estar * rstar, 0.0);
  } else if (NORMAL == RP_NORMAL_JKR) {
    fn = jkr_fast<JS>(overlap, nrm, mc.pp, rstar, &jfail, k, sd, acc, far_pp);
  } else {
    fn = sjkr(overlap, nrm, vrel, mc.pp, rstar);
  }
  acc.njkr += (unsigned int)(jfail * mult);
  dv3 fc;
  if (TANG == RP_TANGENTIAL_COULOMB) {
    fc = add(fn, coulomb(len(fn), vrel, nrm, mc.mu));
  } else if constexpr (HIST) {
    const double fnmag = len(fn);
    dv3 ft;
    // the stored displacement of (i, j): the old table is sorted by partner
    // id and the contacts come in ascending id, so a cursor finds it
    const uint32_t sj = T.sid[j];
    const size_t cap = (size_t)P.cap, st = (size_t)P.hcap * cap;
    dv3 xi = zero3();
    while (acc.hp < acc.hold && HR.key[(size_t)acc.hp * cap + i] < sj) ++acc.hp;
    if (acc.hp < acc.hold && HR.key[(size_t)acc.hp * cap + i] == sj) {
      const size_t b = (size_t)acc.hp * cap + i;
      xi = mk(HR.xi[b], HR.xi[b + st], HR.xi[b + 2 * st]);
      ++acc.hp;
    }
    const double kt = TANG == RP_TANGENTIAL_MINDLIN ? mindlin_kt_fix(mc.gstar_pp, rstar, overlap)
                                                    : mc.gstar_pp * rstar;
    ft = spring_coulomb_fix(fnmag, vrel, nrm, xi, kt, mc.mu, P.dt);
    if (kt > 0.0 && (xi.x != 0.0 || xi.y != 0.0 || xi.z != 0.0)) {
      const size_t b = (size_t)acc.hn * cap + i;
      HW.key[b] = sj;
      HW.xi[b] = xi.x;
      HW.xi[b + st] = xi.y;
      HW.xi[b + 2 * st] = xi.z;
      ++acc.hn;
    }
    fc = add(fn, ft);
  } else {
    // Mindlin kt = 8*G*sqrt(R* delta), LSC kt = G*R*, with G* = 0 (material.rs:49).
    // Where that root is finite and the friction coefficient is finite and
    // >= 0, kt = +0 and the law returns (-0, -0, -0), which adds nothing to
    // the normal force (x + -0 == x for every x, zeros included): the sum is
    // the normal force as it stands, and neither root nor |F_n| is needed.
    // Otherwise (a negative overlap's NaN root) the law runs as written.
    const double ro = rstar * overlap;
    const bool inert = mc.mu >= 0.0 && mc.mu < HUGE_VAL &&
                       (TANG == RP_TANGENTIAL_LSC || (ro >= 0.0 && ro < HUGE_VAL));
    if (inert) {
      fc = fn;
    } else {
      const double kt = TANG == RP_TANGENTIAL_MINDLIN ? (8.0 * 0.0) * sqrt(ro) : 0.0 * rstar;
      fc = add(fn, spring_coulomb_g0(len(fn), kt, mc.mu));
    }
  }
  const dv3 tc = cross(ra, fc);
  for (int rep = 0; rep < mult; ++rep) {
    acc.f = add(acc.f, fc);
    acc.t = add(acc.t, tc);
  }
  acc.ncont += (unsigned int)mult;
  if (RECORD) record_pairs(P, T, j, mask0, mask1, mask2, acc, ctl, prow);
}

// Surface velocity of triangle ti at the point c: omega x (c - origin) of
// its mesh's rotation, 0 for a mesh at rest.
__device__ inline dv3 wall_velocity(const KParams& P, unsigned int ti, dv3 c) {
  for (int k = 0; k < P.nmov; ++k) {
    const WallMotion& m = P.mov[k];
    if (ti >= m.lo && ti < m.hi) return cross(mk(m.w[0], m.w[1], m.w[2]), sub(c, mk(m.o[0], m.o[1], m.o[2])));
  }
  return mk(0.0, 0.0, 0.0);
}

// One particle-triangle contact (simulation.rs:465-523).  A moving wall's
// surface velocity at the contact point (the closest point) is subtracted
// from the relative velocity; the reference's walls are at rest.
template <int NORMAL, int JS>
__device__ inline void contact_wall(const KParams& P, const MaterialConsts& mc, double far_pw, const Own& o,
                                    const Tri& tr, unsigned int ti, Acc& acc, unsigned int& nwall,
                                    const SurrDev& sd) {
  // out of reach of the triangle's plane: no closest point can be within r
  if (!(fabs(dot(tr.pn, o.x) - tr.pd) < o.r + tr.ptol)) return;
  const dv3 cp = P.fix ? closest_point_fix(tr, o.x) : closest_point(tr, o.x);
  const dv3 dv = sub(o.x, cp);
  const double dsq = dot(dv, dv);
  if (!(dsq < o.r * o.r)) return;
  const double dist = sqrt(dsq);
  const double pen = o.r - dist;
  const dv3 nrm = dist > 1e-6 ? divs(dv, dist) : tr.n;
  if (pen > 0.2 * o.r) acc.overlap = true;
  const dv3 ra = mul(neg(nrm), o.r);
  dv3 vrel = add(o.v, cross(o.w, ra));
  if (P.nmov) vrel = sub(vrel, wall_velocity(P, ti, cp));
  dv3 fn;
  int jfail = 0;
  if (NORMAL == RP_NORMAL_HERTZ) {
    fn = hertz(pen, nrm, vrel, mc.pw, o.r, o.m, mc.beta_abs);
  } else if (NORMAL == RP_NORMAL_LINEAR) {
    const double kn = mc.pw.estar * o.r;
    const double gn = 0.5 * sqrt(o.m * kn);
    fn = linear_sd(pen, nrm, vrel, kn, gn);
  } else if (NORMAL == RP_NORMAL_HYSTERETIC) {
    fn = linear_sd(pen, nrm, vrel, mc.pw.estar * o.r, 0.0);
  } else if (NORMAL == RP_NORMAL_JKR) {
    // walls take the exact solve (with the no-root shortcut when JS >= 1)
    fn = jkr_fast<JS == 0 ? 0 : 1>(pen, nrm, mc.pw, o.r, &jfail, 64, sd, acc, far_pw);
  } else {
    fn = sjkr(pen, nrm, vrel, mc.pw, o.r);
  }
  acc.njkr += (unsigned int)jfail;
  const double fnmag = len(fn);
  const dv3 ft = coulomb(fnmag, vrel, nrm, mc.mu);
  const dv3 fc = add(fn, ft);
  acc.f = add(acc.f, fc);
  acc.t = add(acc.t, cross(ra, fc));
  ++nwall;
}

__device__ inline int sel4(int q, int a, int b, int c, int d) { return q == 0 ? a : (q == 1 ? b : (q == 2 ? c : d)); }

__device__ inline void copy_slot(const Dyn& S, const Dyn& D, int i, int cap) {
  reinterpret_cast<double4*>(D.xr)[i] = reinterpret_cast<const double4*>(S.xr)[i];
  for (int k = 0; k < 3; ++k) {
    const int o = i + k * cap;
    D.v[o] = S.v[o];
    D.w[o] = S.w[o];
    D.rx[o] = S.rx[o];
    D.rv[o] = S.rv[o];
    D.rw[o] = S.rw[o];
    D.a[o] = S.a[o];
    D.tau[o] = S.tau[o];
  }
}

// Contacts of particle i: its neighbour list (loads issued kNbrBatch at a
// time, so a wave waits for memory a few times per particle instead of once
// per neighbour) and the triangles binned to its cell.
// Gather force kernel + integration epilogue, one thread per particle slot.
// Reads Dyn buffer S and writes buffer D (so no thread can read a neighbour's
// next-step state); the host swaps the two between steps.  All own-state loads are issued before the neighbour loop, and
// neighbours are tested four at a time, so a wave waits for memory a few
// times per particle instead of once per neighbour.
// ADVANCE: also performs the next step's first half (begin_step) and leaves
// the state drifted; otherwise stores a, tau and leaves the at-rest state.
// RECORD: writes the reference-ordered pair entries for rp_engine_pairs.
// JS: JKR solver (0 exact as the reference, 1 with the no-root shortcut,
// 2 with the network's radii from k_jkr_collect + the MLP kernel).
// HIST: physics-fix mode with a tangential history (reads table HR, writes HW).
template <int NORMAL, int TANG, bool ADVANCE, bool RECORD, int JS, bool BATCH, bool DEAL, bool HIST>
__global__ void __launch_bounds__(kForceBlock) RP_FORCE_ATTR k_force(Ctl* ctl, int flag_slot, unsigned long long off, KParams P,
                                                  Dyn S, Dyn D, Stat T, const int* nbr_count, const int* nbr,
                                                  const Tri* tris, const uint32_t* tri_start,
                                                  const uint32_t* tri_idx, const float* xb, BlockStats* bstats,
                                                  unsigned long long* pairbuf, int* paircount, SurrDev sd,
                                                  SurrStats* sstats, Hist HR, Hist HW) {
  const int n = ctl->n;
  // zero-force pass (split_pass), dealt over the wave (DEAL) or per lane
  constexpr bool kSplit = split_pass(NORMAL, TANG, RECORD, JS, HIST);
  // the rows of this step have been read by the MLP kernel; the next step's
  // collect kernel starts from zero
  if (JS >= 2 && blockIdx.x == 0 && threadIdx.x == 0) {
    unsigned mx = 0u;
    for (int r = 0; r < kSurrRanges; ++r) {
      mx = max(mx, sd.counter[r * kSurrCounterStride]);
      sd.counter[r * kSurrCounterStride] = 0u;
    }
    ctl->surr_rows_max = max(ctl->surr_rows_max, mx);
  }
  if (ctl->stall_off <= (long long)off || ctl->overflow) {
    // The lists may miss a pair for this step: keep the state as it is (the
    // host swaps the buffers after every launch, so copy it across) and let
    // the host run this step again after a rebuild.
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) copy_slot(S, D, i, P.cap);
    if (HIST && i < n) copy_hist(HR, HW, i, i, P);
    if (threadIdx.x == 0) bstats[blockIdx.x] = BlockStats{0u, 0u, 0u, 0u, 0u};
    if (JS >= 2 && threadIdx.x == 0) sstats[blockIdx.x] = SurrStats{0u, 0u, 0.f, 0.f, 0.0, 0.0};
    return;
  }
  const int cap = P.cap;
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  unsigned int nwall = 0;
  Acc acc;
  acc.f = zero3();
  acc.t = zero3();
  acc.ncont = 0;
  acc.njkr = 0;
  acc.npair = 0;
  acc.overlap = false;
  acc.srow = -1;
  acc.sok = 0;
  acc.rmask = 0ull;
  acc.nmask = 0ull;
  acc.sserved = 0;
  acc.sfall = 0;
  acc.emax = 0.f;
  acc.rmax = 0.f;
  acc.esum = 0.0;
  acc.rsum = 0.0;
  acc.hp = 0;
  acc.hold = 0;
  acc.hn = 0;
  bool far = false;
  uint8_t flags = 0;
  Own o{};
  int cnt = 0, cnt_f = 0;  // list length; entries that can carry a force
  if (i < n) {
    flags = T.flags[i];
    ldxr(S.xr, i, o.x, o.r);
    o.v = ld3s(S.v, i, cap);
    o.w = ld3s(S.w, i, cap);
    o.m = T.m[i];
    cnt = nbr_count[i];
    // an advancing launch needs only the near part of the list: the far
    // entries have zero force until the next rebuild (k_build), and only
    // the call's last launch counts them
    cnt_f = (ADVANCE && kSplit) ? nbr_count[cap + i] : cnt;
  }
  const bool live = i < n && !(flags & FLAG_DEAD);
  if (live) {
    o.rc0 = ref_coord<0>(o.x.x, P);
    o.rc1 = ref_coord<1>(o.x.y, P);
    o.rc2 = ref_coord<2>(o.x.z, P);
  }
  unsigned long long defer = 0ull;
  if constexpr (kSplit && DEAL) {
    // Zero-force pass over the list entries below 64, shared by the wave:
    // a contact whose overlap is past the model's zero-force threshold adds
    // exactly zero (see contact_pp), so it is only counted (reference-grid
    // multiplicity); the others are marked for the ordered pass below.
    // The wave's entries are dealt to its 64 lanes in turn, so the pass
    // takes (sum of the wave's list lengths) / 64 rounds instead of the
    // longest list; counts and marks go through LDS atomics (integers, so
    // the order does not matter).
    constexpr int kW = kForceBlock / 64;
    __shared__ uint8_t s_owner[kW][64 * 64];     // lane owning each dealt entry
    __shared__ int s_excl[kW][64];               // first dealt entry of each lane
    __shared__ double s_own[kW][5][64];          // x, y, z, r, zero-force threshold
    __shared__ int s_rc[kW][3][64];              // reference-grid cell
    __shared__ unsigned int s_cnt[kW][2][64];    // contacts, JKR failures
    __shared__ unsigned long long s_def[kW][64]; // entries for the ordered pass
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // threshold below which the force is zero; -inf: no zero-force contacts
    double zthr = -HUGE_VAL;
    if (live) {
      const InstConsts* ic = BATCH ? P.itab + T.inst[i] : nullptr;
      zthr = zero_threshold_r(P, BATCH ? ic->mc.pp : P.mc.pp, o.r, NORMAL);
    }
    const int ka = live ? min(cnt_f, 64) : 0;
    int incl = ka;
    for (int sh = 1; sh < 64; sh <<= 1) {
      const int v = __shfl_up(incl, sh, 64);
      if (lane >= sh) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    const int ex = incl - ka;
    s_excl[wv][lane] = ex;
    s_own[wv][0][lane] = o.x.x;
    s_own[wv][1][lane] = o.x.y;
    s_own[wv][2][lane] = o.x.z;
    s_own[wv][3][lane] = o.r;
    s_own[wv][4][lane] = zthr;
    s_rc[wv][0][lane] = o.rc0;
    s_rc[wv][1][lane] = o.rc1;
    s_rc[wv][2][lane] = o.rc2;
    s_cnt[wv][0][lane] = 0u;
    s_cnt[wv][1][lane] = 0u;
    s_def[wv][lane] = 0ull;
    for (int t = 0; t < ka; ++t) s_owner[wv][ex + t] = (uint8_t)lane;
    wave_sync();
    const int ib = i - lane;  // slot of the wave's lane 0
    // each lane takes kDealBatch entries per round, their loads issued together
    for (int f0 = lane; f0 < total; f0 += kDealBatch * 64) {
      int pq[kDealBatch], kq[kDealBatch], jq[kDealBatch];
      bool hq[kDealBatch];
#pragma unroll
      for (int q = 0; q < kDealBatch; ++q) {
        const int f = f0 + 64 * q;
        hq[q] = f < total;
        pq[q] = hq[q] ? (int)s_owner[wv][f] : pq[0];
        kq[q] = hq[q] ? f - s_excl[wv][pq[q]] : kq[0];
      }
#pragma unroll
      for (int q = 0; q < kDealBatch; ++q) jq[q] = hq[q] ? nbr[kq[q] * cap + ib + pq[q]] : jq[0];
      dv3 xq[kDealBatch];
      double rq[kDealBatch];
#pragma unroll
      for (int q = 0; q < kDealBatch; ++q) ldxr(S.xr, jq[q], xq[q], rq[q]);
#pragma unroll
      for (int q = 0; q < kDealBatch; ++q) {
        if (!hq[q]) break;
        const int p = pq[q], k = kq[q];
        const dv3 xo = mk(s_own[wv][0][p], s_own[wv][1][p], s_own[wv][2][p]);
        const double zt = s_own[wv][4][p];
        const dv3 d = pair_delta(P, xo, xq[q]);
        const double dsq = dot(d, d);
        const double rsum = s_own[wv][3][p] + rq[q];
        // distance past (r_sum - threshold)(1 + 1e-6): the overlap is below
        // the threshold by far more than rounding, so no square root is
        // needed; nearer pairs take the exact test
        const double tz = (rsum - zt) * (1.0 + 1e-6);
        // Only the call's last launch (not ADVANCE) keeps the counters the
        // host reads; the others need just the entries with force, so a pair
        // past the zero-force distance is dropped whether in the cutoff or not.
        if (ADVANCE && dsq > tz * tz) continue;
        if (!in_contact(P, dsq, rsum)) continue;
        bool zero = !ADVANCE && dsq > tz * tz;
        if (!zero) {
          const double ov = rsum - sqrt(dsq);
          zero = NORMAL == RP_NORMAL_JKR ? ov < zt - 1e-8 * fabs(ov) : ov < zt;
        }
        if (!zero) {
          atomicOr(&s_def[wv][p], 1ull << k);
          continue;
        }
        if (ADVANCE) continue;  // not counted (above)
        unsigned int mult = 1u;  // physics-fix mode: every pair once
        if (!P.fix) {
          const int cj0 = ref_coord_fast<0>(xq[q].x, P), cj1 = ref_coord_fast<1>(xq[q].y, P),
                    cj2 = ref_coord_fast<2>(xq[q].z, P);
          if (cj0 < 0 || cj0 >= P.ref_dims[0] || cj1 < 0 || cj1 >= P.ref_dims[1] || cj2 < 0 ||
              cj2 >= P.ref_dims[2])
            continue;
          mult = (unsigned int)(__popc(ref_mask_fast<0>(s_rc[wv][0][p], cj0, P)) *
                                __popc(ref_mask_fast<1>(s_rc[wv][1][p], cj1, P)) *
                                __popc(ref_mask_fast<2>(s_rc[wv][2][p], cj2, P)));
        }
        if (mult) {
          atomicAdd(&s_cnt[wv][0][p], mult);
          if (NORMAL == RP_NORMAL_JKR) atomicAdd(&s_cnt[wv][1][p], mult);
        }
      }
    }
    wave_sync();
    acc.ncont += s_cnt[wv][0][lane];
    acc.njkr += s_cnt[wv][1][lane];
    defer = s_def[wv][lane];
  }
  if (i < n) {
    if (flags & FLAG_DEAD) {
      const double qnan = __builtin_nan("");
      stxr(D.xr, i, mk(qnan, qnan, qnan), qnan);
      if (HIST) HW.cnt[i] = 0;
    } else {
      if (HIST) acc.hold = HR.cnt[i];
      const int cx = fine_coord(o.x.x, 0, P, true), cy = fine_coord(o.x.y, 1, P, true),
                cz = fine_coord(o.x.z, 2, P, true);
      const uint32_t cell = cell_linear(P, cx, cy, cz);
      const uint32_t t0 = tri_start[cell], t1 = tri_start[cell + 1];
      unsigned long long* prow = RECORD ? pairbuf + (size_t)i * P.pair_cap : nullptr;
      if (JS >= 2) {
        acc.srow = sd.base[i];
        if (acc.srow >= 0) acc.rmask = sd.rowmask[i];
        acc.sok = BATCH ? (int)sd.inst[T.inst[i]].x : sd.pp_ok;
      }
      const InstConsts* ic = BATCH ? P.itab + T.inst[i] : nullptr;
      const MaterialConsts& mc = BATCH ? ic->mc : P.mc;
      const double far_pp = BATCH ? ic->jkr_far_pp : P.jkr_far_pp;
      const double far_pw = BATCH ? ic->jkr_far_pw : P.jkr_far_pw;
      if constexpr (kSplit) {
        if constexpr (!DEAL) {
          // The zero-force pass in one lane per particle (the tiling the
          // autotuner may pick instead of the wave-dealt one above).
          const double zthr = zero_threshold_r(P, mc.pp, o.r, NORMAL);
          const bool zon = NORMAL == RP_NORMAL_SJKR || mc.pp.gamma > 0.0;
          const int ka = min(cnt_f, 64);
          for (int base = 0; base < ka; base += kNbrBatch) {
            int jj[kNbrBatch];
            dv3 xq[kNbrBatch];
            double rq[kNbrBatch];
#pragma unroll
            for (int q = 0; q < kNbrBatch; ++q) jj[q] = (base + q < ka) ? nbr[(base + q) * cap + i] : -1;
#pragma unroll
            for (int q = 0; q < kNbrBatch; ++q) {
              const int j = jj[q] < 0 ? i : jj[q];
              ldxr(S.xr, j, xq[q], rq[q]);
            }
#pragma unroll
            for (int q = 0; q < kNbrBatch; ++q) {
              const dv3 d = pair_delta(P, o.x, xq[q]);
              const double dsq = dot(d, d);
              const double rsum = o.r + rq[q];
              if (ADVANCE && zon) {
                // not counted (the call's last launch keeps the counters):
                // a pair past the zero-force distance with rounding to spare
                // has no force, whether in the cutoff or not
                const double tz = (rsum - zthr) * (1.0 + 1e-6);
                if (dsq > tz * tz) continue;
              }
              if (!(jj[q] >= 0 && in_contact(P, dsq, rsum))) continue;
              const double ov = rsum - sqrt(dsq);
              const bool zero = zon && (NORMAL == RP_NORMAL_JKR ? ov < zthr - 1e-8 * fabs(ov) : ov < zthr);
              if (!zero) {
                defer |= 1ull << (base + q);
                continue;
              }
              if (ADVANCE) continue;  // counters: the call's last launch only
              unsigned int mult = 1u;  // physics-fix mode: every pair once
              if (!P.fix) {
                const int cj0 = ref_coord_fast<0>(xq[q].x, P), cj1 = ref_coord_fast<1>(xq[q].y, P),
                          cj2 = ref_coord_fast<2>(xq[q].z, P);
                if (cj0 < 0 || cj0 >= P.ref_dims[0] || cj1 < 0 || cj1 >= P.ref_dims[1] || cj2 < 0 ||
                    cj2 >= P.ref_dims[2])
                  continue;
                mult = (unsigned int)(__popc(ref_mask_fast<0>(o.rc0, cj0, P)) *
                                      __popc(ref_mask_fast<1>(o.rc1, cj1, P)) *
                                      __popc(ref_mask_fast<2>(o.rc2, cj2, P)));
              }
              acc.ncont += mult;
              if (NORMAL == RP_NORMAL_JKR) acc.njkr += mult;
            }
          }
        }
        // the marked entries and the entries from 64 on get the full
        // contact, in list order, so the force sums keep the reference's order
        while (defer) {
          const int k = __ffsll((long long)defer) - 1;
          defer &= defer - 1ull;
          contact_pp<NORMAL, TANG, RECORD, JS, HIST>(P, mc, far_pp, S, T, o, nbr[k * cap + i], k, acc, ctl, prow, sd,
                                                     HR, HW, i);
        }
        for (int k = 64; k < cnt_f; ++k) {
          const int j = nbr[k * cap + i];
          dv3 xj;
          double rj;
          ldxr(S.xr, j, xj, rj);
          const dv3 d = pair_delta(P, o.x, xj);
          if (in_contact(P, dot(d, d), o.r + rj))
            contact_pp<NORMAL, TANG, RECORD, JS, HIST>(P, mc, far_pp, S, T, o, j, k, acc, ctl, prow, sd, HR, HW, i);
        }
      } else {
        for (int base = 0; base < cnt; base += kNbrBatch) {
          int jj[kNbrBatch];
          dv3 xq[kNbrBatch];
          double rq[kNbrBatch];
      #pragma unroll
          for (int q = 0; q < kNbrBatch; ++q) jj[q] = (base + q < cnt) ? nbr[(base + q) * cap + i] : -1;
      #pragma unroll
          for (int q = 0; q < kNbrBatch; ++q) {
            const int j = jj[q] < 0 ? i : jj[q];
            ldxr(S.xr, j, xq[q], rq[q]);
          }
          int hit = 0;
      #pragma unroll
          for (int q = 0; q < kNbrBatch; ++q) {
            const dv3 d = pair_delta(P, o.x, xq[q]);
            if (jj[q] >= 0 && in_contact(P, dot(d, d), o.r + rq[q])) hit |= 1 << q;
          }
          while (hit) {
            const int q = __ffs(hit) - 1;
            hit &= hit - 1;
            const int j = kNbrBatch == 2 ? (q == 0 ? jj[0] : jj[kNbrBatch - 1])
                                         : sel4(q, jj[0], jj[1 % kNbrBatch], jj[2 % kNbrBatch], jj[3 % kNbrBatch]);
            contact_pp<NORMAL, TANG, RECORD, JS, HIST>(P, mc, far_pp, S, T, o, j, base + q, acc, ctl, prow, sd, HR,
                                                       HW, i);
          }
        }
      }
      if (RECORD) paircount[i] = min(acc.npair, P.pair_cap);
      if (HIST) HW.cnt[i] = acc.hn;
      for (uint32_t q = t0; q < t1; ++q)
        contact_wall<NORMAL, JS>(P, mc, far_pw, o, tris[tri_idx[q]], tri_idx[q], acc, nwall, sd);



      // apply (simulation.rs:533-536) and second half kick (:543-561)
      const dv3 a = mk(P.g[0] + acc.f.x / o.m, P.g[1] + acc.f.y / o.m, P.g[2] + acc.f.z / o.m);
      const dv3 tau = mk(0.0 + acc.t.x, 0.0 + acc.t.y, 0.0 + acc.t.z);
      dv3 rx = ld3s(S.rx, i, cap), rv = ld3s(S.rv, i, cap), rw = ld3s(S.rw, i, cap);
      dv3 v = o.v, w = o.w;
      const dv3 alpha = angular_acc(tau, o.r, o.m);
      if (!(flags & FLAG_FIXED)) {
        const double dt = P.dt;
        kahan1(v.x, rv.x, (0.5 * a.x) * dt);
        kahan1(v.y, rv.y, (0.5 * a.y) * dt);
        kahan1(v.z, rv.z, (0.5 * a.z) * dt);
        kahan1(w.x, rw.x, (0.5 * alpha.x) * dt);
        kahan1(w.y, rw.y, (0.5 * alpha.y) * dt);
        kahan1(w.z, rw.z, (0.5 * alpha.z) * dt);
      }
      dv3 x = o.x;
      bool alive = true;
      if (ADVANCE) {
        if (!begin_step(P, x, v, w, rx, rv, rw, a, alpha, flags)) {
          mark_removed(ctl, T, D, i, cap, flags);
          alive = false;
        } else {
          far = moved_too_far(P, x, v, a, xb, i, ctl, off + 1);
        }
      } else {
        st3(D.a, i, cap, a);
        st3(D.tau, i, cap, tau);
      }
      if (alive) {
        stxr(D.xr, i, x, o.r);
        st3s(D.v, i, cap, v);
        st3s(D.w, i, cap, w);
        st3s(D.rx, i, cap, rx);
        st3s(D.rv, i, cap, rv);
        st3s(D.rw, i, cap, rw);
      }
    }
  }
  if (JS >= 2 && i < n) sd.nmask[i] = acc.nmask;
  // counters: per-workgroup plain stores; the rebuild request is the only
  // atomic, issued by a wave only when one of its particles needs it
  const unsigned int c1 = wave_sum_u32(acc.ncont), c2 = wave_sum_u32(nwall), c3 = wave_sum_u32(acc.njkr);
  const unsigned int c4 = HIST ? wave_sum_u32((unsigned int)acc.hn) : 0u;
  const bool anyo = __any(acc.overlap), anyf = __any(far);
  __shared__ unsigned int red[kForceBlock / 64][5];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[wv][0] = c1;
    red[wv][1] = c2;
    red[wv][2] = c3;
    red[wv][3] = anyo ? 1u : 0u;
    red[wv][4] = c4;
    if (ADVANCE && anyf) atomicOr(&ctl->rebuild[flag_slot], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockStats b{0u, 0u, 0u, 0u, 0u};
    for (int q = 0; q < kForceBlock / 64; ++q) {
      b.contacts += red[q][0];
      b.wall_contacts += red[q][1];
      b.jkr_failures += red[q][2];
      b.overlap |= red[q][3];
      b.history += red[q][4];
    }
    bstats[blockIdx.x] = b;
    if (b.overlap) {
      // count this step once (simulation.rs:528-530 warns once per step)
      const unsigned long long lab = ctl->label_base + off + 1ull;
      if (atomicMax(&ctl->ovl_last, lab) < lab) atomicAdd(&ctl->ovl_steps, 1ull);
    }
  }
  if (JS >= 2) {
    // surrogate counters: wave reductions, then one store per workgroup
    unsigned int s1 = wave_sum_u32(acc.sserved), s2 = wave_sum_u32(acc.sfall);
    float em = acc.emax, rm = acc.rmax;
    double es = acc.esum, rs = acc.rsum;
    for (int off = 32; off > 0; off >>= 1) {
      em = fmaxf(em, __shfl_down(em, off, 64));
      rm = fmaxf(rm, __shfl_down(rm, off, 64));
      es += __shfl_down(es, off, 64);
      rs += __shfl_down(rs, off, 64);
    }
    __shared__ SurrStats sred[kForceBlock / 64];
    if ((threadIdx.x & 63) == 0) sred[wv] = SurrStats{s1, s2, em, rm, es, rs};
    __syncthreads();
    if (threadIdx.x == 0) {
      SurrStats b = sred[0];
      for (int q = 1; q < kForceBlock / 64; ++q) {
        b.served += sred[q].served;
        b.fallback += sred[q].fallback;
        b.err_max = fmaxf(b.err_max, sred[q].err_max);
        b.raw_max = fmaxf(b.raw_max, sred[q].raw_max);
        b.err_sum += sred[q].err_sum;
        b.raw_sum += sred[q].raw_sum;
      }
      sstats[blockIdx.x] = b;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (JS >= 2) ctl->surr_mask_epoch = ctl->rebuilds + 1ull;
    ctl->steps_done += 1;
    ctl->at_rest = ADVANCE ? 0 : 1;
  }
}

// Routing test of the JKR surrogate for the list entry k of particle i,
// with exactly the force kernel's pair tests and arithmetic (in_contact,
// reference-grid visibility, overlap, R*), so both kernels agree on which
// contacts consume a feature row.  Returns the row's scaled features.
__device__ inline bool collect_pair(const KParams& P, const Dyn& S, const Own& o, int j, int k, const SurrDev& sd,
                                    const PairConsts& c, double far, float f_lne, float f_lng, int ok,
                                    float4* feat) {
  dv3 xj;
  double rj;
  ldxr(S.xr, j, xj, rj);
  const dv3 dd = pair_delta(P, o.x, xj);
  const double dsq = dot(dd, dd);
  const double rsum = o.r + rj;
  if (!in_contact(P, dsq, rsum)) return false;
  if (!P.fix) {  // the physics-fix mode has no reference-grid visibility
    // the division-free forms give the same integers (ref_coord_fast)
    const int cj0 = ref_coord_fast<0>(xj.x, P), cj1 = ref_coord_fast<1>(xj.y, P), cj2 = ref_coord_fast<2>(xj.z, P);
    if (cj0 < 0 || cj0 >= P.ref_dims[0] || cj1 < 0 || cj1 >= P.ref_dims[1] || cj2 < 0 || cj2 >= P.ref_dims[2])
      return false;
    const int mult = __popc(ref_mask_fast<0>(o.rc0, cj0, P)) * __popc(ref_mask_fast<1>(o.rc1, cj1, P)) *
                     __popc(ref_mask_fast<2>(o.rc2, cj2, P));
    if (mult == 0) return false;
  }
  const double overlap = rsum - sqrt(dsq);
  const double rstar = (o.r * rj) / (o.r + rj);
  if (!(c.gamma > 0.0)) return false;
  if (overlap < far - 1e-8 * fabs(overlap)) return false;
  const double s = cbrt(rstar * rstar * c.jc);
  const double t = s * s / rstar;
  if (overlap < kJkrDMin * t - 1e-3 * rstar - 1e-8 * (rstar + t + fabs(overlap))) return false;
  double d = 0.0;
  if (!surr_route(overlap, rstar, s, t, sd, k, d, ok)) return false;
  if (feat) {
    feat->x = scale_in((float)sqrt(d - kJkrDMin), sd.sc.in_lo[0], sd.sc.in_hi[0]);
    feat->y = scale_in((float)log(rstar), sd.sc.in_lo[1], sd.sc.in_hi[1]);
    feat->z = f_lne;
    feat->w = f_lng;
  }
  return true;
}

// Gathers the feature rows of every contact the JKR surrogate serves in
// this step (one thread per particle).  It visits the neighbour-list entries
// the last force launch marked as near the served set (all entries after a
// rebuild, which re-orders the lists), marks those served now, places each
// particle's rows contiguously (a wave-wide prefix sum plus one atomic per
// wave) and writes them.  base[i] = -1 when i has no rows or they would not
// fit; rowmask[i] names the entries that have rows, in row order.  A served
// contact without a row takes the exact solve in the force kernel.  Six
// waves per SIMD (79 VGPRs, no spills): at eight (64 VGPRs) it spilled 44 B
// and collected the 1.97M crystal's rows in 730 us instead of 666 us.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6))) k_jkr_collect(
    Ctl* ctl, unsigned long long off, KParams P, Dyn S, Stat T,
    const int* nbr_count, const int* nbr, SurrDev sd) {
  if (ctl->stall_off <= (long long)off || ctl->overflow) return;  // the force launch keeps the state: no rows
  const int n = ctl->n;
  const bool full = ctl->surr_mask_epoch != ctl->rebuilds + 1ull;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  unsigned long long mask = 0ull;
  Own o;
  // the particle's material pair (per instance in a batched engine)
  PairConsts pc = P.mc.pp;
  double far = P.jkr_far_pp;
  float f_lne = sd.f_lne, f_lng = sd.f_lng;
  int ok = sd.pp_ok;
  if (i < n && !(T.flags[i] & FLAG_DEAD)) {
    if (P.ninst > 1) {
      const int q = T.inst[i];
      pc = P.itab[q].mc.pp;
      far = P.itab[q].jkr_far_pp;
      const float4 f = sd.inst[q];
      ok = (int)f.x;
      f_lne = f.y;
      f_lng = f.z;
    }
    const int kmax = min(nbr_count[i], 64);
    unsigned long long visit = full ? ~0ull : sd.nmask[i];
    if (kmax < 64) visit &= (1ull << kmax) - 1ull;
    if (visit) {
      ldxr(S.xr, i, o.x, o.r);
      o.rc0 = ref_coord<0>(o.x.x, P);
      o.rc1 = ref_coord<1>(o.x.y, P);
      o.rc2 = ref_coord<2>(o.x.z, P);
      while (visit) {
        const int k = __ffsll((long long)visit) - 1;
        visit &= visit - 1ull;
        if (collect_pair(P, S, o, nbr[k * P.cap + i], k, sd, pc, far, f_lne, f_lng, ok, nullptr)) mask |= 1ull << k;
      }
    }
  }
  // rows of this workgroup: wave prefix sums, then one claim per workgroup
  // from range blockIdx.x % kSurrRanges
  const int c = __popcll(mask);
  int incl = c;
  for (int sh = 1; sh < 64; sh <<= 1) {
    const int v = __shfl_up(incl, sh, 64);
    if (lane >= sh) incl += v;
  }
  __shared__ unsigned wsum[kBlock / 64];
  __shared__ unsigned bbase;
  const int wv = threadIdx.x >> 6;
  if (lane == 63) wsum[wv] = (unsigned)incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0u;
    for (int q = 0; q < kBlock / 64; ++q) {
      const unsigned t = wsum[q];
      wsum[q] = tot;
      tot += t;
    }
    const int range = blockIdx.x % kSurrRanges;
    bbase = tot ? (unsigned)range * (unsigned)(sd.rows_cap / kSurrRanges) + atomicAdd(&sd.counter[range * kSurrCounterStride], tot) : 0u;
  }
  __syncthreads();
  if (i >= n) return;
  const long long b = (long long)bbase + wsum[wv] + (incl - c);
  const long long range_end = (long long)(blockIdx.x % kSurrRanges + 1) * (sd.rows_cap / kSurrRanges);
  const bool fits = c > 0 && b + c <= range_end;
  sd.base[i] = fits ? (int)b : -1;
  if (!fits) return;
  sd.rowmask[i] = mask;
  float4* rows = reinterpret_cast<float4*>(sd.x);
  int r = (int)b;
  while (mask) {
    const int k = __ffsll((long long)mask) - 1;
    mask &= mask - 1ull;
    float4 f;
    (void)collect_pair(P, S, o, nbr[k * P.cap + i], k, sd, pc, far, f_lne, f_lng, ok, &f);
    rows[r++] = f;
  }
}

// Plain streaming copy of known size: reference point for hardware traffic
// counters (FETCH_SIZE / WRITE_SIZE) measured in the same profile.
// 16-byte non-temporal loads and stores, grid-stride, 512-thread workgroups:
// the fastest plain copy found on MI355X (~6.2 TB/s at 1 GiB each way).
__global__ void __launch_bounds__(512) k_copy_probe(const double* __restrict__ src, double* __restrict__ dst,
                                                    size_t n) {
  const size_t n2 = n / 2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += (size_t)gridDim.x * blockDim.x) {
    const double a = __builtin_nontemporal_load(src + 2 * i), b = __builtin_nontemporal_load(src + 2 * i + 1);
    __builtin_nontemporal_store(a, dst + 2 * i);
    __builtin_nontemporal_store(b, dst + 2 * i + 1);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1)) dst[n - 1] = src[n - 1];
}

// Moves the label base forward by `by` steps: offsets are relative to it.
__global__ void k_shift_labels(Ctl* ctl, unsigned long long by) {
  if (ctl->stall_off != kNoStall) ctl->stall_off -= (long long)by;
  ctl->label_base += by;
}

// Cells of every slot (rp_engine_cells): the reference-grid cell index of
// grid.rs:30-40 (coords = floor((x - origin) / cs) as i32, range check,
// cx + cy dx + cz dx dy; -1 outside the grid) and the engine's fine-cell key
// as k_keys computes it.  Removed slots: -2 and UINT32_MAX.
__global__ void __launch_bounds__(kBlock) k_cells(const Ctl* ctl, KParams P, Dyn S, Stat T, long long* ref_id,
                                                  uint32_t* key) {
  const int n = ctl->n;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (T.flags[i] & FLAG_DEAD) {
    ref_id[i] = -2;
    key[i] = 0xffffffffu;
    return;
  }
  const dv3 x = ldx(S.xr, i);
  const int c0 = ref_coord<0>(x.x, P), c1 = ref_coord<1>(x.y, P), c2 = ref_coord<2>(x.z, P);
  const bool in = c0 >= 0 && c1 >= 0 && c2 >= 0 && c0 < P.ref_dims[0] && c1 < P.ref_dims[1] && c2 < P.ref_dims[2];
  ref_id[i] = in ? (long long)c0 + (long long)c1 * P.ref_dims[0] + (long long)c2 * P.ref_dims[0] * P.ref_dims[1]
                 : -1ll;
  key[i] = (uint32_t)T.inst[i] * P.ncells +
           cell_linear(P, fine_coord(x.x, 0, P, false), fine_coord(x.y, 1, P, false), fine_coord(x.z, 2, P, false));
}

__global__ void k_request_rebuild(Ctl* ctl, int par) {
  ctl->rebuild[par] = 1;
  ctl->rebuild[par ^ 1] = 0;
}

// ---- host side ----------------------------------------------------------------

struct HostMesh {
  std::vector<float> v9, n3;
  double origin[3] = {0.0, 0.0, 0.0};  // surface motion: rotation omega about origin
  double omega[3] = {0.0, 0.0, 0.0};
};

size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

}  // namespace

struct rp_engine {
  rp_params prm{};
  int device = 0;
  hipStream_t stream = nullptr;
  int cap = 0;         // slot capacity
  int n_uploaded = 0;
  double r_max = 0.0, r_mean = 0.0;
  double v_max = 0.0;  // fastest uploaded particle (sizes the automatic skin)
  uint64_t step_count = 0;
  bool have_particles = false;
  // a skin picked by the autotuner is widened when it turns out too thin
  // for later motion (see skin_guard)
  bool skin_tuned = false;
  uint64_t guard_steps = 0;
  unsigned long long guard_r0 = 0;
  bool geometry_dirty = true;
  bool record = false;
  KParams kp{};
  Ctl host_ctl{};
  Ctl* d_ctl = nullptr;
  // buffers
  void* d_dyn_mem[2] = {nullptr, nullptr};
  void* d_stat_mem[2] = {nullptr, nullptr};
  Dyn dyn[2]{};       // dyn[cur] holds the state; the force kernel writes dyn[cur ^ 1]
  Stat stat[2]{};     // stat[0] holds the constants; stat[1] is rebuild scratch
  int cur = 0;
  uint64_t chain_no = 0;  // rebuild chains launched; chain c reads request slot c & 1
  uint64_t launch_no = 0; // force launches issued (labels for the stall watermark)
  uint64_t label_host = 0; // launch number of the device's label base as of the last enqueued work
  // captured step loops (kGraphSteps launches), one per rebuild-request
  // parity at their start; valid while gen matches
  hipGraphExec_t graph_exec[2] = {nullptr, nullptr};
  uint64_t graph_gen[2] = {0, 0};
  uint64_t gen = 1;        // bumped by every change to what launches capture
  bool use_graph = true;
  int tiling = RP_TILING_WAVE;  // zero-force pass of the force kernel
  uint64_t graph_replays = 0;
  uint64_t steps_done = 0;  // steps that advanced the state (device count at the last sync)
  bool at_rest = true;    // state is at rest (needs k_begin before the next force launch)
  BlockStats* d_bstats = nullptr;
  int force_grid = 0;
  float* d_xb = nullptr;
  uint32_t* d_key = nullptr;
  uint32_t* d_tmp = nullptr;
  uint32_t* d_cell_count = nullptr;
  uint32_t* d_cell_start = nullptr;
  uint32_t* d_block_sums = nullptr;
  float* d_cell_rmax = nullptr;  // largest radius in each cell (k_cell_rmax)
  size_t cells_alloc = 0;
  int* d_nbr_count = nullptr;
  int* d_nbr = nullptr;
  int nbr_cap = 0;
  Tri* d_tris = nullptr;
  uint32_t* d_tri_start = nullptr;
  uint32_t* d_tri_idx = nullptr;
  std::vector<HostMesh> meshes;
  unsigned long long* d_pairbuf = nullptr;
  int* d_paircount = nullptr;
  int pair_cap = 0;
  bool pairs_valid = false;
  // optional per-kernel timing with HIP events on the engine stream
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  struct Mark {
    int kind;
    size_t idx;      // start event; end event = idx + 1
    uint64_t count;  // launches the mark covers
  };
  std::vector<Mark> ev_marks;
  hipEvent_t ext_start = nullptr, ext_stop = nullptr;  // the next force launch's timing pair
  // batched engine: instances, their materials, the device table
  int ninst = 1;
  std::vector<rp_material> inst_particle, inst_wall;
  std::vector<size_t> inst_count;  // uploaded particles per instance
  InstConsts* d_itab = nullptr;
  WallMotion* d_mov = nullptr;  // kMaxMoving entries, allocated by the first motion
  // JKR solver (RP_JKR_*) and the surrogate's network and buffers
  int jkr_solver = RP_JKR_EXACT;
  int surr_check = 0;
  rp_jkr_envelope surr_env{};
  MlpHost surr_net;
  int* d_surr_base = nullptr;  // [cap]
  unsigned long long* d_surr_rowmask = nullptr;  // [cap]
  unsigned long long* d_surr_nmask = nullptr;    // [cap]
  float* d_surr_x = nullptr;   // [surr_rows_cap x 4]
  float* d_surr_y = nullptr;   // [surr_rows_cap x 2]
  int surr_rows_cap = 0;
  SurrStats* d_sstats = nullptr;  // [cap / kForceBlock + 1]
  float4* d_surr_inst = nullptr;  // [ninst] batched engines: per-instance envelope data
  unsigned* d_surr_counters = nullptr;  // [kSurrRanges * kSurrCounterStride] row claims
  bool surr_inst_valid = false;
  bool sstats_valid = false;     // the last force launch wrote d_sstats
  // physics-fix tangential history: hist[cur] holds the tables, the force
  // kernel writes hist[cur ^ 1]; capacity hcap = nbr_cap entries per slot
  bool hist_on = false;
  void* d_hist_mem[2] = {nullptr, nullptr};
  Hist hist[2]{};
  int hcap = 0;
};

namespace {

void free_ptr(void* p) {
  if (p) (void)hipFree(p);
}

// History buffer: keys (u32) [hcap][cap], xi (3 doubles) [hcap][cap], counts (int) [cap].
size_t hist_bytes(int cap, int hcap) {
  const size_t c = (size_t)cap, h = (size_t)hcap;
  return round_up(h * c * sizeof(uint32_t), 256) + round_up(3 * h * c * sizeof(double), 256) +
         round_up(c * sizeof(int), 256);
}
Hist carve_hist(void* mem, int cap, int hcap) {
  Hist H{};
  char* p = (char*)mem;
  const size_t c = (size_t)cap, h = (size_t)hcap;
  H.key = (uint32_t*)p;
  p += round_up(h * c * sizeof(uint32_t), 256);
  H.xi = (double*)p;
  p += round_up(3 * h * c * sizeof(double), 256);
  H.cnt = (int*)p;
  return H;
}

MaterialConsts material_consts(const rp_material& pm, const rp_material& wm) {
  auto pair = [](double e1, double nu1, double e2, double nu2, double g1, double g2) {
    double inv = (1.0 - nu1 * nu1) / e1 + (1.0 - nu2 * nu2) / e2;
    return pair_consts(1.0 / inv, (g1 + g2) * 0.5);
  };
  MaterialConsts m{};
  m.pp = pair(pm.youngs_modulus, pm.poissons_ratio, pm.youngs_modulus, pm.poissons_ratio, pm.surface_energy,
              pm.surface_energy);
  m.pw = pair(pm.youngs_modulus, pm.poissons_ratio, wm.youngs_modulus, wm.poissons_ratio, pm.surface_energy,
              wm.surface_energy);
  m.mu = pm.friction_coefficient;
  // physics-fix shear modulus (the reference's effective_shear_modulus is a
  // placeholder 0, material.rs:39-50): G = E / (2 (1 + nu)), 1/G* = 2 (2 - nu) / G
  const double g = pm.youngs_modulus / (2.0 * (1.0 + pm.poissons_ratio));
  const double ginv = (2.0 - pm.poissons_ratio) / g + (2.0 - pm.poissons_ratio) / g;
  m.gstar_pp = 1.0 / ginv;
  double lne = std::log(pm.restitution_coefficient);
  double beta = lne / std::sqrt(lne * lne + kPi * kPi);
  m.beta_abs = std::fabs(beta);
  return m;
}

// t(R*) = R*^(1/3) (8 pi gamma / E*)^(2/3) grows with R*, so the shortcut
// threshold of the largest possible R* bounds every contact's from below.
double jkr_far(const PairConsts& c, double rs) {
  if (!(c.gamma > 0.0)) return -HUGE_VAL;
  const double t = std::cbrt(rs) * std::pow(c.jc, 2.0 / 3.0);
  return (kJkrDMin * t - 1e-3 * rs - 1e-8 * (rs + t)) * (1.0 + 1e-12);
}

// sJKR's force is zero below dc = -sqrt(3 pi^2 gamma^2 R* / E*^2) (physics.rs:
// 256-262), which falls with R*: the value at the largest R*, widened by
// 1e-9 for rounding, bounds every contact's.  With gamma <= 0 the Hertz
// branch has force only for overlap > 0.
double sjkr_zero(const PairConsts& c, double rs) {
  if (!(c.gamma > 0.0)) return 0.0;
  return -std::sqrt(c.sterm * rs / c.estar_sq) * (1.0 + 1e-9);
}

rp_status validate(const rp_params* p) {
  if (!p) return fail(RP_ERR_INVALID, "params is NULL");
  if (!(p->dt > 0.0) || !std::isfinite(p->dt)) return fail(RP_ERR_INVALID, "dt must be positive and finite");
  for (int k = 0; k < 3; ++k) {
    if (!std::isfinite(p->bounds_min[k]) || !std::isfinite(p->bounds_max[k]))
      return fail(RP_ERR_INVALID, "bounds must be finite");
    if (p->periodic[k] && !(p->bounds_max[k] > p->bounds_min[k]))
      return fail(RP_ERR_INVALID, "periodic axis needs bounds_max > bounds_min");
  }
  if (p->normal_model < 0 || p->normal_model > 4) return fail(RP_ERR_INVALID, "invalid normal model");
  if (p->tangential_model < 0 || p->tangential_model > 2) return fail(RP_ERR_INVALID, "invalid tangential model");
  if (p->mode != RP_MODE_PARITY && p->mode != RP_MODE_PHYSICS_FIX) return fail(RP_ERR_INVALID, "unsupported mode");
  if (p->cell_order < RP_CELL_ORDER_AUTO || p->cell_order > RP_CELL_ORDER_PENCIL)
    return fail(RP_ERR_INVALID, "invalid cell_order");
  if (!(p->ref_cell_size > 0.0)) return fail(RP_ERR_INVALID, "ref_cell_size must be positive");
  if (p->cell_size < 0.0 || p->skin < 0.0) return fail(RP_ERR_INVALID, "cell_size and skin must be >= 0");
  return RP_OK;
}

// Host-side kernel parameters from rp_params + particle radii + meshes.
rp_status compute_geometry(rp_engine* e) {
  e->gen++;
  const rp_params& p = e->prm;
  KParams& K = e->kp;
  K.dt = p.dt;
  K.any_periodic = 0;
  for (int k = 0; k < 3; ++k) {
    K.bmin[k] = p.bounds_min[k];
    K.bmax[k] = p.bounds_max[k];
    K.bsize[k] = p.bounds_max[k] - p.bounds_min[k];
    K.periodic[k] = p.periodic[k] ? 1 : 0;
    K.any_periodic |= K.periodic[k];
    K.g[k] = p.gravity[k];
    K.ref_origin[k] = p.bounds_min[k];
  }
  K.normal = p.normal_model;
  K.tangential = p.tangential_model;
  K.fix = p.mode == RP_MODE_PHYSICS_FIX ? 1 : 0;
  K.mc = material_consts(p.particle, p.wall);
  K.margin_factor = (p.normal_model == RP_NORMAL_JKR || p.normal_model == RP_NORMAL_SJKR) ? 1.5 : 1.0;
  K.ref_cs = p.ref_cell_size;
  K.ref_inv_cs = 1.0 / p.ref_cell_size;
  for (int k = 0; k < 3; ++k) K.ref_dims[k] = sat_i32(std::ceil(K.bsize[k] / K.ref_cs)) + 1;

  // fine grid
  double rmax = e->r_max > 0.0 ? e->r_max : 1e-3;
  double rmean = e->r_mean > 0.0 ? e->r_mean : rmax;
  K.r_max = rmax;
  // t(R*) = R*^(1/3) (8 pi gamma / E*)^(2/3) grows with R*, so the shortcut
  // threshold of the largest possible R* bounds every contact's from below
  K.jkr_far_pp = jkr_far(K.mc.pp, 0.5 * rmax);
  K.jkr_far_pw = jkr_far(K.mc.pw, rmax);
  K.sjkr_zero_pp = sjkr_zero(K.mc.pp, 0.5 * rmax);
  // per-instance constants of a batched engine
  if (e->ninst > 1) {
    std::vector<InstConsts> tab((size_t)e->ninst);
    for (int q = 0; q < e->ninst; ++q) {
      tab[q].mc = material_consts(e->inst_particle[q], e->inst_wall[q]);
      tab[q].jkr_far_pp = jkr_far(tab[q].mc.pp, 0.5 * rmax);
      tab[q].jkr_far_pw = jkr_far(tab[q].mc.pw, rmax);
      tab[q].sjkr_zero_pp = sjkr_zero(tab[q].mc.pp, 0.5 * rmax);
    }
    free_ptr(e->d_itab);
    e->d_itab = nullptr;
    HIP_TRY(hipMalloc(&e->d_itab, tab.size() * sizeof(InstConsts)));
    HIP_TRY(hipMemcpy(e->d_itab, tab.data(), tab.size() * sizeof(InstConsts), hipMemcpyHostToDevice));
  }
  K.itab = e->d_itab;
  // automatic skin: a tenth of the largest radius, and wide enough that the
  // fastest uploaded particle cannot cover half of it between rebuild checks
  K.skin = p.skin > 0.0 ? p.skin : std::max(0.1 * rmax, 4.0 * (double)kRebuildInterval * e->v_max * p.dt);
  K.half_skin = 0.5 * K.skin;
  K.lookahead_steps = (double)kRebuildInterval;
  double cut_max = K.margin_factor * 2.0 * rmax + K.skin;
  double cs = p.cell_size > 0.0 ? p.cell_size : (rmax > 2.0 * rmean ? 0.5 * cut_max : cut_max);
  // non-periodic extent: domain bounds plus mesh bounding box, one cell of margin
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = p.bounds_min[k];
    hi[k] = p.bounds_max[k];
  }
  for (const HostMesh& m : e->meshes)
    for (size_t q = 0; q < m.v9.size(); ++q) {
      int k = (int)(q % 3);
      double v = (double)m.v9[q];
      if (!std::isfinite(v)) continue;
      lo[k] = std::min(lo[k], v);
      hi[k] = std::max(hi[k], v);
    }
  const double kMaxCells = 48.0e6;
  for (int attempt = 0; attempt < 64; ++attempt) {
    double total = 1.0;
    for (int k = 0; k < 3; ++k) {
      if (K.periodic[k]) {
        int n = (int)std::floor(K.bsize[k] / cs);
        if (n < 1) n = 1;
        K.gn[k] = n;
        K.go[k] = p.bounds_min[k];
        K.ginv[k] = (double)n / K.bsize[k];
      } else {
        double l0 = lo[k] - cs, l1 = hi[k] + cs;
        int n = (int)std::ceil((l1 - l0) / cs);
        if (n < 1) n = 1;
        K.gn[k] = n;
        K.go[k] = l0;
        K.ginv[k] = 1.0 / cs;
      }
      total *= (double)K.gn[k];
    }
    if (total * (double)std::max(e->ninst, 1) <= kMaxCells) break;
    cs *= 1.25;
  }
  // Cells well below the largest cutoff (a wide radius distribution) put most
  // of a small particle's search box out of its reach: the list build then
  // skips cells by their largest radius.  With cells of a full cutoff the
  // 27-cell box is nearly all in reach and the test costs more than it saves.
  K.cell_reach = cs < 0.75 * cut_max ? 1 : 0;
  // slowest-varying axis = the one with most cells
  int order[3] = {0, 1, 2};
  std::sort(order, order + 3, [&](int a, int b) { return K.gn[a] < K.gn[b]; });
  int stride = 1;
  for (int q = 0; q < 3; ++q) {
    K.gstride[order[q]] = stride;
    stride *= K.gn[order[q]];
  }
  K.ninst = std::max(e->ninst, 1);
  // Pencil numbering (cell_linear) when a layer of the slow axis holds many
  // particles (cubic boxes): there a contiguous run of cells in row-major
  // order is a thin full-width slab whose neighbours do not fit in L2.
  // rp_params.cell_order overrides the choice.
  const int fa = order[0], ma = order[1], sa = order[2];
  const double per_layer = (double)e->cap / ((double)K.ninst * (double)K.gn[sa]);
  bool pencil = per_layer > 8192.0 && K.gn[fa] >= 16 && K.gn[ma] >= 16;
  if (p.cell_order == RP_CELL_ORDER_ROW) pencil = false;
  if (p.cell_order == RP_CELL_ORDER_PENCIL) pencil = true;
  const double pf = std::ceil(K.gn[fa] / 8.0), pm = std::ceil(K.gn[ma] / 8.0);
  if (pencil && pm * pf * K.gn[sa] * 64.0 * K.ninst > kMaxCells) pencil = false;
  if (pencil) {
    K.gsh[sa] = 0;
    K.ghi[sa] = 64;
    K.glo[sa] = 0;
    K.gsh[fa] = 3;
    K.ghi[fa] = K.gn[sa] * 64;
    K.glo[fa] = 1;
    K.gsh[ma] = 3;
    K.ghi[ma] = (int)pf * K.gn[sa] * 64;
    K.glo[ma] = 8;
    K.ncells = (unsigned int)(pm * pf * K.gn[sa] * 64.0);
  } else {
    for (int k = 0; k < 3; ++k) {
      K.gsh[k] = 0;
      K.ghi[k] = K.gstride[k];
      K.glo[k] = 0;
    }
    K.ncells = (unsigned int)stride;
  }
  K.ncells_all = K.ncells * (unsigned int)K.ninst;
  // build positions are stored as float: |x - float(x)| <= |x| 2^-24 per axis
  double xmax = 0.0;
  for (int k = 0; k < 3; ++k)
    xmax = std::max(xmax, std::max(std::fabs(K.go[k]), std::fabs(K.go[k] + K.gn[k] / K.ginv[k])));
  K.safe_half_skin = K.half_skin - 2.0 * std::sqrt(3.0) * xmax * std::ldexp(1.0, -24);
  K.cap = e->cap;
  K.nbr_cap = e->nbr_cap;
  K.pair_cap = e->pair_cap;
  K.hcap = e->hcap;
  return RP_OK;
}

// The moving-wall table of the kernel parameters from the meshes' motions
// (meshes are concatenated in the order added).
rp_status fill_motion(rp_engine* e) {
  KParams& K = e->kp;
  WallMotion tab[kMaxMoving] = {};
  int nmov = 0;
  size_t first = 0;
  for (const HostMesh& m : e->meshes) {
    const size_t nt = m.n3.size() / 3;
    if (m.omega[0] != 0.0 || m.omega[1] != 0.0 || m.omega[2] != 0.0) {
      if (nmov == kMaxMoving) return fail(RP_ERR_INVALID, "more than 4 meshes move");
      tab[nmov].lo = (unsigned int)first;
      tab[nmov].hi = (unsigned int)(first + nt);
      for (int k = 0; k < 3; ++k) {
        tab[nmov].o[k] = m.origin[k];
        tab[nmov].w[k] = m.omega[k];
      }
      ++nmov;
    }
    first += nt;
  }
  if (nmov > 0) {
    if (!e->d_mov) HIP_TRY(hipMalloc(&e->d_mov, sizeof(tab)));
    HIP_TRY(hipStreamSynchronize(e->stream));  // queued steps may still read the old table
    HIP_TRY(hipMemcpy(e->d_mov, tab, sizeof(tab), hipMemcpyHostToDevice));
  }
  K.nmov = nmov;
  K.mov = e->d_mov;
  e->gen++;  // captured step loops hold the kernel parameters
  return RP_OK;
}

rp_status upload_meshes_binned(rp_engine* e) {
  e->gen++;
  if (rp_status st = fill_motion(e)) return st;
  KParams& K = e->kp;
  std::vector<Tri> tris;
  for (const HostMesh& m : e->meshes) {
    size_t nt = m.n3.size() / 3;
    for (size_t t = 0; t < nt; ++t) {
      Tri tr;
      const float* v = &m.v9[9 * t];
      tr.v0 = mk(v[0], v[1], v[2]);
      tr.v1 = mk(v[3], v[4], v[5]);
      tr.v2 = mk(v[6], v[7], v[8]);
      tr.n = mk(m.n3[3 * t], m.n3[3 * t + 1], m.n3[3 * t + 2]);
      // the plane through the vertices (the file normal may be anything)
      const dv3 c = cross(sub(tr.v1, tr.v0), sub(tr.v2, tr.v0));
      const double cl = len(c);
      tr.pn = (cl > 0.0 && std::isfinite(cl)) ? divs(c, cl) : zero3();
      tr.pd = dot(tr.pn, tr.v0);
      double vmax = 0.0;
      for (int q = 0; q < 9; ++q) vmax = std::max(vmax, std::fabs((double)v[q]));
      // the rounding of pn.x - pd and of any closest point is ~1e-16 of the
      // coordinates' size; 1e-9 of it leaves a wide margin
      tr.ptol = 1e-9 * (1.0 + vmax);
      tr.pad = 0.0;
      tris.push_back(tr);
    }
  }
  // bin: every cell whose (clamped) box meets the triangle's AABB grown by
  // r_max and comes within r_max of the triangle's plane (the edge cells of
  // each axis hold the particles clamped into them, so they reach to infinity
  // on their outer side and are kept)
  std::vector<uint32_t> counts((size_t)K.ncells + 1, 0);
  std::vector<std::pair<uint32_t, uint32_t>> entries;  // (cell, tri)
  const double grow = K.r_max * 1.0001 + 1e-12;
  auto clampc = [&](double x, int k) {
    double q = std::floor((x - K.go[k]) * K.ginv[k]);
    if (!(q == q)) return 0;
    if (q < 0) return 0;
    if (q >= K.gn[k]) return K.gn[k] - 1;
    return (int)q;
  };
  for (size_t t = 0; t < tris.size(); ++t) {
    const Tri& tr = tris[t];
    // The reference's closest_point clamps the edge parameters of regions
    // 2-6 to [0, a] or [0, c] (a = |e0|^2, c = |e1|^2; mesh.rs:67-105), not
    // [0, 1]: its point lies on v0 + [0, max(a, c)] e0, v0 + [0, c] e1 or
    // the triangle.  On edges longer than 1 m that reaches past the triangle,
    // so the binned box covers those segments too.
    const dv3 e0 = sub(tr.v1, tr.v0), e1 = sub(tr.v2, tr.v0);
    const double ea = dot(e0, e0), ec = dot(e1, e1);
    const dv3 x0 = add(tr.v0, smul(std::max(1.0, std::max(ea, ec)), e0));
    const dv3 x1 = add(tr.v0, smul(std::max(1.0, ec), e1));
    int c0[3], c1[3];
    for (int k = 0; k < 3; ++k) {
      double a = comp(tr.v0, k), b = comp(tr.v1, k), c = comp(tr.v2, k);
      double mn = std::min(a, std::min(b, c)), mx = std::max(a, std::max(b, c));
      if (std::isfinite(comp(x0, k)) && std::isfinite(comp(x1, k))) {
        mn = std::min(mn, std::min(comp(x0, k), comp(x1, k)));
        mx = std::max(mx, std::max(comp(x0, k), comp(x1, k)));
      }
      mn -= grow;
      mx += grow;
      c0[k] = clampc(mn, k);
      c1[k] = clampc(mx, k);
    }
    double ncov = (double)(c1[0] - c0[0] + 1) * (c1[1] - c0[1] + 1) * (c1[2] - c0[2] + 1);
    if (ncov > 2.0e8) return fail(RP_ERR_NOMEM, "triangle covers too many grid cells");
    const double pn[3] = {tr.pn.x, tr.pn.y, tr.pn.z};
    const bool planar = pn[0] != 0.0 || pn[1] != 0.0 || pn[2] != 0.0;
    for (int z = c0[2]; z <= c1[2]; ++z)
      for (int y = c0[1]; y <= c1[1]; ++y)
        for (int x = c0[0]; x <= c1[0]; ++x) {
          if (planar) {
            const int cc[3] = {x, y, z};
            double sdist = -tr.pd, reach = 0.0;
            bool open = false;
            for (int k = 0; k < 3; ++k) {
              if (pn[k] == 0.0) continue;
              if (cc[k] == 0 || cc[k] == K.gn[k] - 1) open = true;
              const double w = 1.0 / K.ginv[k];
              sdist += pn[k] * (K.go[k] + (cc[k] + 0.5) * w);
              reach += std::fabs(pn[k]) * 0.5 * w;
            }
            if (!open && std::fabs(sdist) > (reach + grow) * (1.0 + 1e-9) + 2.0 * tr.ptol) continue;
          }
          uint32_t c = cell_linear(K, x, y, z);
          entries.emplace_back(c, (uint32_t)t);
        }
  }
  std::stable_sort(entries.begin(), entries.end(),
                   [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                     return a.first < b.first;
                   });
  std::vector<uint32_t> start((size_t)K.ncells + 1, 0), idx(entries.size());
  for (auto& en : entries) counts[en.first]++;
  uint32_t run = 0;
  for (size_t c = 0; c < K.ncells; ++c) {
    start[c] = run;
    run += counts[c];
  }
  start[K.ncells] = run;
  for (size_t q = 0; q < entries.size(); ++q) idx[q] = entries[q].second;
  free_ptr(e->d_tris);
  free_ptr(e->d_tri_start);
  free_ptr(e->d_tri_idx);
  e->d_tris = nullptr;
  e->d_tri_start = nullptr;
  e->d_tri_idx = nullptr;
  HIP_TRY(hipMalloc(&e->d_tris, std::max<size_t>(1, tris.size()) * sizeof(Tri)));
  HIP_TRY(hipMalloc(&e->d_tri_start, start.size() * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&e->d_tri_idx, std::max<size_t>(1, idx.size()) * sizeof(uint32_t)));
  if (!tris.empty())
    HIP_TRY(hipMemcpyAsync(e->d_tris, tris.data(), tris.size() * sizeof(Tri), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipMemcpyAsync(e->d_tri_start, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                         e->stream));
  if (!idx.empty())
    HIP_TRY(hipMemcpyAsync(e->d_tri_idx, idx.data(), idx.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return RP_OK;
}

// Buffer carving: Dyn = xr (4 doubles per slot) then v w rx rv rw a tau
// (21 planes of doubles);
// Stat = r m (2 planes of doubles), flags (bytes), sid (u32).
size_t dyn_bytes(int cap) { return round_up((size_t)cap * 25 * sizeof(double), 256); }
size_t stat_bytes(int cap) {
  size_t c = (size_t)cap;
  return round_up(c * 2 * sizeof(double), 256) + round_up(c, 256) + round_up(c * sizeof(uint32_t), 256) +
         round_up(c * sizeof(uint16_t), 256);
}

Dyn carve_dyn(void* mem, int cap) {
  Dyn s{};
  size_t c = (size_t)cap;
  double* d = (double*)mem;
  s.xr = d;
  s.v = d + 4 * c;
  s.w = d + 7 * c;
  s.rx = d + 10 * c;
  s.rv = d + 13 * c;
  s.rw = d + 16 * c;
  s.a = d + 19 * c;
  s.tau = d + 22 * c;
  return s;
}

Stat carve_stat(void* mem, int cap) {
  Stat s{};
  char* p = (char*)mem;
  size_t c = (size_t)cap;
  s.r = (double*)p;
  s.m = (double*)p + c;
  p += round_up(c * 2 * sizeof(double), 256);
  s.flags = (uint8_t*)p;
  p += round_up(c, 256);
  s.sid = (uint32_t*)p;
  p += round_up(c * sizeof(uint32_t), 256);
  s.inst = (uint16_t*)p;
  return s;
}

void free_particles(rp_engine* e) {
  for (int b = 0; b < 2; ++b) {
    free_ptr(e->d_dyn_mem[b]);
    free_ptr(e->d_stat_mem[b]);
    e->d_dyn_mem[b] = nullptr;
    e->d_stat_mem[b] = nullptr;
  }
  free_ptr(e->d_xb);
  free_ptr(e->d_key);
  free_ptr(e->d_tmp);
  free_ptr(e->d_nbr_count);
  free_ptr(e->d_nbr);
  free_ptr(e->d_pairbuf);
  free_ptr(e->d_paircount);
  free_ptr(e->d_bstats);
  e->d_bstats = nullptr;
  for (int b = 0; b < 2; ++b) {
    free_ptr(e->d_hist_mem[b]);
    e->d_hist_mem[b] = nullptr;
    e->hist[b] = Hist{};
  }
  e->hcap = 0;
  e->kp.hcap = 0;
  free_ptr(e->d_surr_base);
  free_ptr(e->d_surr_rowmask);
  free_ptr(e->d_surr_nmask);
  free_ptr(e->d_sstats);
  e->d_surr_base = nullptr;
  e->d_surr_rowmask = e->d_surr_nmask = nullptr;
  e->d_sstats = nullptr;
  e->sstats_valid = false;
  e->d_xb = nullptr;
  e->d_key = e->d_tmp = nullptr;
  e->d_nbr_count = e->d_nbr = nullptr;
  e->d_pairbuf = nullptr;
  e->d_paircount = nullptr;
  e->cap = 0;
}

rp_status alloc_cells(rp_engine* e) {
  e->gen++;
  size_t need = (size_t)e->kp.ncells_all + 2;
  if (need <= e->cells_alloc) return RP_OK;
  free_ptr(e->d_cell_count);
  free_ptr(e->d_cell_start);
  free_ptr(e->d_block_sums);
  free_ptr(e->d_cell_rmax);
  e->d_cell_count = e->d_cell_start = e->d_block_sums = nullptr;
  e->d_cell_rmax = nullptr;
  HIP_TRY(hipMalloc(&e->d_cell_count, need * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&e->d_cell_start, need * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&e->d_cell_rmax, need * sizeof(float)));
  size_t nb = (need + kScanChunk - 1) / kScanChunk + 1;
  HIP_TRY(hipMalloc(&e->d_block_sums, nb * sizeof(uint32_t)));
  HIP_TRY(hipMemsetAsync(e->d_cell_count, 0, need * sizeof(uint32_t), e->stream));
  e->cells_alloc = need;
  return RP_OK;
}

// Moves the current history tables (hist[cur], old capacity) into a buffer of
// the new capacity: a list can hold at most nbr_cap contacts, so a table
// never needs more entries than the lists have.
__global__ void __launch_bounds__(kBlock) k_hist_regrow(Hist A, int hcap_a, Hist B, int hcap_b, int cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap) return;
  const int c = A.cnt[i];
  B.cnt[i] = c;
  const size_t sa = (size_t)hcap_a * cap, sb = (size_t)hcap_b * cap;
  for (int k = 0; k < c; ++k) {
    const size_t o = (size_t)k * cap + i;
    B.key[o] = A.key[o];
    B.xi[o] = A.xi[o];
    B.xi[o + sb] = A.xi[o + sa];
    B.xi[o + 2 * sb] = A.xi[o + 2 * sa];
  }
}

// (Re)allocates the history buffers with capacity hcap per slot, keeping the
// current tables (fresh buffers start empty).
rp_status alloc_hist(rp_engine* e, int hcap) {
  e->gen++;
  void* mem[2] = {nullptr, nullptr};
  const size_t bytes = hist_bytes(e->cap, hcap);
  for (int b = 0; b < 2; ++b) HIP_TRY(hipMalloc(&mem[b], bytes));
  Hist nh[2] = {carve_hist(mem[0], e->cap, hcap), carve_hist(mem[1], e->cap, hcap)};
  HIP_TRY(hipMemsetAsync(nh[0].cnt, 0, (size_t)e->cap * sizeof(int), e->stream));
  HIP_TRY(hipMemsetAsync(nh[1].cnt, 0, (size_t)e->cap * sizeof(int), e->stream));
  if (e->d_hist_mem[e->cur]) {
    const int c = e->cur;
    hipLaunchKernelGGL(k_hist_regrow, dim3((e->cap + kBlock - 1) / kBlock), dim3(kBlock), 0, e->stream, e->hist[c],
                       e->hcap, nh[c], hcap, e->cap);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (int b = 0; b < 2; ++b) {
    free_ptr(e->d_hist_mem[b]);
    e->d_hist_mem[b] = mem[b];
    e->hist[b] = nh[b];
  }
  e->hcap = hcap;
  e->kp.hcap = hcap;
  return RP_OK;
}

rp_status alloc_nbr(rp_engine* e, int nbr_cap) {
  // the kernels address list entry k of slot i as k * cap + i in 32 bits
  if ((size_t)nbr_cap * (size_t)e->cap >= ((size_t)1 << 31))
    return fail(RP_ERR_NOMEM, "neighbour lists would exceed 2^31 entries (" + std::to_string(nbr_cap) +
                                  " per particle x " + std::to_string(e->cap) + " slots)");
  e->gen++;
  free_ptr(e->d_nbr);
  e->d_nbr = nullptr;
  HIP_TRY(hipMalloc(&e->d_nbr, (size_t)nbr_cap * (size_t)e->cap * sizeof(int)));
  e->nbr_cap = nbr_cap;
  e->kp.nbr_cap = nbr_cap;
  if (e->hist_on) return alloc_hist(e, nbr_cap);
  return RP_OK;
}

int grid_for(size_t n) {
  size_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > (size_t)kMaxGrid) g = kMaxGrid;
  return (int)g;
}

// Launches rebuild chain number e->chain_no; it acts only if its request
// slot (chain_no & 1) is set, and it clears the other slot for the kernels
// that follow it.
rp_status launch_rebuild(rp_engine* e) {
  const int par = (int)(e->chain_no & 1);
  e->chain_no++;
  const KParams& K = e->kp;
  int g = grid_for((size_t)e->cap);
  int m = (int)K.ncells_all + 1;  // buckets incl. removed sentinel
  int nb = (m + kScanChunk - 1) / kScanChunk;
  int gs = std::min(nb, kMaxGrid);
  const Dyn W = e->dyn[e->cur], D = e->dyn[e->cur ^ 1];
  const Stat V = e->stat[0], U = e->stat[1];
  hipLaunchKernelGGL(k_keys, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, par, K, W, V, e->d_key,
                     e->d_cell_count);
  hipLaunchKernelGGL(k_scan_sums, dim3(gs), dim3(kBlock), 0, e->stream, e->d_ctl, par, e->d_cell_count, m,
                     e->d_block_sums);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, e->stream, e->d_ctl, par, e->d_block_sums, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(gs), dim3(kBlock), 0, e->stream, e->d_ctl, par, e->d_cell_count, m,
                     e->d_block_sums, e->d_cell_start);
  hipLaunchKernelGGL(k_scatter, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, par, e->d_key, e->d_cell_start,
                     e->d_cell_count, e->d_tmp);
  const Hist HS = e->hist_on ? e->hist[e->cur] : Hist{}, HD = e->hist_on ? e->hist[e->cur ^ 1] : Hist{};
  hipLaunchKernelGGL(k_rank_permute, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, par, K, W, V, D, U, e->d_key,
                     e->d_cell_start, e->d_tmp, HS, HD);
  if (K.cell_reach) {
    const int gc = std::min((int)((K.ncells_all + kBlock - 1) / kBlock), kMaxGrid);
    hipLaunchKernelGGL(k_cell_rmax, dim3(std::max(gc, 1)), dim3(kBlock), 0, e->stream, e->d_ctl, par, K, D,
                       e->d_cell_start, e->d_cell_rmax);
    hipLaunchKernelGGL(k_build<true>, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, par, K, D, U, W, V,
                       e->d_cell_start, e->d_cell_rmax, e->d_nbr_count, e->d_nbr, e->d_xb, HD, HS);
  } else {
    hipLaunchKernelGGL(k_build<false>, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, par, K, D, U, W, V,
                       e->d_cell_start, e->d_cell_rmax, e->d_nbr_count, e->d_nbr, e->d_xb, HD, HS);
  }
  HIP_TRY(hipGetLastError());
  return RP_OK;
}

rp_status read_ctl(rp_engine* e) {
  HIP_TRY(hipMemcpyAsync(&e->host_ctl, e->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  return RP_OK;
}

// Recompute geometry, rebin meshes and make sure the neighbour list is large
// enough; ends with a synchronous rebuild so capacity is known.
rp_status prepare(rp_engine* e) {
  if (!e->have_particles) return fail(RP_ERR_STATE, "no particles uploaded");
  if (!e->geometry_dirty) return RP_OK;
  const bool hist = e->prm.mode == RP_MODE_PHYSICS_FIX && e->prm.tangential_model != RP_TANGENTIAL_COULOMB;
  if (hist != e->hist_on) {
    e->hist_on = hist;
    if (hist) {
      if (rp_status st = alloc_hist(e, e->nbr_cap)) return st;
    } else {
      for (int b = 0; b < 2; ++b) {
        free_ptr(e->d_hist_mem[b]);
        e->d_hist_mem[b] = nullptr;
        e->hist[b] = Hist{};
      }
      e->hcap = 0;
    }
  } else if (hist && !e->d_hist_mem[0]) {
    if (rp_status st = alloc_hist(e, e->nbr_cap)) return st;
  }
  rp_status st = compute_geometry(e);
  if (st) return st;
  if ((st = alloc_cells(e))) return st;
  if ((st = upload_meshes_binned(e))) return st;
  // trial rebuild to size the neighbour lists; repeat while it overflows
  for (int attempt = 0; attempt < 8; ++attempt) {
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(1), 0, e->stream, e->d_ctl, (int)(e->chain_no & 1));
    if ((st = launch_rebuild(e))) return st;
    if ((st = read_ctl(e))) return st;
    if (!e->host_ctl.overflow) break;
    int want = (int)round_up((size_t)(e->host_ctl.max_count * 3 / 2 + 8), 8);
    int zero = 0;
    HIP_TRY(hipMemcpyAsync(&e->d_ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    if ((st = alloc_nbr(e, want))) return st;
  }
  if (e->host_ctl.overflow) return fail(RP_ERR_NOMEM, "neighbour list capacity could not be sized");
  // headroom for later rebuilds
  int want = (int)round_up((size_t)(e->host_ctl.max_count * 3 / 2 + 16), 8);
  if (want > e->nbr_cap) {
    if ((st = alloc_nbr(e, want))) return st;
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(1), 0, e->stream, e->d_ctl, (int)(e->chain_no & 1));
    if ((st = launch_rebuild(e))) return st;
  }
  // the list is valid now; a later k_begin decides about the next rebuild
  int clear[2] = {0, 0};
  HIP_TRY(hipMemcpyAsync(&e->d_ctl->rebuild[0], clear, sizeof(clear), hipMemcpyHostToDevice, e->stream));
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->geometry_dirty = false;
  return RP_OK;
}

rp_status ensure_pairs(rp_engine* e) {
  if (e->d_pairbuf && e->pair_cap >= e->nbr_cap + 16) return RP_OK;
  free_ptr(e->d_pairbuf);
  free_ptr(e->d_paircount);
  e->d_pairbuf = nullptr;
  e->d_paircount = nullptr;
  e->pair_cap = e->nbr_cap + 16;
  e->kp.pair_cap = e->pair_cap;
  HIP_TRY(hipMalloc(&e->d_pairbuf, (size_t)e->pair_cap * e->cap * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&e->d_paircount, (size_t)e->cap * sizeof(int)));
  HIP_TRY(hipMemsetAsync(e->d_paircount, 0, (size_t)e->cap * sizeof(int), e->stream));
  return RP_OK;
}

// kind: 0 = force kernel, 1 = rebuild chain, 2 = first half (k_begin),
// 3 = surrogate feature collection, 4 = surrogate network
hipEvent_t next_event(rp_engine* e) {
  if (e->ev_used == e->ev_pool.size()) {
    hipEvent_t ev;
    if (hipEventCreate(&ev) != hipSuccess) return nullptr;
    e->ev_pool.push_back(ev);
  }
  return e->ev_pool[e->ev_used++];
}

// Offset of the next launch's label from the device's label base.
unsigned long long label_off(const rp_engine* e) { return (unsigned long long)(e->launch_no - e->label_host); }

// Timing marks: an event pair around the launches of one kind.  A force
// launch (kind 0) is timed by its own dispatch (hipExtLaunchKernel records
// the pair at the kernel's start and end), so its time is the kernel's
// execution alone, without the gaps that separate event records add.
void mark_begin(rp_engine* e, int kind) {
  if (!e->timing) return;
  size_t idx = e->ev_used;
  hipEvent_t a = next_event(e);
  hipEvent_t b = next_event(e);
  if (!a || !b) return;
  if (kind == 0) {
    e->ext_start = a;
    e->ext_stop = b;
  } else {
    (void)hipEventRecord(a, e->stream);
  }
  e->ev_marks.push_back(rp_engine::Mark{kind, idx, 1});
}

void mark_end(rp_engine* e) {
  if (!e->timing || e->ev_marks.empty()) return;
  if (e->ev_marks.back().kind == 0) return;  // recorded by the launch itself
  (void)hipEventRecord(e->ev_pool[e->ev_marks.back().idx + 1], e->stream);
}

// Per-instance envelope table of a batched engine's surrogate: built here,
// before any launch of a step call, never while a stream capture records
// (surr_dev only reads it: an allocation or synchronous copy inside a
// capture would invalidate it).
rp_status ensure_surr_inst(rp_engine* e) {
  if (e->jkr_solver != RP_JKR_SURROGATE || e->ninst <= 1 || e->surr_inst_valid) return RP_OK;
  const rp_jkr_envelope& v = e->surr_env;
  const MlpScale sc = make_scale(v);
  auto inside = [](double x, double lo, double hi) { return x >= lo * (1.0 - 1e-12) && x <= hi * (1.0 + 1e-12); };
  std::vector<float4> tab((size_t)e->ninst);
  for (int q = 0; q < e->ninst; ++q) {
    const PairConsts c = material_consts(e->inst_particle[q], e->inst_wall[q]).pp;
    tab[q].x = inside(c.estar, v.estar_lo, v.estar_hi) && inside(c.gamma, v.gamma_lo, v.gamma_hi) ? 1.f : 0.f;
    tab[q].y = scale_in((float)std::log(c.estar), sc.in_lo[2], sc.in_hi[2]);
    tab[q].z = scale_in((float)std::log(std::max(c.gamma, 1e-300)), sc.in_lo[3], sc.in_hi[3]);
    tab[q].w = 0.f;
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  free_ptr(e->d_surr_inst);
  e->d_surr_inst = nullptr;
  HIP_TRY(hipMalloc(&e->d_surr_inst, tab.size() * sizeof(float4)));
  HIP_TRY(hipMemcpy(e->d_surr_inst, tab.data(), tab.size() * sizeof(float4), hipMemcpyHostToDevice));
  e->surr_inst_valid = true;
  e->gen++;  // captured loops hold the table's address
  return RP_OK;
}

// Device view of the surrogate state (null pointers when it is not in use).
// Reads only: ensure_surr_inst has built the per-instance table.
SurrDev surr_dev(rp_engine* e) {
  SurrDev sd{};
  if (e->jkr_solver != RP_JKR_SURROGATE) return sd;
  const rp_jkr_envelope& v = e->surr_env;
  sd.d_lo = v.d_lo;
  sd.d_hi = v.d_hi;
  sd.r_lo = v.rstar_lo;
  sd.r_hi = v.rstar_hi;
  const PairConsts& pp = e->kp.mc.pp;
  auto inside = [](double x, double lo, double hi) { return x >= lo * (1.0 - 1e-12) && x <= hi * (1.0 + 1e-12); };
  sd.pp_ok = inside(pp.estar, v.estar_lo, v.estar_hi) && inside(pp.gamma, v.gamma_lo, v.gamma_hi) ? 1 : 0;
  sd.check = e->surr_check;
  sd.sc = make_scale(v);
  sd.f_lne = scale_in((float)std::log(pp.estar), sd.sc.in_lo[2], sd.sc.in_hi[2]);
  sd.f_lng = scale_in((float)std::log(std::max(pp.gamma, 1e-300)), sd.sc.in_lo[3], sd.sc.in_hi[3]);
  if (e->ninst > 1) {
    sd.inst = e->d_surr_inst;
    if (!e->surr_inst_valid) sd.pp_ok = 0;  // not built: no instance is served
  }
  sd.base = e->d_surr_base;
  sd.rowmask = e->d_surr_rowmask;
  sd.nmask = e->d_surr_nmask;
  sd.md_lo = v.d_lo - kSurrMaskMargin;
  sd.md_hi = v.d_hi + kSurrMaskMargin;
  sd.x = e->d_surr_x;
  sd.y = e->d_surr_y;
  sd.counter = e->d_surr_counters;
  sd.rows_cap = e->surr_rows_cap;
  return sd;
}

rp_status alloc_surr_rows(rp_engine* e, int rows) {
  e->gen++;
  free_ptr(e->d_surr_x);
  free_ptr(e->d_surr_y);
  e->d_surr_x = e->d_surr_y = nullptr;
  e->surr_rows_cap = 0;
  HIP_TRY(hipMalloc(&e->d_surr_x, (size_t)rows * 4 * sizeof(float)));
  HIP_TRY(hipMalloc(&e->d_surr_y, (size_t)rows * 2 * sizeof(float)));
  e->surr_rows_cap = rows;
  return RP_OK;
}

bool surr_active(const rp_engine* e) { return e->jkr_solver == RP_JKR_SURROGATE && e->kp.normal == RP_NORMAL_JKR; }

// Feature rows of the surrogate-served contacts of the next force launch,
// then the network on them.
// The surrogate's two launches before a force launch: feature rows of the
// served contacts, then the network on them.  Timing mode brackets each.
void launch_surrogate(rp_engine* e) {
  const SurrDev sd = surr_dev(e);
  const int g = (int)((e->cap + kBlock - 1) / kBlock);
  mark_begin(e, 3);
  hipLaunchKernelGGL(k_jkr_collect, dim3(g), dim3(kBlock), 0, e->stream, e->d_ctl, label_off(e),
                     e->kp, e->dyn[e->cur], e->stat[0], e->d_nbr_count, e->d_nbr, sd);
  mark_end(e);
  hipEvent_t a = nullptr, b = nullptr;
  if (e->timing) {  // timed by its own dispatch, like the force kernel
    const size_t idx = e->ev_used;
    a = next_event(e);
    b = next_event(e);
    if (a && b) e->ev_marks.push_back(rp_engine::Mark{4, idx, 1});
    else a = b = nullptr;
  }
  mlp_infer_counted(e->surr_net.dev(), sd.x, sd.y, sd.counter, sd.rows_cap, e->stream, a, b);
}

template <int NORMAL, int TANG, bool ADV, bool REC, int JS, bool BATCH, bool DEAL, bool HIST>
void launch_force_d(rp_engine* e, int flag_slot) {
  int g = (int)((e->cap + kForceBlock - 1) / kForceBlock);
  const Hist hr = HIST ? e->hist[e->cur] : Hist{}, hw = HIST ? e->hist[e->cur ^ 1] : Hist{};
  if (e->ext_start) {
    // timing mode: the dispatch records the kernel's start and end
    hipExtLaunchKernelGGL((k_force<NORMAL, TANG, ADV, REC, JS, BATCH, DEAL, HIST>), dim3(g), dim3(kForceBlock), 0,
                          e->stream, e->ext_start, e->ext_stop, 0u, e->d_ctl, flag_slot, label_off(e), e->kp,
                          e->dyn[e->cur], e->dyn[e->cur ^ 1], e->stat[0], (const int*)e->d_nbr_count,
                          (const int*)e->d_nbr, (const Tri*)e->d_tris, (const uint32_t*)e->d_tri_start,
                          (const uint32_t*)e->d_tri_idx, (const float*)e->d_xb, e->d_bstats, e->d_pairbuf,
                          e->d_paircount, JS >= 2 ? surr_dev(e) : SurrDev{}, e->d_sstats, hr, hw);
    e->ext_start = e->ext_stop = nullptr;
  } else {
    hipLaunchKernelGGL((k_force<NORMAL, TANG, ADV, REC, JS, BATCH, DEAL, HIST>), dim3(g), dim3(kForceBlock), 0,
                       e->stream, e->d_ctl, flag_slot, label_off(e), e->kp, e->dyn[e->cur], e->dyn[e->cur ^ 1],
                       e->stat[0], e->d_nbr_count, e->d_nbr, e->d_tris, e->d_tri_start, e->d_tri_idx, e->d_xb,
                       e->d_bstats, e->d_pairbuf, e->d_paircount, JS >= 2 ? surr_dev(e) : SurrDev{}, e->d_sstats,
                       hr, hw);
  }
  e->force_grid = g;
  e->sstats_valid = JS >= 2;
  e->cur ^= 1;
  e->launch_no++;
}

// Kernels with a zero-force pass come in both tilings of it.
template <int NORMAL, int TANG, bool ADV, bool REC, int JS, bool BATCH, bool HIST>
void launch_force_t(rp_engine* e, int flag_slot) {
  if constexpr (split_pass(NORMAL, TANG, REC, JS, HIST)) {
    if (e->tiling == RP_TILING_WAVE) launch_force_d<NORMAL, TANG, ADV, REC, JS, BATCH, true, HIST>(e, flag_slot);
    else launch_force_d<NORMAL, TANG, ADV, REC, JS, BATCH, false, HIST>(e, flag_slot);
  } else {
    launch_force_d<NORMAL, TANG, ADV, REC, JS, BATCH, false, HIST>(e, flag_slot);
  }
}

// Single engines: the JKR solver variants (compiled for the JKR model only; a
// recording launch uses the exact solve with the shortcut, no surrogate rows).
template <int NORMAL, int TANG, bool ADV, bool REC, bool HIST>
void launch_force_s(rp_engine* e, int flag_slot) {
  if constexpr (NORMAL != RP_NORMAL_JKR) {
    launch_force_t<NORMAL, TANG, ADV, REC, 0, false, HIST>(e, flag_slot);
  } else {
    constexpr int kSurr = REC ? 1 : 2, kSurrCheck = REC ? 1 : 3;
    switch (e->jkr_solver) {
      case RP_JKR_SHORTCUT: launch_force_t<NORMAL, TANG, ADV, REC, 1, false, HIST>(e, flag_slot); break;
      case RP_JKR_SURROGATE:
        if (e->surr_check) launch_force_t<NORMAL, TANG, ADV, REC, kSurrCheck, false, HIST>(e, flag_slot);
        else launch_force_t<NORMAL, TANG, ADV, REC, kSurr, false, HIST>(e, flag_slot);
        break;
      default: launch_force_t<NORMAL, TANG, ADV, REC, 0, false, HIST>(e, flag_slot); break;
    }
  }
}

// Batched engines (several instances): the JKR solver variants with the
// per-instance material table.
template <int NORMAL, int TANG, bool ADV, bool HIST>
void launch_force_b(rp_engine* e, int flag_slot) {
  if constexpr (NORMAL == RP_NORMAL_JKR) {
    if (e->jkr_solver == RP_JKR_SURROGATE) {
      if (e->surr_check) launch_force_t<NORMAL, TANG, ADV, false, 3, true, HIST>(e, flag_slot);
      else launch_force_t<NORMAL, TANG, ADV, false, 2, true, HIST>(e, flag_slot);
      return;
    }
    if (e->jkr_solver == RP_JKR_SHORTCUT) {
      launch_force_t<NORMAL, TANG, ADV, false, 1, true, HIST>(e, flag_slot);
      return;
    }
  }
  launch_force_t<NORMAL, TANG, ADV, false, 0, true, HIST>(e, flag_slot);
}

// Batched engines have no pair recording; in the physics-fix mode they keep
// the tangential history as single engines do (stable ids are unique across
// instances, so they key the tables unchanged).
template <int NORMAL, int TANG, bool ADV, bool REC>
void launch_force_j(rp_engine* e, int flag_slot) {
  if constexpr (!REC) {
    if (e->ninst > 1) {
      if constexpr (TANG != RP_TANGENTIAL_COULOMB) {
        if (e->hist_on) {
          launch_force_b<NORMAL, TANG, ADV, true>(e, flag_slot);
          return;
        }
      }
      launch_force_b<NORMAL, TANG, ADV, false>(e, flag_slot);
      return;
    }
  }
  if constexpr (TANG != RP_TANGENTIAL_COULOMB) {
    if (e->hist_on) {
      launch_force_s<NORMAL, TANG, ADV, REC, true>(e, flag_slot);
      return;
    }
  }
  launch_force_s<NORMAL, TANG, ADV, REC, false>(e, flag_slot);
}

// Tangential kinds compiled: Coulomb, and the spring law (LSC / Mindlin) with
// the placeholder shear modulus.
template <int NORMAL, bool ADV, bool REC>
void launch_force_n(rp_engine* e, int slot) {
  switch (e->kp.tangential) {
    case RP_TANGENTIAL_COULOMB: launch_force_j<NORMAL, RP_TANGENTIAL_COULOMB, ADV, REC>(e, slot); break;
    case RP_TANGENTIAL_LSC: launch_force_j<NORMAL, RP_TANGENTIAL_LSC, ADV, REC>(e, slot); break;
    default: launch_force_j<NORMAL, RP_TANGENTIAL_MINDLIN, ADV, REC>(e, slot); break;
  }
}

template <bool ADV, bool REC>
void launch_force(rp_engine* e, int slot) {
  switch (e->kp.normal) {
    case RP_NORMAL_LINEAR: launch_force_n<RP_NORMAL_LINEAR, ADV, REC>(e, slot); break;
    case RP_NORMAL_HERTZ: launch_force_n<RP_NORMAL_HERTZ, ADV, REC>(e, slot); break;
    case RP_NORMAL_HYSTERETIC: launch_force_n<RP_NORMAL_HYSTERETIC, ADV, REC>(e, slot); break;
    case RP_NORMAL_JKR: launch_force_n<RP_NORMAL_JKR, ADV, REC>(e, slot); break;
    default: launch_force_n<RP_NORMAL_SJKR, ADV, REC>(e, slot); break;
  }
}

// Records kGraphSteps advancing launches (rebuild chain every
// kRebuildInterval launches, surrogate rows if active, force kernel) and the
// label advance into a graph.  Capture enqueues nothing: the host counters
// are restored afterwards and advanced by each replay instead.
rp_status capture_steps(rp_engine* e, hipGraphExec_t* exec) {
  const uint64_t launch0 = e->launch_no, chain0 = e->chain_no, label0 = e->label_host;
  const int cur0 = e->cur;
  e->label_host = e->launch_no;  // offsets 0 .. kGraphSteps-1 relative to the label at replay
  hipGraph_t graph = nullptr;
  HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  rp_status st = RP_OK;
  for (uint64_t s = 0; s < kGraphSteps && st == RP_OK; ++s) {
    if (e->launch_no % kRebuildInterval == 0) st = launch_rebuild(e);
    if (st == RP_OK && surr_active(e)) launch_surrogate(e);
    if (st == RP_OK) launch_force<true, false>(e, (int)(e->chain_no & 1));
  }
  hipLaunchKernelGGL(k_shift_labels, dim3(1), dim3(1), 0, e->stream, e->d_ctl, (unsigned long long)kGraphSteps);
  const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
  e->launch_no = launch0;
  e->chain_no = chain0;
  e->label_host = label0;
  e->cur = cur0;
  if (st) {
    if (graph) (void)hipGraphDestroy(graph);
    return st;
  }
  if (ce != hipSuccess) return fail(RP_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
  const hipError_t ie = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(RP_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  return RP_OK;
}

// One replay of the captured loop for the current request parity
// (captured first if what the launches bake in has changed).
rp_status replay_steps(rp_engine* e) {
  const int slot = (int)(e->chain_no & 1);
  if (!e->graph_exec[slot] || e->graph_gen[slot] != e->gen) {
    if (e->graph_exec[slot]) {
      // an earlier replay of this executable may still run on the stream
      HIP_TRY(hipStreamSynchronize(e->stream));
      (void)hipGraphExecDestroy(e->graph_exec[slot]);
    }
    e->graph_exec[slot] = nullptr;
    rp_status st = capture_steps(e, &e->graph_exec[slot]);
    if (st) return st;
    e->graph_gen[slot] = e->gen;
  }
  if (e->label_host != e->launch_no) {
    hipLaunchKernelGGL(k_shift_labels, dim3(1), dim3(1), 0, e->stream, e->d_ctl,
                       (unsigned long long)(e->launch_no - e->label_host));
    e->label_host = e->launch_no;
  }
  HIP_TRY(hipGraphLaunch(e->graph_exec[slot], e->stream));
  e->launch_no += kGraphSteps;
  e->chain_no += kGraphSteps / kRebuildInterval;
  e->label_host += kGraphSteps;
  e->graph_replays++;
  return RP_OK;
}

}  // namespace

// ---- C ABI ----------------------------------------------------------------------

extern "C" {

void rp_params_default(rp_params* p, double dt, const double bmin[3], const double bmax[3]) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->dt = dt;
  for (int k = 0; k < 3; ++k) {
    p->bounds_min[k] = bmin ? bmin[k] : 0.0;
    p->bounds_max[k] = bmax ? bmax[k] : 0.0;
  }
  p->normal_model = RP_NORMAL_HERTZ;
  p->tangential_model = RP_TANGENTIAL_MINDLIN;
  p->mode = RP_MODE_PARITY;
  p->particle = rp_material{1e7, 0.3, 2500.0, 0.5, 0.5, 0.0};
  p->wall = rp_material{1e9, 0.3, 7800.0, 0.5, 0.5, 0.0};
  p->gravity[0] = 0.0;
  p->gravity[1] = -9.81;
  p->gravity[2] = 0.0;
  p->ref_cell_size = 1.0;
}

int32_t rp_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* rp_last_error(void) { return g_last_error.c_str(); }

rp_status rp_engine_create(const rp_params* params, rp_engine** out) {
  if (!out) return fail(RP_ERR_INVALID, "out is NULL");
  *out = nullptr;
  rp_status st = validate(params);
  if (st) return st;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(RP_ERR_DEVICE, "no HIP device available");
  rp_engine* e = new rp_engine();
  e->prm = *params;
  HIP_TRY(hipGetDevice(&e->device));
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete e;
    return fail(RP_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  he = hipMalloc(&e->d_ctl, sizeof(Ctl));
  if (he != hipSuccess) {
    (void)hipStreamDestroy(e->stream);
    delete e;
    return fail(RP_ERR_NOMEM, "hipMalloc(ctl) failed");
  }
  Ctl c{};
  (void)hipMemcpy(e->d_ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice);
  *out = e;
  g_last_error.clear();
  return RP_OK;
}

rp_status rp_engine_set_params(rp_engine* e, const rp_params* params) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  rp_status st = validate(params);
  if (st) return st;
  e->prm = *params;
  e->skin_tuned = false;
  e->geometry_dirty = true;
  return RP_OK;
}

}  // extern "C"

namespace {

// Shared by the single and the batched upload: inst (NULL = all 0) gives
// each particle's instance.
rp_status upload_impl(rp_engine* e, size_t n, const double* x3, const double* v3, const double* w3, const double* a3,
                      const double* tau3, const double* rx3, const double* rv3, const double* rw3,
                      const double* radius, const double* mass, const uint8_t* fixed, const uint16_t* inst) {
  if (n > 0 && (!x3 || !radius || !mass)) return fail(RP_ERR_INVALID, "x3, radius and mass are required");
  if (n > (size_t)1 << 30) return fail(RP_ERR_INVALID, "too many particles");
  HIP_TRY(hipStreamSynchronize(e->stream));
  free_particles(e);
  e->gen++;
  e->label_host = 0;  // the control block below has no stall recorded
  int cap = (int)std::max<size_t>(round_up(n, 256), 256);
  e->cap = cap;
  for (int b = 0; b < 2; ++b) {
    HIP_TRY(hipMalloc(&e->d_dyn_mem[b], dyn_bytes(cap)));
    HIP_TRY(hipMalloc(&e->d_stat_mem[b], stat_bytes(cap)));
  }
  HIP_TRY(hipMalloc(&e->d_xb, (size_t)cap * 3 * sizeof(float)));
  HIP_TRY(hipMalloc(&e->d_key, (size_t)cap * sizeof(uint32_t)));
  HIP_TRY(hipMalloc(&e->d_tmp, (size_t)cap * sizeof(uint32_t)));
  // list length of slot i at [i], its near part (k_build) at [cap + i]
  HIP_TRY(hipMalloc(&e->d_nbr_count, (size_t)cap * 2 * sizeof(int)));
  HIP_TRY(hipMalloc(&e->d_bstats, ((size_t)cap / kForceBlock + 1) * sizeof(BlockStats)));
  HIP_TRY(hipMemset(e->d_bstats, 0, ((size_t)cap / kForceBlock + 1) * sizeof(BlockStats)));
  HIP_TRY(hipMalloc(&e->d_surr_base, (size_t)cap * sizeof(int)));
  if (!e->d_surr_counters)
    HIP_TRY(hipMalloc(&e->d_surr_counters, (size_t)kSurrRanges * kSurrCounterStride * sizeof(unsigned)));
  HIP_TRY(hipMemset(e->d_surr_counters, 0, (size_t)kSurrRanges * kSurrCounterStride * sizeof(unsigned)));
  HIP_TRY(hipMalloc(&e->d_surr_rowmask, (size_t)cap * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&e->d_surr_nmask, (size_t)cap * sizeof(unsigned long long)));
  HIP_TRY(hipMalloc(&e->d_sstats, ((size_t)cap / kForceBlock + 1) * sizeof(SurrStats)));
  HIP_TRY(hipMemset(e->d_sstats, 0, ((size_t)cap / kForceBlock + 1) * sizeof(SurrStats)));
  rp_status st = alloc_nbr(e, 32);
  if (st) return st;
  // host staging in the device layout: Dyn = xr records + 7 vec3 fields as
  // planes (25 doubles per slot), Stat = r and m planes
  const size_t cs = (size_t)cap;
  std::vector<double> dynh(cs * 25, 0.0), stath(cs * 2, 0.0);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) dynh[4 * i + k] = x3[3 * i + k];
    dynh[4 * i + 3] = radius[i];
  }
  auto put3 = [&](size_t first_plane, const double* src) {
    if (!src) return;
    for (size_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) dynh[(first_plane + k) * cs + i] = src[3 * i + k];
  };
  put3(4, v3);
  put3(7, w3);
  put3(10, rx3);
  put3(13, rv3);
  put3(16, rw3);
  put3(19, a3);
  put3(22, tau3);
  double rsum = 0.0, rmax = 0.0, vmax2 = 0.0;
  if (v3)
    for (size_t i = 0; i < n; ++i) {
      double q = v3[3 * i] * v3[3 * i] + v3[3 * i + 1] * v3[3 * i + 1] + v3[3 * i + 2] * v3[3 * i + 2];
      if (q > vmax2) vmax2 = q;
    }
  e->v_max = std::sqrt(vmax2);
  for (size_t i = 0; i < n; ++i) {
    stath[i] = radius[i];
    stath[cs + i] = mass[i];
    rsum += radius[i];
    if (radius[i] > rmax) rmax = radius[i];
  }
  std::vector<uint8_t> fl((size_t)cap, 0);
  std::vector<uint32_t> sid((size_t)cap, 0);
  for (size_t i = 0; i < n; ++i) {
    fl[i] = (fixed && fixed[i]) ? FLAG_FIXED : 0;
    sid[i] = (uint32_t)i;
  }
  Dyn d0 = carve_dyn(e->d_dyn_mem[0], cap);
  Stat s0 = carve_stat(e->d_stat_mem[0], cap);
  HIP_TRY(hipMemcpy(d0.xr, dynh.data(), dynh.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s0.r, stath.data(), stath.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s0.flags, fl.data(), (size_t)cap, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(s0.sid, sid.data(), (size_t)cap * sizeof(uint32_t), hipMemcpyHostToDevice));
  std::vector<uint16_t> ins((size_t)cap, 0);
  if (inst) std::copy(inst, inst + n, ins.begin());
  HIP_TRY(hipMemcpy(s0.inst, ins.data(), (size_t)cap * sizeof(uint16_t), hipMemcpyHostToDevice));
  e->dyn[0] = d0;
  e->dyn[1] = carve_dyn(e->d_dyn_mem[1], cap);
  e->stat[0] = s0;
  e->stat[1] = carve_stat(e->d_stat_mem[1], cap);
  e->cur = 0;
  e->force_grid = 0;
  Ctl c{};
  c.n = (int)n;
  c.alive = (int)n;
  c.stall_off = kNoStall;
  c.at_rest = 1;
  HIP_TRY(hipMemcpy(e->d_ctl, &c, sizeof(Ctl), hipMemcpyHostToDevice));
  e->n_uploaded = (int)n;
  e->steps_done = 0;
  e->at_rest = true;
  e->r_max = rmax;
  e->r_mean = n ? rsum / (double)n : 0.0;
  e->have_particles = true;
  e->geometry_dirty = true;
  e->pairs_valid = false;
  return RP_OK;
}

}  // namespace

extern "C" {

rp_status rp_engine_upload_full(rp_engine* e, size_t n, const double* x3, const double* v3, const double* w3,
                                const double* a3, const double* tau3, const double* rx3, const double* rv3,
                                const double* rw3, const double* radius, const double* mass,
                                const uint8_t* fixed) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  e->ninst = 1;
  e->inst_particle.clear();
  e->inst_wall.clear();
  e->inst_count.assign(1, n);
  return upload_impl(e, n, x3, v3, w3, a3, tau3, rx3, rv3, rw3, radius, mass, fixed, nullptr);
}

rp_status rp_engine_upload_batch(rp_engine* e, int32_t ninst, const size_t* counts, const rp_material* particle,
                                 const rp_material* wall, const double* x3, const double* v3, const double* w3,
                                 const double* a3, const double* tau3, const double* rx3, const double* rv3,
                                 const double* rw3, const double* radius, const double* mass,
                                 const uint8_t* fixed) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (ninst < 1 || ninst > 65535) return fail(RP_ERR_INVALID, "ninst must lie in [1, 65535]");
  if (!counts || !particle) return fail(RP_ERR_INVALID, "counts and particle materials are required");
  if (e->record && ninst > 1) return fail(RP_ERR_STATE, "pair recording is not available for batched engines");
  size_t n = 0;
  for (int q = 0; q < ninst; ++q) n += counts[q];
  std::vector<uint16_t> inst(n);
  size_t o = 0;
  for (int q = 0; q < ninst; ++q)
    for (size_t k = 0; k < counts[q]; ++k) inst[o++] = (uint16_t)q;
  e->ninst = ninst;
  e->inst_particle.assign(particle, particle + ninst);
  e->inst_wall.assign((size_t)ninst, e->prm.wall);
  if (wall) e->inst_wall.assign(wall, wall + ninst);
  e->inst_count.assign(counts, counts + ninst);
  e->surr_inst_valid = false;
  return upload_impl(e, n, x3, v3, w3, a3, tau3, rx3, rv3, rw3, radius, mass, fixed, inst.data());
}

rp_status rp_engine_instance_counts(rp_engine* e, size_t* counts) {
  if (!e || !counts) return fail(RP_ERR_INVALID, "NULL argument");
  for (int q = 0; q < e->ninst; ++q) counts[q] = 0;
  if (!e->have_particles) return RP_OK;
  rp_status st = read_ctl(e);
  if (st) return st;
  const int n = e->host_ctl.n;
  std::vector<uint16_t> ins((size_t)n);
  std::vector<uint8_t> fl((size_t)n);
  HIP_TRY(hipMemcpy(ins.data(), e->stat[0].inst, (size_t)n * sizeof(uint16_t), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(fl.data(), e->stat[0].flags, (size_t)n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; ++i)
    if (!(fl[i] & FLAG_DEAD)) counts[ins[i]]++;
  return RP_OK;
}

rp_status rp_engine_upload(rp_engine* e, size_t n, const double* x3, const double* v3, const double* w3,
                           const double* a3, const double* tau3, const double* radius, const double* mass,
                           const uint8_t* fixed) {
  return rp_engine_upload_full(e, n, x3, v3, w3, a3, tau3, nullptr, nullptr, nullptr, radius, mass, fixed);
}

rp_status rp_engine_add_mesh(rp_engine* e, size_t ntri, const float* v9, const float* n3) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (ntri > 0 && (!v9 || !n3)) return fail(RP_ERR_INVALID, "v9 and n3 are required");
  HostMesh m;
  m.v9.assign(v9, v9 + ntri * 9);
  m.n3.assign(n3, n3 + ntri * 3);
  e->meshes.push_back(std::move(m));
  e->geometry_dirty = true;
  return RP_OK;
}

rp_status rp_engine_set_mesh_motion(rp_engine* e, int32_t mesh, const double* origin, const double* omega) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (!origin || !omega) return fail(RP_ERR_INVALID, "origin and omega are required");
  if (mesh < 0 || (size_t)mesh >= e->meshes.size()) return fail(RP_ERR_INVALID, "no such mesh");
  for (int k = 0; k < 3; ++k)
    if (!std::isfinite(origin[k]) || !std::isfinite(omega[k]))
      return fail(RP_ERR_INVALID, "origin and omega must be finite");
  HostMesh& m = e->meshes[(size_t)mesh];
  const HostMesh old = m;
  for (int k = 0; k < 3; ++k) {
    m.origin[k] = origin[k];
    m.omega[k] = omega[k];
  }
  if (rp_status st = fill_motion(e)) {
    m = old;
    (void)fill_motion(e);
    return st;
  }
  return RP_OK;
}

rp_status rp_engine_add_mesh_stl(rp_engine* e, const char* path) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  std::vector<float> v9, n3;
  std::string err;
  if (!rp_stl_load(path ? path : "", v9, n3, err)) return fail(RP_ERR_IO, err);
  return rp_engine_add_mesh(e, n3.size() / 3, v9.data(), n3.data());
}

rp_status rp_engine_set_record(rp_engine* e, int on) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (on && e->ninst > 1) return fail(RP_ERR_STATE, "pair recording is not available for batched engines");
  e->record = on != 0;
  return RP_OK;
}

constexpr uint64_t kGuardChunk = 200;  // steps between checks of a tuned skin

// The autotuner measures the motion of its moment: particles released at
// rest make a thin skin look best, and once they move it rebuilds the lists
// at nearly every check.  After a tuned skin has run kGuardChunk steps or
// more, if the lists were rebuilt at more than a quarter of the checks (one
// every kRebuildInterval launches), the skin is doubled, up to the automatic
// skin, and the lists are rebuilt with it.  Skin never changes a result.
void skin_guard(rp_engine* e, uint64_t executed) {
  if (!e->skin_tuned) return;
  e->guard_steps += executed;
  if (e->guard_steps < kGuardChunk) return;
  const unsigned long long r = e->host_ctl.rebuilds - e->guard_r0;
  const double auto_skin = std::max(0.1 * e->r_max, 4.0 * (double)kRebuildInterval * e->v_max * e->prm.dt);
  if (r * 4ull * (unsigned long long)kRebuildInterval > e->guard_steps && e->prm.skin < auto_skin) {
    e->prm.skin = std::min(2.0 * e->prm.skin, auto_skin);
    e->geometry_dirty = true;
  }
  e->guard_steps = 0;
  e->guard_r0 = e->host_ctl.rebuilds;
}

rp_status rp_engine_step(rp_engine* e, uint64_t nsteps) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (nsteps == 0) return RP_OK;
  rp_status st = prepare(e);
  if (st) return st;
  if ((st = ensure_surr_inst(e))) return st;
  const int gcap = (int)((e->cap + kBlock - 1) / kBlock);
  uint64_t remaining = nsteps;
  int stalled = 0;  // rounds that ended with stalled launches
  while (remaining > 0) {
    if (stalled > 64) return fail(RP_ERR_STATE, "time step made no progress after 64 neighbour-list rebuilds");
    if (e->record && (st = ensure_pairs(e))) return st;
    if (e->at_rest) {
      mark_begin(e, 2);
      hipLaunchKernelGGL(k_begin, dim3(gcap), dim3(kBlock), 0, e->stream, e->d_ctl, (int)(e->chain_no & 1),
                         e->kp, e->dyn[e->cur], e->stat[0], e->d_xb, label_off(e));
      mark_end(e);
    }
    // A tuned skin is checked every kGuardChunk steps (skin_guard).
    const uint64_t chunk = e->skin_tuned ? std::min<uint64_t>(remaining, kGuardChunk) : remaining;
    const bool final_chunk = chunk == remaining;
    for (uint64_t s = 0; s < chunk; ++s) {
      // Whole captured loops where they fit (aligned to the rebuild cadence,
      // never the call's last launch, which leaves the at-rest state).
      if (e->use_graph && !e->timing && s > 0 && e->launch_no % kGraphSteps == 0 &&
          (final_chunk ? chunk - s > kGraphSteps : chunk - s >= kGraphSteps)) {
        if ((st = replay_steps(e))) return st;
        s += kGraphSteps - 1;
        continue;
      }
      // The rebuild chain runs at the first launch of a round and every
      // kRebuildInterval launches; the force kernels' look-ahead covers the gap.
      if (s == 0 || e->launch_no % kRebuildInterval == 0) {
        mark_begin(e, 1);
        if ((st = launch_rebuild(e))) return st;
        mark_end(e);
      }
      const int slot = (int)(e->chain_no & 1);
      const bool last = final_chunk && s + 1 == chunk;
      if (surr_active(e) && !(last && e->record)) launch_surrogate(e);
      mark_begin(e, 0);
      if (last) {
        if (e->record) launch_force<false, true>(e, slot);
        else launch_force<false, false>(e, slot);
      } else {
        launch_force<true, false>(e, slot);
      }
      mark_end(e);
    }
    HIP_TRY(hipGetLastError());
    // One sync per round (the call, or a chunk of it while a tuned skin is
    // checked): count the launches that advanced the state.  Launches
    // stalled by a skin violation or a list overflow are issued again.
    HIP_TRY(hipMemcpyAsync(&e->host_ctl, e->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const uint64_t executed = e->host_ctl.steps_done - e->steps_done;
    e->steps_done = e->host_ctl.steps_done;
    if (surr_active(e) && e->host_ctl.surr_rows_max > (unsigned)(e->surr_rows_cap / kSurrRanges)) {
      // some particles' contacts took the exact solve for want of rows: grow
      const size_t want = round_up((size_t)e->host_ctl.surr_rows_max * kSurrRanges * 3 / 2 + 1024,
                                   (size_t)64 * kSurrRanges);
      HIP_TRY(hipStreamSynchronize(e->stream));
      if ((st = alloc_surr_rows(e, (int)std::min<size_t>(want, (size_t)1 << 30)))) return st;
    }
    e->at_rest = e->host_ctl.at_rest != 0;
    remaining -= executed;
    skin_guard(e, executed);
    if (remaining == 0) break;
    if (executed == chunk) {
      // the chunk ran whole: go on (with the lists of a widened skin)
      if (e->geometry_dirty && (st = prepare(e))) return st;
      continue;
    }
    ++stalled;
    if (e->host_ctl.overflow) {
      // a rebuild found more candidates than the lists hold: grow them
      int want = (int)round_up((size_t)(e->host_ctl.max_count * 3 / 2 + 16), 8);
      if ((st = alloc_nbr(e, std::max(want, e->nbr_cap + 8)))) return st;
      int zero = 0;
      HIP_TRY(hipMemcpyAsync(&e->d_ctl->overflow, &zero, sizeof(int), hipMemcpyHostToDevice, e->stream));
    } else if (executed == 0) {
      // a particle crosses half the skin within one step: widen the skin
      e->prm.skin = 2.0 * (e->prm.skin > 0.0 ? e->prm.skin : e->kp.skin);
      e->geometry_dirty = true;
      if ((st = prepare(e))) return st;
    }
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(1), 0, e->stream, e->d_ctl, (int)(e->chain_no & 1));
  }
  e->step_count += nsteps;
  e->pairs_valid = e->record;
  return RP_OK;
}

rp_status rp_engine_synchronize(rp_engine* e) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  HIP_TRY(hipStreamSynchronize(e->stream));
  return RP_OK;
}

rp_status rp_engine_count(rp_engine* e, size_t* n) {
  if (!e || !n) return fail(RP_ERR_INVALID, "NULL argument");
  if (!e->have_particles) {
    *n = 0;
    return RP_OK;
  }
  rp_status st = read_ctl(e);
  if (st) return st;
  *n = (size_t)e->host_ctl.alive;
  return RP_OK;
}

namespace {
// Slot order -> API order (ascending stable id among alive particles).
rp_status slot_order(rp_engine* e, std::vector<int>& order, Dyn& cur, Stat& cst, int& nslots) {
  rp_status st = read_ctl(e);
  if (st) return st;
  cur = e->dyn[e->cur];
  cst = e->stat[0];
  nslots = e->host_ctl.n;
  std::vector<uint32_t> sid((size_t)nslots);
  std::vector<uint8_t> fl((size_t)nslots);
  if (nslots > 0) {
    HIP_TRY(hipMemcpy(sid.data(), cst.sid, (size_t)nslots * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(fl.data(), cst.flags, (size_t)nslots, hipMemcpyDeviceToHost));
  }
  order.clear();
  for (int i = 0; i < nslots; ++i)
    if (!(fl[i] & FLAG_DEAD)) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return sid[a] < sid[b]; });
  return RP_OK;
}

rp_status fetch3(const double* dev, int cap, int nslots, const std::vector<int>& order, double* out) {
  if (!out) return RP_OK;
  std::vector<double> tmp((size_t)cap * 3);
  if (nslots > 0) {
    for (int k = 0; k < 3; ++k)
      HIP_TRY(hipMemcpy(tmp.data() + (size_t)k * cap, dev + (size_t)k * cap, (size_t)nslots * sizeof(double),
                        hipMemcpyDeviceToHost));
  }
  for (size_t q = 0; q < order.size(); ++q)
    for (int k = 0; k < 3; ++k) out[3 * q + k] = tmp[(size_t)k * cap + order[q]];
  return RP_OK;
}
}  // namespace

rp_status rp_engine_download(rp_engine* e, double* x3, double* v3, double* w3, double* a3, double* tau3) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (!e->have_particles) return RP_OK;
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  rp_status st = slot_order(e, order, cur, cst, ns);
  if (st) return st;
  if (x3 && ns > 0) {
    std::vector<double> tmp((size_t)ns * 4);
    HIP_TRY(hipMemcpy(tmp.data(), cur.xr, tmp.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < order.size(); ++q)
      for (int k = 0; k < 3; ++k) x3[3 * q + k] = tmp[4 * (size_t)order[q] + k];
  }
  if ((st = fetch3(cur.v, e->cap, ns, order, v3))) return st;
  if ((st = fetch3(cur.w, e->cap, ns, order, w3))) return st;
  if ((st = fetch3(cur.a, e->cap, ns, order, a3))) return st;
  if ((st = fetch3(cur.tau, e->cap, ns, order, tau3))) return st;
  return RP_OK;
}

rp_status rp_engine_download_aux(rp_engine* e, double* rx3, double* rv3, double* rw3, double* radius,
                                 double* mass, uint8_t* fixed) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (!e->have_particles) return RP_OK;
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  rp_status st = slot_order(e, order, cur, cst, ns);
  if (st) return st;
  if ((st = fetch3(cur.rx, e->cap, ns, order, rx3))) return st;
  if ((st = fetch3(cur.rv, e->cap, ns, order, rv3))) return st;
  if ((st = fetch3(cur.rw, e->cap, ns, order, rw3))) return st;
  std::vector<double> tmp((size_t)std::max(ns, 1));
  std::vector<uint8_t> fl((size_t)std::max(ns, 1));
  if (radius && ns > 0) {
    HIP_TRY(hipMemcpy(tmp.data(), cst.r, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < order.size(); ++q) radius[q] = tmp[order[q]];
  }
  if (mass && ns > 0) {
    HIP_TRY(hipMemcpy(tmp.data(), cst.m, (size_t)ns * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < order.size(); ++q) mass[q] = tmp[order[q]];
  }
  if (fixed && ns > 0) {
    HIP_TRY(hipMemcpy(fl.data(), cst.flags, (size_t)ns, hipMemcpyDeviceToHost));
    for (size_t q = 0; q < order.size(); ++q) fixed[q] = (fl[order[q]] & FLAG_FIXED) ? 1 : 0;
  }
  return RP_OK;
}

rp_status rp_engine_history(rp_engine* e, int64_t* offsets, int64_t* partner, double* xi3, size_t cap,
                            size_t* total) {
  if (!e || !total) return fail(RP_ERR_INVALID, "NULL argument");
  *total = 0;
  if (!e->have_particles || !e->hist_on || !e->d_hist_mem[0]) {
    if (offsets) {
      size_t n = 0;
      if (e->have_particles) {
        rp_status st = rp_engine_count(e, &n);
        if (st) return st;
      }
      for (size_t q = 0; q <= n; ++q) offsets[q] = 0;
    }
    return RP_OK;
  }
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  rp_status st = slot_order(e, order, cur, cst, ns);
  if (st) return st;
  const Hist& H = e->hist[e->cur];
  const size_t c = (size_t)e->cap, h = (size_t)e->hcap;
  std::vector<int> cnt((size_t)std::max(ns, 1));
  std::vector<uint32_t> sid((size_t)std::max(ns, 1)), key(h * c);
  std::vector<double> xi(3 * h * c);
  if (ns > 0) {
    HIP_TRY(hipMemcpy(cnt.data(), H.cnt, (size_t)ns * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(sid.data(), cst.sid, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(key.data(), H.key, key.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(xi.data(), H.xi, xi.size() * sizeof(double), hipMemcpyDeviceToHost));
  }
  std::vector<std::pair<uint32_t, int64_t>> api;
  api.reserve(order.size());
  for (size_t q = 0; q < order.size(); ++q) api.emplace_back(sid[order[q]], (int64_t)q);
  size_t tot = 0;
  for (size_t q = 0; q < order.size(); ++q) tot += (size_t)cnt[order[q]];
  *total = tot;
  if (!offsets || !partner || !xi3) return RP_OK;
  if (cap < tot) return fail(RP_ERR_INVALID, "history buffer too small");
  size_t o = 0;
  for (size_t q = 0; q < order.size(); ++q) {
    const size_t i = (size_t)order[q];
    offsets[q] = (int64_t)o;
    for (int k = 0; k < cnt[i]; ++k, ++o) {
      const size_t b = (size_t)k * c + i;
      auto it = std::lower_bound(api.begin(), api.end(), std::make_pair(key[b], (int64_t)-1));
      partner[o] = (it != api.end() && it->first == key[b]) ? it->second : -1;
      xi3[3 * o] = xi[b];
      xi3[3 * o + 1] = xi[b + h * c];
      xi3[3 * o + 2] = xi[b + 2 * h * c];
    }
  }
  offsets[order.size()] = (int64_t)o;
  return RP_OK;
}

rp_status rp_engine_set_history(rp_engine* e, const int64_t* offsets, const int64_t* partner, const double* xi3,
                                size_t total) {
  if (!e || !offsets) return fail(RP_ERR_INVALID, "NULL argument");
  if (!e->have_particles) return fail(RP_ERR_STATE, "no particles uploaded");
  if (total > 0 && (!partner || !xi3)) return fail(RP_ERR_INVALID, "partner and xi3 are required");
  // sizes the lists (and with them the tables) and sorts the slots, as the
  // first step would; a later step finds the engine prepared
  rp_status st = prepare(e);
  if (st) return st;
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  if ((st = slot_order(e, order, cur, cst, ns))) return st;
  const size_t n = order.size();
  if (offsets[0] != 0 || offsets[n] < 0 || (size_t)offsets[n] != total)
    return fail(RP_ERR_INVALID, "offsets must run from 0 to total");
  for (size_t q = 0; q < n; ++q)
    if (offsets[q + 1] < offsets[q]) return fail(RP_ERR_INVALID, "offsets must not decrease");
  if (!e->hist_on) {
    if (total > 0) return fail(RP_ERR_STATE, "no tangential history is kept (parity mode or Coulomb friction)");
    return RP_OK;
  }
  const size_t c = (size_t)e->cap, h = (size_t)e->hcap;
  std::vector<uint32_t> sid((size_t)std::max(ns, 1));
  if (ns > 0) HIP_TRY(hipMemcpy(sid.data(), cst.sid, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost));
  std::vector<int> cnt(c, 0);
  std::vector<uint32_t> key(h * c, 0u);
  std::vector<double> xi(3 * h * c, 0.0);
  for (size_t q = 0; q < n; ++q) {
    const size_t i = (size_t)order[q];
    size_t k = 0;
    int64_t prev = -1;
    for (int64_t o = offsets[q]; o < offsets[q + 1]; ++o) {
      const int64_t pj = partner[o];
      if (pj < 0) continue;  // a partner removed since the history was read
      if ((size_t)pj >= n || (size_t)pj == q) return fail(RP_ERR_INVALID, "history partner out of range");
      if (pj <= prev) return fail(RP_ERR_INVALID, "history partners must ascend");
      prev = pj;
      if (k >= h) return fail(RP_ERR_INVALID, "more history entries than the neighbour lists hold");
      // API order is ascending stable id, so the keys ascend as the kernel's cursor needs
      const size_t b = k * c + i;
      key[b] = sid[(size_t)order[(size_t)pj]];
      xi[b] = xi3[3 * o];
      xi[b + h * c] = xi3[3 * o + 1];
      xi[b + 2 * h * c] = xi3[3 * o + 2];
      ++k;
    }
    cnt[i] = (int)k;
  }
  const Hist& H = e->hist[e->cur];
  HIP_TRY(hipMemcpy(H.cnt, cnt.data(), c * sizeof(int), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(H.key, key.data(), key.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(H.xi, xi.data(), xi.size() * sizeof(double), hipMemcpyHostToDevice));
  return RP_OK;
}

rp_status rp_engine_cells(rp_engine* e, int32_t rebuild, int64_t* ref_cell, uint32_t* fine_key, int64_t* slot_api,
                          double* grid) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (!e->have_particles) return fail(RP_ERR_STATE, "no particles uploaded");
  rp_status st = prepare(e);
  if (st) return st;
  if (rebuild) {
    // a rebuild never changes a result: the lists only have to be supersets
    hipLaunchKernelGGL(k_request_rebuild, dim3(1), dim3(1), 0, e->stream, e->d_ctl, (int)(e->chain_no & 1));
    if ((st = launch_rebuild(e))) return st;
    int clear[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&e->d_ctl->rebuild[0], clear, sizeof(clear), hipMemcpyHostToDevice, e->stream));
  }
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  if ((st = slot_order(e, order, cur, cst, ns))) return st;
  long long* d_ref = nullptr;
  uint32_t* d_key = nullptr;
  HIP_TRY(hipMalloc(&d_ref, (size_t)std::max(ns, 1) * sizeof(long long)));
  HIP_TRY(hipMalloc(&d_key, (size_t)std::max(ns, 1) * sizeof(uint32_t)));
  std::vector<long long> ref((size_t)std::max(ns, 1));
  std::vector<uint32_t> key((size_t)std::max(ns, 1));
  std::vector<uint32_t> sid((size_t)std::max(ns, 1));
  if (ns > 0) {
    hipLaunchKernelGGL(k_cells, dim3((ns + kBlock - 1) / kBlock), dim3(kBlock), 0, e->stream, e->d_ctl, e->kp, cur,
                       cst, d_ref, d_key);
    HIP_TRY(hipMemcpyAsync(ref.data(), d_ref, (size_t)ns * sizeof(long long), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(key.data(), d_key, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipMemcpyAsync(sid.data(), cst.sid, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
  }
  (void)hipFree(d_ref);
  (void)hipFree(d_key);
  // API index of each slot (order lists the alive slots by stable id)
  std::vector<int64_t> api((size_t)std::max(ns, 1), -1);
  for (size_t q = 0; q < order.size(); ++q) api[(size_t)order[q]] = (int64_t)q;
  for (size_t q = 0; q < order.size(); ++q) {
    if (ref_cell) ref_cell[q] = ref[(size_t)order[q]];
    if (fine_key) fine_key[q] = key[(size_t)order[q]];
  }
  if (slot_api) {
    size_t o = 0;
    for (int i = 0; i < ns; ++i)
      if (api[(size_t)i] >= 0) slot_api[o++] = api[(size_t)i];
  }
  if (grid) {
    const KParams& K = e->kp;
    for (int k = 0; k < 3; ++k) {
      grid[k] = K.go[k];
      grid[3 + k] = K.ginv[k];
      grid[6 + k] = (double)K.gn[k];
      grid[9 + k] = (double)K.gsh[k];
      grid[12 + k] = (double)K.ghi[k];
      grid[15 + k] = (double)K.glo[k];
      grid[18 + k] = (double)K.periodic[k];
    }
    grid[21] = (double)K.ncells;
  }
  return RP_OK;
}

rp_status rp_engine_pairs(rp_engine* e, int64_t* offsets, int64_t* js, size_t cap, size_t* total) {
  if (!e || !total) return fail(RP_ERR_INVALID, "NULL argument");
  if (!e->pairs_valid) return fail(RP_ERR_STATE, "no recorded step: call rp_engine_set_record(e,1) and step");
  std::vector<int> order;
  Dyn cur;
  Stat cst;
  int ns;
  rp_status st = slot_order(e, order, cur, cst, ns);
  if (st) return st;
  if (e->host_ctl.pair_overflow) return fail(RP_ERR_NOMEM, "pair record overflow");
  std::vector<uint32_t> sid((size_t)std::max(ns, 1));
  std::vector<int> pc((size_t)std::max(ns, 1));
  std::vector<unsigned long long> pb((size_t)std::max(ns, 1) * e->pair_cap);
  if (ns > 0) {
    HIP_TRY(hipMemcpy(sid.data(), cst.sid, (size_t)ns * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pc.data(), e->d_paircount, (size_t)ns * sizeof(int), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(pb.data(), e->d_pairbuf, (size_t)ns * e->pair_cap * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost));
  }
  // stable id -> API index
  std::vector<std::pair<uint32_t, int64_t>> api;
  api.reserve(order.size());
  for (size_t q = 0; q < order.size(); ++q) api.emplace_back(sid[order[q]], (int64_t)q);
  auto to_api = [&](uint32_t s) -> int64_t {
    auto it = std::lower_bound(api.begin(), api.end(), std::make_pair(s, (int64_t)-1));
    return (it != api.end() && it->first == s) ? it->second : -1;
  };
  size_t tot = 0;
  for (size_t q = 0; q < order.size(); ++q) tot += (size_t)pc[order[q]];
  *total = tot;
  if (offsets == nullptr || js == nullptr) return RP_OK;
  if (cap < tot) return fail(RP_ERR_INVALID, "pair buffer too small");
  size_t o = 0;
  for (size_t q = 0; q < order.size(); ++q) {
    int slot = order[q];
    offsets[q] = (int64_t)o;
    std::vector<unsigned long long> ent(pb.begin() + (size_t)slot * e->pair_cap,
                                        pb.begin() + (size_t)slot * e->pair_cap + pc[slot]);
    std::sort(ent.begin(), ent.end());  // (rank of the 27-cell walk, stable id) = reference order
    for (unsigned long long v : ent) js[o++] = to_api((uint32_t)(v & 0xffffffffu));
  }
  offsets[order.size()] = (int64_t)o;
  return RP_OK;
}

rp_status rp_engine_stats(rp_engine* e, rp_stats* out) {
  if (!e || !out) return fail(RP_ERR_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  if (!e->have_particles) return RP_OK;
  rp_status st = read_ctl(e);
  if (st) return st;
  const Ctl& c = e->host_ctl;
  out->steps = e->host_ctl.steps_done;
  out->rebuilds = c.rebuilds;
  out->removed = (uint64_t)(e->n_uploaded - c.alive);
  out->neighbour_capacity = e->nbr_cap;
  out->max_neighbours = c.max_count;
  out->overlap_steps = c.ovl_steps;
  if (e->force_grid > 0) {
    std::vector<BlockStats> bs((size_t)e->force_grid);
    HIP_TRY(hipMemcpy(bs.data(), e->d_bstats, bs.size() * sizeof(BlockStats), hipMemcpyDeviceToHost));
    for (const BlockStats& b : bs) {
      out->contacts += b.contacts;
      out->wall_contacts += b.wall_contacts;
      out->jkr_failures += b.jkr_failures;
      out->excessive_overlap |= (int32_t)b.overlap;
      out->history_contacts += b.history;
    }
  }
  return RP_OK;
}

rp_status rp_engine_set_jkr_solver(rp_engine* e, int32_t solver, const rp_jkr_envelope* env, const float* params,
                                   int32_t check) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (solver < RP_JKR_EXACT || solver > RP_JKR_SURROGATE) return fail(RP_ERR_INVALID, "unknown JKR solver");
  HIP_TRY(hipStreamSynchronize(e->stream));
  if (solver == RP_JKR_SURROGATE) {
    if (!env || !params) return fail(RP_ERR_INVALID, "the surrogate needs an envelope and parameters");
    const char* why = "";
    if (envelope_check(*env, &why) != RP_OK) return fail(RP_ERR_INVALID, why);
    for (int k = 0; k < RP_MLP_PARAMS; ++k)
      if (!std::isfinite(params[k])) return fail(RP_ERR_INVALID, "surrogate parameters must be finite");
    if (mlp_upload(e->surr_net, params, e->stream) != RP_OK) return fail(RP_ERR_DEVICE, "surrogate weight upload failed");
    e->surr_env = *env;
    if (e->surr_rows_cap == 0) {
      rp_status st = alloc_surr_rows(e, 1 << 17);
      if (st) return st;
    }
  }
  e->jkr_solver = solver;
  e->surr_check = check ? 1 : 0;
  e->surr_inst_valid = false;
  e->gen++;
  if (e->d_ctl) {  // near masks of an earlier surrogate period are stale: the next collect scans everything
    const unsigned long long zero = 0ull;
    HIP_TRY(hipMemcpy(&e->d_ctl->surr_mask_epoch, &zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  e->sstats_valid = false;
  return RP_OK;
}

rp_status rp_engine_jkr_stats(rp_engine* e, rp_jkr_stats* out) {
  if (!e || !out) return fail(RP_ERR_INVALID, "NULL argument");
  std::memset(out, 0, sizeof(*out));
  out->rows_capacity = (uint64_t)e->surr_rows_cap;
  if (!e->have_particles) return RP_OK;
  rp_status st = read_ctl(e);
  if (st) return st;
  out->rows_max = e->host_ctl.surr_rows_max;
  if (!e->sstats_valid || e->force_grid <= 0) return RP_OK;
  std::vector<SurrStats> ss((size_t)e->force_grid);
  HIP_TRY(hipMemcpy(ss.data(), e->d_sstats, ss.size() * sizeof(SurrStats), hipMemcpyDeviceToHost));
  double es = 0.0, rs = 0.0;
  for (const SurrStats& b : ss) {
    out->served += b.served;
    out->fallback += b.fallback;
    out->err_max = std::max(out->err_max, (double)b.err_max);
    out->raw_max = std::max(out->raw_max, (double)b.raw_max);
    es += b.err_sum;
    rs += b.raw_sum;
  }
  if (out->served) {
    out->err_mean = es / (double)out->served;
    out->raw_mean = rs / (double)out->served;
  }
  return RP_OK;
}

rp_status rp_engine_autotune(rp_engine* e, uint32_t steps, double* skin_out, double* cell_out) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (steps == 0) return fail(RP_ERR_INVALID, "steps must be positive");
  rp_status st = prepare(e);
  if (st) return st;
  // the tangential history (physics-fix mode) in API order, so that it can
  // be loaded into tables of any capacity after each candidate's rebuild
  std::vector<int64_t> h_off, h_part;
  std::vector<double> h_xi;
  size_t h_tot = 0;
  if (e->hist_on) {
    size_t n = 0;
    if ((st = rp_engine_count(e, &n))) return st;
    if ((st = rp_engine_history(e, nullptr, nullptr, nullptr, 0, &h_tot))) return st;
    h_off.resize(n + 1);
    h_part.resize(std::max<size_t>(h_tot, 1));
    h_xi.resize(3 * std::max<size_t>(h_tot, 1));
    if ((st = rp_engine_history(e, h_off.data(), h_part.data(), h_xi.data(), h_part.size(), &h_tot))) return st;
  }
  e->skin_tuned = false;  // no guard while the candidates run
  HIP_TRY(hipStreamSynchronize(e->stream));
  // snapshot of everything a step reads: state buffers, control block, host
  // counters (the lists are rebuilt from the state)
  const size_t db = dyn_bytes(e->cap), sb = stat_bytes(e->cap);
  void *sdyn = nullptr, *sstat = nullptr;
  HIP_TRY(hipMalloc(&sdyn, db));
  HIP_TRY(hipMalloc(&sstat, sb));
  const int cur0 = e->cur;
  HIP_TRY(hipMemcpy(sdyn, e->d_dyn_mem[cur0], db, hipMemcpyDeviceToDevice));
  HIP_TRY(hipMemcpy(sstat, e->d_stat_mem[0], sb, hipMemcpyDeviceToDevice));
  Ctl c0{};
  HIP_TRY(hipMemcpy(&c0, e->d_ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  const uint64_t steps_done0 = e->steps_done, step_count0 = e->step_count, launch0 = e->launch_no,
                 label0 = e->label_host, chain0 = e->chain_no;
  const bool at_rest0 = e->at_rest, timing0 = e->timing;
  const double skin0 = e->prm.skin, cell0 = e->prm.cell_size;
  // after restore() and the candidate's skin and cell: the lists are built
  // with them and the history goes back into the tables
  auto reload_hist = [&]() -> rp_status {
    if (!e->hist_on) return RP_OK;
    return rp_engine_set_history(e, h_off.data(), h_part.data(), h_xi.data(), h_tot);
  };
  auto restore = [&]() -> rp_status {
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->cur = cur0;
    HIP_TRY(hipMemcpy(e->d_dyn_mem[cur0], sdyn, db, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(e->d_stat_mem[0], sstat, sb, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(e->d_ctl, &c0, sizeof(Ctl), hipMemcpyHostToDevice));
    e->steps_done = steps_done0;
    e->step_count = step_count0;
    e->launch_no = launch0;
    e->label_host = label0;
    e->chain_no = chain0;
    e->at_rest = at_rest0;
    e->geometry_dirty = true;  // lists and grid are rebuilt from the restored state
    return RP_OK;
  };
  // candidates: the current setting first, then skins as a fraction of the
  // largest radius (never below the motion bound the automatic skin uses,
  // nor the fall bound below)
  // with automatic or half-cutoff cells.  Each runs in calls of `steps`
  // steps (at least 100, so that the at-rest start and end of a call weigh
  // little) until it has rebuilt its lists 3 times or run 10 calls, so a
  // small skin pays for its frequent rebuilds; another setting must beat the
  // current one by 3 % to replace it.
  // Speeds grow after tuning under gravity: a skin must also hold the fall of
  // a particle released at rest over 1,000 steps, |g| (1000 dt)^2 (twice the
  // distance, as the rebuild check allows half the skin).
  const double gmag = std::sqrt(e->prm.gravity[0] * e->prm.gravity[0] + e->prm.gravity[1] * e->prm.gravity[1] +
                                e->prm.gravity[2] * e->prm.gravity[2]);
  const double fall = gmag * (1000.0 * e->prm.dt) * (1000.0 * e->prm.dt);
  const double vmin = std::max(4.0 * (double)kRebuildInterval * e->v_max * e->prm.dt, fall);
  const double rmax = e->r_max > 0.0 ? e->r_max : 1e-3;
  const double fr[] = {0.2, 0.1, 0.05, 0.02, 0.01, 0.005};
  const double cut = e->kp.margin_factor * 2.0 * rmax;
  std::vector<std::pair<double, double>> cand{{skin0, cell0}};
  for (double f : fr) {
    const double skin = std::max(f * rmax, vmin);
    cand.emplace_back(skin, 0.0);
    cand.emplace_back(skin, 0.5 * (cut + skin));
  }
  double best_t = HUGE_VAL, best_skin = skin0, best_cell = cell0;
  e->timing = false;
  steps = std::max<uint32_t>(steps, 100u);
  for (size_t q = 0; q < cand.size() && st == RP_OK; ++q) {
    if ((st = restore())) break;
    e->prm.skin = cand[q].first;
    e->prm.cell_size = cand[q].second;
    if ((st = reload_hist())) break;
    if ((st = rp_engine_step(e, 2))) break;
    if ((st = read_ctl(e))) break;
    const unsigned long long r0 = e->host_ctl.rebuilds;
    double t = 0.0;
    uint64_t done = 0;
    for (int chunk = 0; chunk < 10; ++chunk) {
      const auto t0 = std::chrono::steady_clock::now();
      if ((st = rp_engine_step(e, steps))) break;
      if ((st = read_ctl(e))) break;  // synchronises
      t += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      done += steps;
      if (e->host_ctl.rebuilds - r0 >= 3) break;
    }
    if (st) break;
    const double per = t / (double)done * (q == 0 ? 0.97 : 1.0);
    if (per < best_t) {
      best_t = per;
      best_skin = cand[q].first;
      best_cell = cand[q].second;
    }
  }
  // the other tiling of the zero-force pass, where the force kernel has one,
  // with the chosen skin and cell
  const int tiling0 = e->tiling;
  int best_tiling = tiling0;
  const int js = e->kp.normal == RP_NORMAL_JKR ? (e->jkr_solver == RP_JKR_EXACT ? 0 : 1) : 0;
  const bool has_split = split_pass(e->kp.normal, e->kp.tangential, false, js);
  if (st == RP_OK && has_split) {
    const int other = tiling0 == RP_TILING_WAVE ? RP_TILING_LANE : RP_TILING_WAVE;
    double t_other = HUGE_VAL;
    for (int pass = 0; pass < 2 && st == RP_OK; ++pass) {
      // pass 0: the chosen setting again with the current tiling (same
      // warm-up as the other), pass 1: the other tiling
      if ((st = restore())) break;
      e->prm.skin = best_skin;
      e->prm.cell_size = best_cell;
      (void)rp_engine_set_tiling(e, pass == 0 ? tiling0 : other);
      if ((st = reload_hist())) break;
      if ((st = rp_engine_step(e, 2))) break;
      if ((st = read_ctl(e))) break;
      const unsigned long long r0 = e->host_ctl.rebuilds;
      double t = 0.0;
      uint64_t done = 0;
      for (int chunk = 0; chunk < 10; ++chunk) {
        const auto t0 = std::chrono::steady_clock::now();
        if ((st = rp_engine_step(e, steps))) break;
        if ((st = read_ctl(e))) break;
        t += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        done += steps;
        if (e->host_ctl.rebuilds - r0 >= 3) break;
      }
      const double per = t / (double)std::max<uint64_t>(done, 1);
      if (pass == 0) best_t = per;
      else t_other = per;
    }
    if (st == RP_OK && t_other < 0.97 * best_t) best_tiling = other;
  }
  (void)rp_engine_set_tiling(e, st ? tiling0 : best_tiling);
  e->timing = timing0;
  rp_status rs = restore();
  e->prm.skin = st ? skin0 : best_skin;
  e->prm.cell_size = st ? cell0 : best_cell;
  if (!rs) rs = reload_hist();
  if (!st && !rs && best_skin != skin0) {
    e->skin_tuned = true;
    e->guard_steps = 0;
    e->guard_r0 = c0.rebuilds;
  }
  (void)hipFree(sdyn);
  (void)hipFree(sstat);
  if (st) return st;
  if (rs) return rs;
  if (skin_out) *skin_out = best_skin;
  if (cell_out) *cell_out = best_cell;
  return RP_OK;
}

rp_status rp_engine_set_tiling(rp_engine* e, int32_t tiling) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  if (tiling != RP_TILING_LANE && tiling != RP_TILING_WAVE) return fail(RP_ERR_INVALID, "unknown tiling");
  if (tiling != e->tiling) {
    e->tiling = tiling;
    e->gen++;  // captured step loops hold the other kernel
  }
  return RP_OK;
}

rp_status rp_engine_get_tiling(rp_engine* e, int32_t* tiling) {
  if (!e || !tiling) return fail(RP_ERR_INVALID, "NULL argument");
  *tiling = e->tiling;
  return RP_OK;
}

rp_status rp_engine_set_graph(rp_engine* e, int on) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  e->use_graph = on != 0;
  return RP_OK;
}

rp_status rp_engine_set_timing(rp_engine* e, int on) {
  if (!e) return fail(RP_ERR_INVALID, "engine is NULL");
  HIP_TRY(hipStreamSynchronize(e->stream));
  e->timing = on != 0;
  e->ev_used = 0;
  e->ev_marks.clear();
  e->ext_start = e->ext_stop = nullptr;
  return RP_OK;
}

rp_status rp_engine_timing(rp_engine* e, double* ms, uint64_t* count) {
  if (!e || !ms || !count) return fail(RP_ERR_INVALID, "NULL argument");
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (int k = 0; k < RP_TIMING_KINDS; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  for (auto& mk : e->ev_marks) {
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e->ev_pool[mk.idx], e->ev_pool[mk.idx + 1]));
    ms[mk.kind] += (double)t;
    count[mk.kind] += mk.count;
  }
  e->ev_used = 0;
  e->ev_marks.clear();
  e->ext_start = e->ext_stop = nullptr;
  return RP_OK;
}

rp_status rp_probe_copy(size_t bytes, double* ms) {
  if (!ms) return fail(RP_ERR_INVALID, "ms is NULL");
  size_t n = bytes / sizeof(double);
  double *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, n * sizeof(double)));
  HIP_TRY(hipMalloc(&b, n * sizeof(double)));
  HIP_TRY(hipMemset(a, 0, n * sizeof(double)));
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  // the fastest of 5 launches (the first pays for first-touch effects)
  float best = 0.f;
  for (int rep = 0; rep < 5; ++rep) {
    HIP_TRY(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_copy_probe, dim3(32768), dim3(512), 0, nullptr, a, b, n);
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    if (rep == 0 || t < best) best = t;
  }
  *ms = best;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIP_TRY(hipFree(a));
  HIP_TRY(hipFree(b));
  return RP_OK;
}

void rp_engine_destroy(rp_engine* e) {
  if (!e) return;
  (void)hipStreamSynchronize(e->stream);
  for (hipEvent_t ev : e->ev_pool) (void)hipEventDestroy(ev);
  for (hipGraphExec_t g : e->graph_exec)
    if (g) (void)hipGraphExecDestroy(g);
  free_particles(e);
  free_ptr(e->d_cell_count);
  free_ptr(e->d_cell_start);
  free_ptr(e->d_block_sums);
  free_ptr(e->d_cell_rmax);
  free_ptr(e->d_tris);
  free_ptr(e->d_tri_start);
  free_ptr(e->d_tri_idx);
  free_ptr(e->d_ctl);
  free_ptr(e->d_itab);
  free_ptr(e->d_mov);
  free_ptr(e->d_surr_inst);
  free_ptr(e->d_surr_counters);
  free_ptr(e->d_surr_x);
  free_ptr(e->d_surr_y);
  mlp_free(e->surr_net);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
