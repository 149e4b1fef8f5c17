// rp_physics.hpp — FP64 contact laws, triangle routine and integrator of the
// DEM step, for the gfx950 kernels in rp_engine.hip.
//
// Each function evaluates the reference's expression in the reference's
// operation order (glam Vec3, Cargo.toml:7): dot = (x*x'+y*y')+z*z',
// normalize = v*(1/|v|), per-component scalar ops.  The translation unit is
// built with -ffp-contract=off so no multiply-add is fused, which keeps the
// pair predicate bit-exact and every contact force identical, bit for bit, to
// a plain IEEE evaluation of the same expression.
//
// Per-material factors that the reference recomputes per contact (E*, the
// Hertz damping ratio, JKR's 8*pi*gamma/E*, ...) are evaluated once on the host
// with the same operation order (MaterialConsts below), which yields the same
// double values.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rp {

struct dv3 {
  double x, y, z;
};

__host__ __device__ inline dv3 mk(double x, double y, double z) { return dv3{x, y, z}; }
__host__ __device__ inline dv3 add(dv3 a, dv3 b) { return dv3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline dv3 sub(dv3 a, dv3 b) { return dv3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline dv3 neg(dv3 a) { return dv3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline dv3 mul(dv3 a, double s) { return dv3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline dv3 smul(double s, dv3 a) { return dv3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline dv3 divs(dv3 a, double s) { return dv3{a.x / s, a.y / s, a.z / s}; }
__host__ __device__ inline double dot(dv3 a, dv3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__host__ __device__ inline double len(dv3 a) { return sqrt(dot(a, a)); }
__host__ __device__ inline dv3 cross(dv3 a, dv3 b) {
  return dv3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ inline dv3 zero3() { return dv3{0.0, 0.0, 0.0}; }

// f32::clamp semantics (NaN passes through)
__host__ __device__ inline double clampd(double v, double lo, double hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

// delta^1.5 as delta*sqrt(delta)
__host__ __device__ inline double pow15(double d) { return d * sqrt(d); }

// Rust `as i32` of a floored value: saturating, NaN -> 0
__host__ __device__ inline int32_t sat_i32(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (-2147483647 - 1);
  return (int32_t)v;
}

__host__ __device__ inline int32_t wrap_add_i32(int32_t c, int32_t o) {
  return (int32_t)(uint32_t)((uint32_t)c + (uint32_t)o);
}

// Factors of one material pair (particle-particle or particle-wall).
struct PairConsts {
  double estar;     // effective_youngs_modulus   material.rs:33-37
  double k43e;      // (4/3)*E*                    physics.rs:101,160,245,309
  double twoe;      // 2*E*                        physics.rs:121
  double gamma;     // (g1+g2)*0.5                 simulation.rs:433,503
  double jc;        // ((8*pi)*gamma)/E*           physics.rs:165
  double jsqrtc;    // sqrt(jc)                    physics.rs:166
  double j8pg_e;    // ((8*pi)*gamma)*E*           physics.rs:207
  double jpull;     // ((-2/3)*pi)*gamma           physics.rs:212
  double spull;     // (1.5*pi)*gamma              physics.rs:252
  double sterm;     // (3*(pi*pi))*(gamma*gamma)   physics.rs:256
  double estar_sq;  // E*·E*                       physics.rs:256
  double jc23;      // jc^(2/3): the JKR length t = R*^(1/3) jc^(2/3)
};

// PairConsts from E* and the pair's surface energy (same operation order as
// the reference's per-contact expressions).
__host__ __device__ inline PairConsts pair_consts(double estar, double gamma) {
  const double pi = 3.14159265358979323846;
  PairConsts c{};
  c.estar = estar;
  c.k43e = 4.0 / 3.0 * estar;
  c.twoe = 2.0 * estar;
  c.gamma = gamma;
  const double e8pg = 8.0 * pi * gamma;
  c.jc = e8pg / estar;
  c.jsqrtc = sqrt(c.jc);
  c.j8pg_e = e8pg * estar;
  c.jpull = -2.0 / 3.0 * pi * gamma;
  c.spull = 1.5 * pi * gamma;
  c.sterm = 3.0 * (pi * pi) * (gamma * gamma);
  c.estar_sq = estar * estar;
  c.jc23 = pow(c.jc, 2.0 / 3.0);
  return c;
}

struct MaterialConsts {
  PairConsts pp, pw;
  double mu;         // particle friction (walls use it too, simulation.rs:514)
  double beta_abs;   // |ln e / sqrt(ln^2 e + pi^2)|, particle restitution (physics.rs:111-112)
  double gstar_pp;   // physics-fix mode: 1 / ((2 - nu)/G + (2 - nu)/G), G = E / (2 (1 + nu))
};

// ---- normal laws -----------------------------------------------------------

// physics.rs:86-129
__device__ inline dv3 hertz(double ov, dv3 n, dv3 vrel, const PairConsts& c, double rstar, double mstar,
                            double beta_abs) {
  if (ov <= 0.0) return zero3();
  double kn = c.k43e * sqrt(rstar);
  double fe = kn * pow15(ov);
  double kt = c.twoe * sqrt(rstar * ov);
  double cn = 2.0 * sqrt(mstar * kt) * beta_abs;
  double vn = dot(vrel, n);
  double fd = -cn * vn;
  double tot = fmax(fe + fd, 0.0);
  return smul(tot, n);
}

// physics.rs:49-70
__device__ inline dv3 linear_sd(double ov, dv3 n, dv3 vrel, double kn, double gn) {
  double vn = dot(vrel, n);
  double f = kn * ov - gn * vn;
  f = fmax(f, 0.0);
  return smul(f, n);
}

// physics.rs:132-219.  Returns the force; *failed = 1 when the residual
// check rejects the Newton root (the reference's zero-force exit).
// Residual check, pull-off clip and force of a JKR contact radius a
// (physics.rs:191-219); shared by the exact solve and the surrogate path.
__device__ inline dv3 jkr_finish(double a, double ov, dv3 n, const PairConsts& c, double rstar, int* failed) {
  double ts = sqrt(c.jc * a);
  double f = a * a / rstar - ts - ov;
  if (fabs(f) > 1e-3 * rstar) {
    *failed = 1;
    return zero3();
  }
  if (a <= 0.0) return zero3();
  double a3 = a * a * a;
  double term1 = 4.0 * c.estar * a3 / (3.0 * rstar);
  double term2 = sqrt(c.j8pg_e * a3);
  double fm = term1 - term2;
  double fpull = c.jpull * rstar;
  if (fm < fpull) return zero3();
  return smul(fm, n);
}

__device__ inline dv3 jkr(double ov, dv3 n, const PairConsts& c, double rstar, int* failed,
                          double* a_out = nullptr) {
  if (c.gamma <= 0.0) {
    double kn = c.k43e * sqrt(rstar);
    double fm = ov > 0.0 ? kn * pow15(ov) : 0.0;
    return smul(fm, n);
  }
  double a = ov > 0.0 ? sqrt(rstar * ov) : rstar * 0.5;
  for (int it = 0; it < 20; ++it) {
    if (a <= 0.0) a = 1e-9;
    double ts = sqrt(c.jc * a);
    double f = a * a / rstar - ts - ov;
    double df = 2.0 * a / rstar - 0.5 * c.jsqrtc / sqrt(a);
    if (fabs(df) < 1e-9) break;
    double da = f / df;
    double ms = a * 0.5;
    double st = clampd(da, -ms, ms);
    a -= st;
    if (fabs(st) < 1e-6 * rstar) break;
  }
  if (a_out) *a_out = a;
  return jkr_finish(a, ov, n, c, rstar, failed);
}

// Minimum over a > 0 of a^2/R* - sqrt(C a) - delta, in units of
// t = (R*^2 C)^(2/3) / R*, plus the overlap: min_alpha (alpha^2 - sqrt(alpha))
// = 4^(-4/3) - 4^(-1/3) at alpha = 4^(-2/3).
constexpr double kJkrDMin = -0.47247039371057744;

// True when the Newton solve of jkr() cannot pass its residual check: the
// residual a^2/R* - sqrt(C a) - delta is >= t kJkrDMin - delta for every
// a > 0, so delta < t kJkrDMin - 1e-3 R* rejects every iterate.  The 1e-8
// term covers the rounding of both evaluations (the iterates stay within
// 1.5^20 of the start, so the residual's rounding is below 1e-9 of
// R* + t + |delta|).  Such a contact has zero force and counts as a rejected
// root, exactly as the 20 iterations would end.
__device__ inline bool jkr_no_root(double ov, double rstar, const PairConsts& c) {
  if (!(c.gamma > 0.0)) return false;
  const double s = cbrt(rstar * rstar * c.jc);
  const double t = s * s / rstar;
  return ov < kJkrDMin * t - 1e-3 * rstar - 1e-8 * (rstar + t + fabs(ov));
}

// physics.rs:234-325
__device__ inline dv3 sjkr(double ov, dv3 n, dv3 vrel, const PairConsts& c, double rstar) {
  if (c.gamma <= 0.0) {
    double kn = c.k43e * sqrt(rstar);
    double fm = ov > 0.0 ? kn * pow15(ov) : 0.0;
    return smul(fm, n);
  }
  double fpull = c.spull * rstar;
  double term = c.sterm * rstar / c.estar_sq;
  double dc = -sqrt(term);
  if (ov < dc) return zero3();
  double vn = dot(vrel, n);
  double coh = vn > 0.0 ? 1.0 : 0.0;
  double fh = ov > 0.0 ? c.k43e * sqrt(rstar) * pow15(ov) : 0.0;
  double fad = coh * fpull;
  double tot = fh - fad;
  return smul(tot, n);
}

// ---- tangential laws --------------------------------------------------------

// physics.rs:385-397
__device__ inline dv3 coulomb(double fn, dv3 vrel, dv3 n, double mu) {
  dv3 vt = sub(vrel, smul(dot(vrel, n), n));
  double l2 = dot(vt, vt);
  if (l2 < 1e-8) return zero3();
  dv3 dir = mul(vt, 1.0 / sqrt(l2));
  return smul(-mu * fn, dir);
}

// linear_spring_coulomb (physics.rs:400-434) with the trial force -kt*xi.
// With the placeholder shear modulus 0 (material.rs:49) kt is 0 or NaN, and
// -kt*xi == -kt*0 for every finite history xi, so the history is not carried.
__device__ inline dv3 spring_coulomb_g0(double fn, double kt, double mu) {
  dv3 trial = smul(-kt, zero3());
  double tm = len(trial);
  double fc = mu * fn;
  if (tm > fc) {
    dv3 dir = tm > 1e-6 ? divs(trial, tm) : zero3();
    return smul(fc, dir);
  }
  return trial;
}

// ---- physics-fix mode: tangential spring with history -------------------------
// linear_spring_coulomb (physics.rs:400-434) with the stored displacement xi
// projected onto the current tangent plane first, no force and no history
// for kt <= 0 (Mindlin at delta <= 0, where physics.rs:452 is NaN), and the
// slip direction for any non-zero trial force.
__device__ inline dv3 spring_coulomb_fix(double fn, dv3 vrel, dv3 n, dv3& xi, double kt, double mu, double dt) {
  if (!(kt > 0.0)) {
    xi = zero3();
    return zero3();
  }
  xi = sub(xi, smul(dot(xi, n), n));
  const dv3 vt = sub(vrel, smul(dot(vrel, n), n));
  xi = add(xi, mul(vt, dt));
  const dv3 trial = smul(-kt, xi);
  const double tm = len(trial);
  const double fc = mu * fn;
  if (tm > fc) {
    const dv3 dir = tm > 0.0 ? divs(trial, tm) : zero3();
    const dv3 ft = smul(fc, dir);
    xi = divs(neg(ft), kt);
    return ft;
  }
  return trial;
}

// Mindlin stiffness 8 G* sqrt(R* delta) for delta > 0, else 0 (physics.rs:443-459)
__device__ inline double mindlin_kt_fix(double gstar, double rstar, double ov) {
  return ov > 0.0 ? 8.0 * gstar * sqrt(rstar * ov) : 0.0;
}

// ---- triangle ---------------------------------------------------------------

// A triangle as the STL gives it (vertices, file normal), plus the unit
// normal pn and offset pd of its plane (pn.x = pd on the plane; pn = 0 for a
// degenerate triangle) and a rounding allowance ptol for tests against it.
// Every closest point -- the reference's clamped ones of regions 2-6 too --
// is v0 + s e0 + t e1 and lies in that plane, so a centre farther than r from
// the plane has no contact with the triangle.
struct Tri {
  dv3 v0, v1, v2, n;
  dv3 pn;
  double pd, ptol, pad;
};

// mesh.rs:45-117, with the reference's clamps in regions 2-6 (mesh.rs:67,71,75,78,95,105).
__device__ inline dv3 closest_point(const Tri& tr, dv3 p) {
  dv3 e0 = sub(tr.v1, tr.v0);
  dv3 e1 = sub(tr.v2, tr.v0);
  dv3 v0 = sub(tr.v0, p);
  double a = dot(e0, e0), b = dot(e0, e1), c = dot(e1, e1), d = dot(e0, v0), e = dot(e1, v0);
  double det = a * c - b * b;
  double s = b * e - c * d;
  double t = b * d - a * e;
  double sc = s, tc = t;
  if (s + t < det) {
    if (s < 0.0) {
      if (t < 0.0) {
        if (d < 0.0) {
          sc = clampd(-d, 0.0, a);
          tc = 0.0;
        } else {
          sc = 0.0;
          tc = clampd(-e, 0.0, c);
        }
      } else {
        sc = 0.0;
        tc = clampd(-e, 0.0, c);
      }
    } else if (t < 0.0) {
      sc = clampd(-d, 0.0, a);
      tc = 0.0;
    } else {
      double inv_det = 1.0 / det;
      sc *= inv_det;
      tc *= inv_det;
    }
  } else {
    if (s < 0.0) {
      double tmp0 = b + d;
      double tmp1 = c + e;
      if (tmp1 > tmp0) {
        double numer = tmp1 - tmp0;
        double denom = a - 2.0 * b + c;
        sc = clampd(numer / denom, 0.0, 1.0);
        tc = 1.0 - sc;
      } else {
        tc = clampd(-e, 0.0, c);
        sc = 0.0;
      }
    } else if (t < 0.0) {
      if (a + d > b + e) {
        double numer = c + e - b - d;
        double denom = a - 2.0 * b + c;
        sc = clampd(numer / denom, 0.0, 1.0);
        tc = 1.0 - sc;
      } else {
        sc = clampd(-e, 0.0, c);
        tc = 0.0;
      }
    } else {
      double numer = c + e - b - d;
      double denom = a - 2.0 * b + c;
      sc = clampd(numer / denom, 0.0, 1.0);
      tc = 1.0 - sc;
    }
  }
  return add(add(tr.v0, smul(sc, e0)), smul(tc, e1));
}

// Closest point of a triangle, physics-fix mode (Eberly, "Distance Between
// Point and Triangle in 3D", Geometric Tools): the regions of mesh.rs:63-113
// with each edge parameter normalised by its edge's squared length before
// the clamp to [0, 1].
__device__ inline double edge_param(double num, double den) {
  return num <= 0.0 ? 0.0 : (num >= den ? 1.0 : num / den);
}
__device__ inline dv3 closest_point_fix(const Tri& tr, dv3 p) {
  const dv3 e0 = sub(tr.v1, tr.v0);
  const dv3 e1 = sub(tr.v2, tr.v0);
  const dv3 v0 = sub(tr.v0, p);
  const double a = dot(e0, e0), b = dot(e0, e1), c = dot(e1, e1), d = dot(e0, v0), e = dot(e1, v0);
  const double det = a * c - b * b;
  double s = b * e - c * d;
  double t = b * d - a * e;
  if (s + t <= det) {
    if (s < 0.0) {
      if (t < 0.0 && d < 0.0) {
        s = edge_param(-d, a);
        t = 0.0;
      } else {
        s = 0.0;
        t = edge_param(-e, c);
      }
    } else if (t < 0.0) {
      s = edge_param(-d, a);
      t = 0.0;
    } else {
      s = s / det;
      t = t / det;
    }
  } else if (s < 0.0) {
    const double tmp0 = b + d, tmp1 = c + e;
    if (tmp1 > tmp0) {
      s = edge_param(tmp1 - tmp0, a - 2.0 * b + c);
      t = 1.0 - s;
    } else {
      s = 0.0;
      t = edge_param(-e, c);
    }
  } else if (t < 0.0) {
    const double tmp0 = b + e, tmp1 = a + d;
    if (tmp1 > tmp0) {
      t = edge_param(tmp1 - tmp0, a - 2.0 * b + c);
      s = 1.0 - t;
    } else {
      t = 0.0;
      s = edge_param(-d, a);
    }
  } else {
    s = edge_param((c + e) - (b + d), a - 2.0 * b + c);
    t = 1.0 - s;
  }
  return add(add(tr.v0, smul(s, e0)), smul(t, e1));
}

// ---- integrator -------------------------------------------------------------

// Kahan update (simulation.rs:238-242): y = d - res; t = val + y; res = (t - val) - y; val = t
__device__ inline void kahan1(double& val, double& res, double d) {
  double y = d - res;
  double t = val + y;
  res = (t - val) - y;
  val = t;
}

}  // namespace rp
