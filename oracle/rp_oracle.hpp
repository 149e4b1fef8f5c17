// rp_oracle.hpp — CPU restatement of the reference DEM time step.
//
// TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP engine:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product (librp_hip.so, rusty_particles_amd) never calls it.
//
// It restates, operation by operation, the reference's f32 algorithm for the
// path, templated on the scalar type R (float reproduces the reference's own
// precision, double is the FP64 parity oracle for the HIP engine):
//   Simulation::step_cpu        src/simulation.rs:173-562
//   Grid                        src/grid.rs:12-87
//   force laws                  src/physics.rs:49-459
//   effective properties        src/material.rs:25-54
//   Triangle::intersect_sphere  src/mesh.rs:20-117
//   Particle                    src/particle.rs:3-37
// glam 0.24 Vec3 semantics (Cargo.toml:7, scalar-math): dot = (x*x'+y*y')+z*z',
// normalize = v * (1/length), scalar ops per component.  No FMA contraction
// (build with -ffp-contract=off).  Rust f32::max/min are IEEE maxNum/minNum
// (std::fmax/fmin); `as i32` saturates and maps NaN to 0 (sat_i32).
//
// Parity pins (what checks each function; see tests/):
//   force laws       closed forms + known-answer vectors (tests/golden) and the
//                    reference's pydem analytical tests restated
//   JKR              exact inverse known answers (pick a, derive delta)
//   closest_point    mesh.rs:146-172 known answer + brute-force sampling of the
//                    true closest point for regions 0 and 1
//   grid             brute-force O(N^2) pair enumeration
//   integrator       free-fall closed form, Kahan algebra
// Regions 2-6 of closest_point are transcribed with the reference's clamps
// (mesh.rs:63-113); only the transcription defines them.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace rpo {

template <class R>
struct V3 {
  R x, y, z;
};

template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
// Vec3 * f32 and f32 * Vec3 (glam: per component)
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline R length_sq(V3<R> a) { return dot(a, a); }
template <class R> inline R length(V3<R> a) { return std::sqrt(dot(a, a)); }
// glam Vec3::cross
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
  return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// glam Vec3::normalize = self * (1 / length)
template <class R> inline V3<R> normalize(V3<R> a) { return a * (R(1) / length(a)); }

template <class R> inline R pi() { return R(3.14159265358979323846); }

// Rust `as i32` on an already floored value: saturating, NaN -> 0.
template <class R> inline int32_t sat_i32(R v) {
  if (std::isnan(v)) return 0;
  if (v >= R(2147483647.0)) return 2147483647;
  if (v <= R(-2147483648.0)) return (-2147483647 - 1);
  return (int32_t)v;
}

// f32::clamp: NaN passes through.
template <class R> inline R clampr(R v, R lo, R hi) {
  if (v < lo) return lo;
  if (v > hi) return hi;
  return v;
}

// overlap.powf(1.5) written as d*sqrt(d) (identical value up to rounding; the
// HIP engine evaluates the same expression).
template <class R> inline R pow15(R d) { return d * std::sqrt(d); }

template <class R>
struct Material {
  R E, nu, rho, mu, e, gamma;
};

// material.rs:25-54
template <class R> inline R effective_radius(R r1, R r2) { return (r1 * r2) / (r1 + r2); }
template <class R> inline R effective_mass(R m1, R m2) { return (m1 * m2) / (m1 + m2); }
template <class R> inline R effective_youngs(R e1, R nu1, R e2, R nu2) {
  R inv = (R(1) - nu1 * nu1) / e1 + (R(1) - nu2 * nu2) / e2;
  return R(1) / inv;
}
// material.rs:39-50: placeholder returns 0.
template <class R> inline R effective_shear(R, R) { return R(0); }

// physics.rs:49-70
template <class R>
inline V3<R> linear_spring_dashpot(R overlap, V3<R> n, V3<R> vrel, R kn, R gn) {
  R vn = dot(vrel, n);
  R f = kn * overlap - gn * vn;
  f = std::fmax(f, R(0));
  return f * n;
}

// physics.rs:86-129
template <class R>
inline V3<R> hertzian_contact(R overlap, V3<R> n, V3<R> vrel, R estar, R rstar, R mstar, R rest) {
  if (overlap <= R(0)) return {R(0), R(0), R(0)};
  R kn = R(4) / R(3) * estar * std::sqrt(rstar);
  R fe = kn * pow15(overlap);
  R lne = std::log(rest);
  R beta = lne / std::sqrt(lne * lne + pi<R>() * pi<R>());
  R ktan = R(2) * estar * std::sqrt(rstar * overlap);
  R cn = R(2) * std::sqrt(mstar * ktan) * std::fabs(beta);
  R vn = dot(vrel, n);
  R fd = -cn * vn;
  R tot = std::fmax(fe + fd, R(0));
  return tot * n;
}

struct JkrInfo {
  int iterations;
  int converged;  // 0 when the residual check rejected the root (zero force)
};

// physics.rs:132-219
template <class R>
inline V3<R> jkr_force(R overlap, V3<R> n, V3<R> /*vrel*/, R estar, R rstar, R gamma,
                       JkrInfo* info = nullptr, R* a_out = nullptr) {
  if (info) { info->iterations = 0; info->converged = 1; }
  if (a_out) *a_out = R(0);
  if (gamma <= R(0)) {
    R kn = R(4) / R(3) * estar * std::sqrt(rstar);
    R fm = overlap > R(0) ? kn * pow15(overlap) : R(0);
    return fm * n;
  }
  R c = R(8) * pi<R>() * gamma / estar;
  R sqrt_c = std::sqrt(c);
  R a = overlap > R(0) ? std::sqrt(rstar * overlap) : rstar * R(0.5);
  int it = 0;
  for (; it < 20; ++it) {
    if (a <= R(0)) a = R(1e-9);
    R ts = std::sqrt(c * a);
    R f = a * a / rstar - ts - overlap;
    R df = R(2) * a / rstar - R(0.5) * sqrt_c / std::sqrt(a);
    if (std::fabs(df) < R(1e-9)) break;
    R da = f / df;
    R max_step = a * R(0.5);
    R step = clampr(da, -max_step, max_step);
    a -= step;
    if (std::fabs(step) < R(1e-6) * rstar) { ++it; break; }
  }
  if (info) info->iterations = it;
  if (a_out) *a_out = a;
  R ts = std::sqrt(c * a);
  R f = a * a / rstar - ts - overlap;
  if (std::fabs(f) > R(1e-3) * rstar) {
    if (info) info->converged = 0;
    return {R(0), R(0), R(0)};
  }
  if (a <= R(0)) return {R(0), R(0), R(0)};
  R a3 = a * a * a;
  R term1 = R(4) * estar * a3 / (R(3) * rstar);
  R term2 = std::sqrt(R(8) * pi<R>() * gamma * estar * a3);
  R fm = term1 - term2;
  R fpull = R(-2) / R(3) * pi<R>() * gamma * rstar;
  if (fm < fpull) return {R(0), R(0), R(0)};
  return fm * n;
}

// physics.rs:234-325
template <class R>
inline V3<R> sjkr_force(R overlap, V3<R> n, V3<R> vrel, R estar, R rstar, R gamma) {
  if (gamma <= R(0)) {
    R kn = R(4) / R(3) * estar * std::sqrt(rstar);
    R fm = overlap > R(0) ? kn * pow15(overlap) : R(0);
    return fm * n;
  }
  R fpull = R(1.5) * pi<R>() * gamma * rstar;
  R term = R(3) * (pi<R>() * pi<R>()) * (gamma * gamma) * rstar / (estar * estar);
  R dc = -std::sqrt(term);
  if (overlap < dc) return {R(0), R(0), R(0)};
  R vn = dot(vrel, n);
  R coh = vn > R(0) ? R(1) : R(0);
  R fh = overlap > R(0) ? R(4) / R(3) * estar * std::sqrt(rstar) * pow15(overlap) : R(0);
  R fad = coh * fpull;
  R tot = fh - fad;
  return tot * n;
}

// physics.rs:385-397
template <class R>
inline V3<R> coulomb_friction(R fn, V3<R> vrel, V3<R> n, R mu) {
  V3<R> vt = vrel - dot(vrel, n) * n;
  if (length_sq(vt) < R(1e-8)) return {R(0), R(0), R(0)};
  V3<R> dir = normalize(vt);
  return (-mu * fn) * dir;
}

// physics.rs:400-434 (xi = Contact.tangential_displacement, updated in place)
template <class R>
inline V3<R> linear_spring_coulomb(R fn, V3<R> vrel, V3<R> n, V3<R>& xi, R kt, R mu, R dt) {
  V3<R> vt = vrel - dot(vrel, n) * n;
  xi = xi + vt * dt;
  V3<R> trial = (-kt) * xi;
  R tm = length(trial);
  R fc = mu * fn;
  if (tm > fc) {
    V3<R> dir = tm > R(1e-6) ? trial / tm : V3<R>{R(0), R(0), R(0)};
    V3<R> ft = fc * dir;
    xi = (-ft) / kt;
    return ft;
  }
  return trial;
}

// physics.rs:443-459
template <class R>
inline V3<R> mindlin_contact(R fn, R overlap, V3<R> vrel, V3<R> n, V3<R>& xi, R gstar, R rstar, R mu,
                             R dt) {
  R kt = R(8) * gstar * std::sqrt(rstar * overlap);
  return linear_spring_coulomb(fn, vrel, n, xi, kt, mu, dt);
}

template <class R>
struct Triangle {
  V3<R> v0, v1, v2, n;
};

// mesh.rs:45-117, transcribed with the reference's clamps in regions 2-6.
template <class R>
inline V3<R> closest_point(const Triangle<R>& tr, V3<R> p) {
  V3<R> e0 = tr.v1 - tr.v0;
  V3<R> e1 = tr.v2 - tr.v0;
  V3<R> v0 = tr.v0 - p;
  R a = dot(e0, e0);
  R b = dot(e0, e1);
  R c = dot(e1, e1);
  R d = dot(e0, v0);
  R e = dot(e1, v0);
  R det = a * c - b * b;
  R s = b * e - c * d;
  R t = b * d - a * e;
  R sc = s, tc = t;
  if (s + t < det) {
    if (s < R(0)) {
      if (t < R(0)) {
        if (d < R(0)) {
          sc = clampr(-d, R(0), a);
          tc = R(0);
        } else {
          sc = R(0);
          tc = clampr(-e, R(0), c);
        }
      } else {
        sc = R(0);
        tc = clampr(-e, R(0), c);
      }
    } else if (t < R(0)) {
      sc = clampr(-d, R(0), a);
      tc = R(0);
    } else {
      R inv_det = R(1) / det;
      sc *= inv_det;
      tc *= inv_det;
    }
  } else {
    if (s < R(0)) {
      R tmp0 = b + d;
      R tmp1 = c + e;
      if (tmp1 > tmp0) {
        R numer = tmp1 - tmp0;
        R denom = a - R(2) * b + c;
        sc = clampr(numer / denom, R(0), R(1));
        tc = R(1) - sc;
      } else {
        tc = clampr(-e, R(0), c);
        sc = R(0);
      }
    } else if (t < R(0)) {
      if (a + d > b + e) {
        R numer = c + e - b - d;
        R denom = a - R(2) * b + c;
        sc = clampr(numer / denom, R(0), R(1));
        tc = R(1) - sc;
      } else {
        sc = clampr(-e, R(0), c);
        tc = R(0);
      }
    } else {
      R numer = c + e - b - d;
      R denom = a - R(2) * b + c;
      sc = clampr(numer / denom, R(0), R(1));
      tc = R(1) - sc;
    }
  }
  return tr.v0 + sc * e0 + tc * e1;
}

// mesh.rs:20-43
template <class R>
inline bool intersect_sphere(const Triangle<R>& tr, V3<R> center, R radius, R* pen, V3<R>* normal) {
  V3<R> cp = closest_point(tr, center);
  V3<R> dv = center - cp;
  R dsq = length_sq(dv);
  if (dsq < radius * radius) {
    R dist = std::sqrt(dsq);
    *pen = radius - dist;
    *normal = dist > R(1e-6) ? dv / dist : tr.n;
    return true;
  }
  return false;
}

// ---- physics-fix mode (no reference counterpart: corrections of its
// placeholders, see rp_abi.h RP_MODE_PHYSICS_FIX) -------------------------

// Shear modulus G = E / (2 (1 + nu)) (material.rs:52-54) and the Mindlin
// effective shear modulus 1/G* = (2 - nu1)/G1 + (2 - nu2)/G2, in place of
// the placeholder 0 of material.rs:39-50.
template <class R> inline R shear_modulus(R e, R nu) { return e / (R(2) * (R(1) + nu)); }
template <class R> inline R effective_shear_fix(R e1, R nu1, R e2, R nu2) {
  R inv = (R(2) - nu1) / shear_modulus(e1, nu1) + (R(2) - nu2) / shear_modulus(e2, nu2);
  return R(1) / inv;
}

// Tangential spring with Coulomb cap and history (linear_spring_coulomb,
// physics.rs:400-434, with three corrections): the stored displacement is
// first projected onto the current tangent plane (the contact normal turns
// with the particles); kt <= 0 (Mindlin at zero or negative overlap, where
// physics.rs:452 takes the square root of a negative number) carries no
// force and no history; the slip direction uses any non-zero trial force
// (the reference's 1e-6 N threshold zeroes the slip force of small grains).
template <class R>
inline V3<R> spring_coulomb_fix(R fn, V3<R> vrel, V3<R> n, V3<R>& xi, R kt, R mu, R dt) {
  if (!(kt > R(0))) {
    xi = V3<R>{R(0), R(0), R(0)};
    return xi;
  }
  xi = xi - dot(xi, n) * n;
  V3<R> vt = vrel - dot(vrel, n) * n;
  xi = xi + vt * dt;
  V3<R> trial = (-kt) * xi;
  R tm = length(trial);
  R fc = mu * fn;
  if (tm > fc) {
    V3<R> dir = tm > R(0) ? trial / tm : V3<R>{R(0), R(0), R(0)};
    V3<R> ft = fc * dir;
    xi = (-ft) / kt;
    return ft;
  }
  return trial;
}

// Mindlin stiffness kt = 8 G* sqrt(R* delta) for delta > 0 and 0 otherwise
// (physics.rs:443-459 without the NaN); the spring law's kt = G* R*
// (simulation.rs:450).
template <class R> inline R mindlin_kt_fix(R gstar, R rstar, R overlap) {
  return overlap > R(0) ? R(8) * gstar * std::sqrt(rstar * overlap) : R(0);
}

// Closest point of a triangle (Eberly, "Distance Between Point and Triangle
// in 3D", Geometric Tools): the same regions as mesh.rs:63-113 with each
// edge parameter divided by its edge's squared length before the clamp to
// [0, 1], and region 2/6's fallback onto the correct edge.
template <class R>
inline V3<R> closest_point_fix(const Triangle<R>& tr, V3<R> p) {
  V3<R> e0 = tr.v1 - tr.v0;
  V3<R> e1 = tr.v2 - tr.v0;
  V3<R> v0 = tr.v0 - p;
  R a = dot(e0, e0), b = dot(e0, e1), c = dot(e1, e1), d = dot(e0, v0), e = dot(e1, v0);
  R det = a * c - b * b;
  R s = b * e - c * d;
  R t = b * d - a * e;
  auto edge = [](R num, R den) { return num <= R(0) ? R(0) : (num >= den ? R(1) : num / den); };
  if (s + t <= det) {
    if (s < R(0)) {
      if (t < R(0) && d < R(0)) {  // region 4, towards edge v0-v1
        s = edge(-d, a);
        t = R(0);
      } else {  // regions 3 and 4: edge v0-v2
        s = R(0);
        t = edge(-e, c);
      }
    } else if (t < R(0)) {  // region 5: edge v0-v1
      s = edge(-d, a);
      t = R(0);
    } else {  // region 0: inside
      s = s / det;
      t = t / det;
    }
  } else if (s < R(0)) {  // region 2
    R tmp0 = b + d, tmp1 = c + e;
    if (tmp1 > tmp0) {  // edge v1-v2
      s = edge(tmp1 - tmp0, a - R(2) * b + c);
      t = R(1) - s;
    } else {  // edge v0-v2
      s = R(0);
      t = edge(-e, c);
    }
  } else if (t < R(0)) {  // region 6
    R tmp0 = b + e, tmp1 = a + d;
    if (tmp1 > tmp0) {  // edge v1-v2
      t = edge(tmp1 - tmp0, a - R(2) * b + c);
      s = R(1) - t;
    } else {  // edge v0-v1
      t = R(0);
      s = edge(-d, a);
    }
  } else {  // region 1: edge v1-v2
    s = edge((c + e) - (b + d), a - R(2) * b + c);
    t = R(1) - s;
  }
  return tr.v0 + s * e0 + t * e1;
}

// intersect_sphere (mesh.rs:20-43) on closest_point_fix
template <class R>
inline bool intersect_sphere_fix(const Triangle<R>& tr, V3<R> center, R radius, R* pen, V3<R>* normal) {
  V3<R> cp = closest_point_fix(tr, center);
  V3<R> dv = center - cp;
  R dsq = length_sq(dv);
  if (dsq < radius * radius) {
    R dist = std::sqrt(dsq);
    *pen = radius - dist;
    *normal = dist > R(1e-6) ? dv / dist : tr.n;
    return true;
  }
  return false;
}

// Kahan update, simulation.rs:238-242: y = d - res; t = val + y; res = (t - val) - y; val = t
template <class R>
inline void kahan(V3<R>& val, V3<R>& res, V3<R> d) {
  V3<R> y = d - res;
  V3<R> t = val + y;
  res = (t - val) - y;
  val = t;
}

}  // namespace rpo
