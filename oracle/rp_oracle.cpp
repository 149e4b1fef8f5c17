// rp_oracle.cpp — CPU restatement of Simulation::step_cpu (src/simulation.rs:173-562).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the timed CPU
// baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load liboracle.so.  Exported symbols carry the rpo_ prefix so they can
// never be confused with the product's rp_ ABI.
//
// Parallelism mirrors the reference's rayon loops (per-particle par_iter maps,
// simulation.rs:184,235,268,317,533,543) with OpenMP; the grid insert stays
// serial (simulation.rs:273-277).  There are no cross-particle float
// reductions, so results do not depend on the thread count.
//
// Contact history (simulation.rs:30,397-399, contact.rs:3-18): the reference
// shares one Contact per unordered pair in a DashMap.  In this semantics the
// shear modulus is the placeholder 0 (material.rs:49), so the tangential
// spring stiffness kt is 0 (or NaN when sqrt(R*delta) is NaN) and the trial
// force -kt*xi equals -kt*0 for every finite xi: the history value never
// reaches a force.  The oracle therefore evaluates the tangential laws with
// xi = 0 and keeps no map; see tangential() below.
#include "rp_oracle.hpp"

#include <omp.h>

#include <array>
#include <cstdio>
#include <memory>
#include <unordered_map>

#include "../include/rp_abi.h"

namespace rpo {

// One directed contact's tangential history in the physics-fix mode: the
// partner's stable id (its insertion index) and the stored displacement.
template <class R>
struct HistEntry {
  uint32_t key;
  V3<R> xi;
};

template <class R>
struct Particle {
  V3<R> x, v, a, rx, rv, w, tau, rw;
  R r, m;
  uint8_t fixed;
  uint32_t sid;                     // stable id = insertion index
  std::vector<HistEntry<R>> hist;   // physics-fix mode: sorted by key
};

// grid.rs:4-87
template <class R>
struct Grid {
  R cs{1};
  V3<R> origin{};
  int32_t dims[3]{1, 1, 1};
  std::vector<std::vector<int64_t>> cells;
  std::vector<size_t> used;  // cells that received a particle since the last clear

  void init(R cell_size, V3<R> mn, V3<R> mx) {
    cs = cell_size;
    origin = mn;
    V3<R> q = (mx - mn) / cs;
    dims[0] = sat_i32(std::ceil(q.x)) + 1;
    dims[1] = sat_i32(std::ceil(q.y)) + 1;
    dims[2] = sat_i32(std::ceil(q.z)) + 1;
    size_t total = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
    cells.assign(total, {});
    used.clear();
  }
  // grid.rs:24-28 empties every cell; emptying the used ones is equivalent.
  void clear() {
    for (size_t c : used) cells[c].clear();
    used.clear();
  }
  void coords(V3<R> p, int32_t c[3]) const {
    V3<R> q = (p - origin) / cs;
    c[0] = sat_i32(std::floor(q.x));
    c[1] = sat_i32(std::floor(q.y));
    c[2] = sat_i32(std::floor(q.z));
  }
  // grid.rs:30-40
  bool cell_index(V3<R> p, size_t* idx) const {
    int32_t c[3];
    coords(p, c);
    if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= dims[0] || c[1] >= dims[1] || c[2] >= dims[2])
      return false;
    *idx = (size_t)(c[0] + c[1] * dims[0] + c[2] * dims[0] * dims[1]);
    return true;
  }
  void insert(V3<R> p, int64_t id) {
    size_t idx;
    if (cell_index(p, &idx)) {
      if (cells[idx].empty()) used.push_back(idx);
      cells[idx].push_back(id);
    }
  }
  static int32_t wrap_add(int32_t c, int32_t o) {  // i32 wrapping add (release-mode Rust)
    return (int32_t)(uint32_t)((uint32_t)c + (uint32_t)o);
  }
  // grid.rs:48-87: z, y, x offsets; periodic wrap once; range check; append cell
  void neighbours(V3<R> p, const int periodic[3], std::vector<int64_t>& out) const {
    out.clear();
    int32_t c[3];
    coords(p, c);
    for (int oz = -1; oz <= 1; ++oz)
      for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
          int32_t n[3] = {wrap_add(c[0], ox), wrap_add(c[1], oy), wrap_add(c[2], oz)};
          for (int k = 0; k < 3; ++k)
            if (periodic[k]) {
              if (n[k] < 0)
                n[k] += dims[k];
              else if (n[k] >= dims[k])
                n[k] -= dims[k];
            }
          if (n[0] >= 0 && n[0] < dims[0] && n[1] >= 0 && n[1] < dims[1] && n[2] >= 0 && n[2] < dims[2]) {
            size_t idx = (size_t)(n[0] + n[1] * dims[0] + n[2] * dims[0] * dims[1]);
            const auto& cell = cells[idx];
            out.insert(out.end(), cell.begin(), cell.end());
          }
        }
  }
};

class SimBase {
 public:
  virtual ~SimBase() = default;
  virtual void set_params(const rp_params& p) = 0;
  virtual void set_state(size_t n, const double* x, const double* v, const double* w, const double* a,
                         const double* tau, const double* rx, const double* rv, const double* rw,
                         const double* r, const double* m, const uint8_t* fixed) = 0;
  virtual void get_state(double* x, double* v, double* w, double* a, double* tau, double* rx, double* rv,
                         double* rw) const = 0;
  virtual size_t count() const = 0;
  virtual void add_mesh(size_t ntri, const float* v9, const float* n3) = 0;
  virtual int set_mesh_motion(int mesh, const double* origin, const double* omega) = 0;
  virtual void step(uint64_t nsteps) = 0;
  virtual void set_cell_size(double cs) = 0;
  virtual void set_record(int on) = 0;
  virtual size_t pair_total() const = 0;
  virtual void get_pairs(int64_t* offsets, int64_t* js) const = 0;
  virtual void get_forces(double* f3, double* t3, double* fscale) const = 0;
  virtual int overlap_flag() const = 0;
  virtual uint64_t jkr_failures() const = 0;
  virtual void set_fast(bool on) = 0;
  virtual void cell_ids(int64_t* out) const = 0;
  virtual size_t history_total() const = 0;
  virtual void history(int64_t* offsets, int64_t* partner, double* xi3) const = 0;
  virtual double time_gather(size_t nsample, double cell) const = 0;
};

template <class R>
class Sim : public SimBase {
 public:
  explicit Sim(const rp_params& p) {
    set_params(p);
    grid_.init(R(p.ref_cell_size), bmin_, bmax_);
  }

  void set_params(const rp_params& p) override {
    dt_ = R(p.dt);
    bmin_ = v3<R>(R(p.bounds_min[0]), R(p.bounds_min[1]), R(p.bounds_min[2]));
    bmax_ = v3<R>(R(p.bounds_max[0]), R(p.bounds_max[1]), R(p.bounds_max[2]));
    for (int k = 0; k < 3; ++k) periodic_[k] = p.periodic[k] ? 1 : 0;
    normal_ = p.normal_model;
    tangential_ = p.tangential_model;
    pm_ = mat(p.particle);
    wm_ = mat(p.wall);
    fix_ = p.mode == RP_MODE_PHYSICS_FIX;
    gstar_pp_ = effective_shear_fix(pm_.E, pm_.nu, pm_.E, pm_.nu);
    g_ = v3<R>(R(p.gravity[0]), R(p.gravity[1]), R(p.gravity[2]));
  }

  void set_cell_size(double cs) override {
    grid_.init(R(cs), grid_.origin, grid_.origin + (bmax_ - bmin_));
  }

  void set_state(size_t n, const double* x, const double* v, const double* w, const double* a,
                 const double* tau, const double* rx, const double* rv, const double* rw, const double* r,
                 const double* m, const uint8_t* fixed) override {
    ps_.assign(n, Particle<R>{});
    for (size_t i = 0; i < n; ++i) {
      Particle<R>& q = ps_[i];
      q.x = ld(x, i);
      q.v = ld(v, i);
      q.w = ld(w, i);
      q.a = ld(a, i);
      q.tau = ld(tau, i);
      q.rx = ld(rx, i);
      q.rv = ld(rv, i);
      q.rw = ld(rw, i);
      q.r = R(r[i]);
      q.m = R(m[i]);
      q.fixed = fixed ? (fixed[i] ? 1 : 0) : 0;
      q.sid = (uint32_t)i;
      q.hist.clear();
    }
  }

  void get_state(double* x, double* v, double* w, double* a, double* tau, double* rx, double* rv,
                 double* rw) const override {
    for (size_t i = 0; i < ps_.size(); ++i) {
      const Particle<R>& q = ps_[i];
      st(x, i, q.x);
      st(v, i, q.v);
      st(w, i, q.w);
      st(a, i, q.a);
      st(tau, i, q.tau);
      st(rx, i, q.rx);
      st(rv, i, q.rv);
      st(rw, i, q.rw);
    }
  }

  size_t count() const override { return ps_.size(); }

  void add_mesh(size_t ntri, const float* v9, const float* n3) override {
    for (size_t t = 0; t < ntri; ++t) {
      Triangle<R> tr;
      tr.v0 = v3<R>(R(v9[9 * t + 0]), R(v9[9 * t + 1]), R(v9[9 * t + 2]));
      tr.v1 = v3<R>(R(v9[9 * t + 3]), R(v9[9 * t + 4]), R(v9[9 * t + 5]));
      tr.v2 = v3<R>(R(v9[9 * t + 6]), R(v9[9 * t + 7]), R(v9[9 * t + 8]));
      tr.n = v3<R>(R(n3[3 * t + 0]), R(n3[3 * t + 1]), R(n3[3 * t + 2]));
      tris_.push_back(tr);
      tri_mesh_.push_back((int)motion_.size());
    }
    motion_.push_back({0.0, 0.0, 0.0, 0.0, 0.0, 0.0});
  }

  // Extension, no reference counterpart (its walls are at rest,
  // simulation.rs:475-479): mesh `mesh` has the surface velocity of the
  // rigid rotation omega about origin; the triangles stay where they are.
  int set_mesh_motion(int mesh, const double* origin, const double* omega) override {
    if (mesh < 0 || (size_t)mesh >= motion_.size()) return -1;
    for (int k = 0; k < 3; ++k) {
      motion_[mesh][k] = origin[k];
      motion_[mesh][3 + k] = omega[k];
    }
    return 0;
  }

  void set_record(int on) override { record_ = on != 0; }

  size_t pair_total() const override {
    size_t t = 0;
    for (auto& v : pairs_) t += v.size();
    return t;
  }
  void get_pairs(int64_t* offsets, int64_t* js) const override {
    int64_t o = 0;
    for (size_t i = 0; i < pairs_.size(); ++i) {
      offsets[i] = o;
      for (int64_t j : pairs_[i]) js[o++] = j;
    }
    offsets[pairs_.size()] = o;
  }
  void get_forces(double* f3, double* t3, double* fscale) const override {
    for (size_t i = 0; i < ftot_.size(); ++i) {
      st(f3, i, ftot_[i]);
      st(t3, i, ttot_[i]);
      if (fscale) fscale[i] = (double)fscale_[i];
    }
  }
  int overlap_flag() const override { return overlap_flag_; }
  uint64_t jkr_failures() const override { return jkr_fail_; }

  void step(uint64_t nsteps) override {
    for (uint64_t s = 0; s < nsteps; ++s) step_once();
  }

 private:
  static Material<R> mat(const rp_material& m) {
    return Material<R>{R(m.youngs_modulus), R(m.poissons_ratio), R(m.density),
                       R(m.friction_coefficient), R(m.restitution_coefficient), R(m.surface_energy)};
  }
  static V3<R> ld(const double* p, size_t i) {
    if (!p) return v3<R>(R(0), R(0), R(0));
    return v3<R>(R(p[3 * i]), R(p[3 * i + 1]), R(p[3 * i + 2]));
  }
  static void st(double* p, size_t i, V3<R> v) {
    if (!p) return;
    p[3 * i] = (double)v.x;
    p[3 * i + 1] = (double)v.y;
    p[3 * i + 2] = (double)v.z;
  }

  // simulation.rs:402-440 (particle-particle) / 488-510 (wall)
  V3<R> normal_force(R overlap, V3<R> n, V3<R> vrel, R estar, R rstar, R mstar, R gamma, bool wall,
                     uint64_t* fails) const {
    switch (normal_) {
      case RP_NORMAL_HERTZ:
        return hertzian_contact(overlap, n, vrel, estar, rstar, mstar, pm_.e);
      case RP_NORMAL_LINEAR: {
        R kn = estar * rstar;
        R gn = R(0.5) * std::sqrt(mstar * kn);
        return linear_spring_dashpot(overlap, n, vrel, kn, gn);
      }
      case RP_NORMAL_HYSTERETIC: {
        R kn = estar * rstar;
        if (wall) return linear_spring_dashpot(overlap, n, vrel, kn, R(0));
        // hysteretic_contact (physics.rs:334-380) falls back to linear with
        // zero velocity and zero damping.
        return linear_spring_dashpot(overlap, n, v3<R>(R(0), R(0), R(0)), kn, R(0));
      }
      case RP_NORMAL_JKR: {
        JkrInfo info;
        V3<R> f = jkr_force(overlap, n, vrel, estar, rstar, gamma, &info);
        if (!info.converged) ++*fails;
        return f;
      }
      case RP_NORMAL_SJKR:
      default:
        return sjkr_force(overlap, n, vrel, estar, rstar, gamma);
    }
  }

  // simulation.rs:445-452; xi = 0, see the file header.
  V3<R> tangential(R fnmag, R overlap, V3<R> vrel, V3<R> n, R rstar) const {
    R gstar = effective_shear(R(0), R(0));
    V3<R> xi = v3<R>(R(0), R(0), R(0));
    switch (tangential_) {
      case RP_TANGENTIAL_MINDLIN: {
        R kt = R(8) * gstar * std::sqrt(rstar * overlap);
        V3<R> trial = (-kt) * xi;
        R tm = length(trial);
        R fc = pm_.mu * fnmag;
        if (tm > fc) {
          // Slip branch of linear_spring_coulomb; reachable only when mu*|Fn| < 0.
          V3<R> dir = tm > R(1e-6) ? trial / tm : v3<R>(R(0), R(0), R(0));
          return fc * dir;
        }
        return trial;
      }
      case RP_TANGENTIAL_LSC: {
        R kt = gstar * rstar;
        V3<R> trial = (-kt) * xi;
        R tm = length(trial);
        R fc = pm_.mu * fnmag;
        if (tm > fc) {
          V3<R> dir = tm > R(1e-6) ? trial / tm : v3<R>(R(0), R(0), R(0));
          return fc * dir;
        }
        return trial;
      }
      case RP_TANGENTIAL_COULOMB:
      default:
        return coulomb_friction(fnmag, vrel, n, pm_.mu);
    }
  }

  void step_once() {
    const R dt = dt_;
    const V3<R> bmin = bmin_, bmax = bmax_;
    const V3<R> bsize = bmax - bmin;
    const int* per = periodic_;
    const int64_t n0 = (int64_t)ps_.size();

    // 0. boundaries: wrap (simulation.rs:182-197) and retain (:200-214)
    if (per[0] || per[1] || per[2]) {
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n0; ++i) {
        Particle<R>& p = ps_[i];
        if (per[0]) {
          if (p.x.x < bmin.x) p.x.x += bsize.x;
          else if (p.x.x >= bmax.x) p.x.x -= bsize.x;
        }
        if (per[1]) {
          if (p.x.y < bmin.y) p.x.y += bsize.y;
          else if (p.x.y >= bmax.y) p.x.y -= bsize.y;
        }
        if (per[2]) {
          if (p.x.z < bmin.z) p.x.z += bsize.z;
          else if (p.x.z >= bmax.z) p.x.z -= bsize.z;
        }
      }
      size_t k = 0;
      for (size_t i = 0; i < ps_.size(); ++i) {
        const V3<R>& x = ps_[i].x;
        bool keep = (per[0] || (x.x >= bmin.x - R(1) && x.x <= bmax.x + R(1))) &&
                    (per[1] || (x.y >= bmin.y - R(1))) &&
                    (per[2] || (x.z >= bmin.z - R(1) && x.z <= bmax.z + R(1)));
        if (keep) ps_[k++] = ps_[i];
      }
      ps_.resize(k);
    } else {
      size_t k = 0;
      for (size_t i = 0; i < ps_.size(); ++i) {
        const V3<R>& x = ps_[i].x;
        bool keep = x.y >= bmin.y - R(1) && x.x >= bmin.x - R(1) && x.x <= bmax.x + R(1) &&
                    x.z >= bmin.z - R(1) && x.z <= bmax.z + R(1);
        if (keep) ps_[k++] = ps_[i];
      }
      ps_.resize(k);
    }
    // Re-index (simulation.rs:216-219): the id is the vector position.
    const int64_t n = (int64_t)ps_.size();

    // 1. first half kick, drift, angular half kick (simulation.rs:235-263)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      Particle<R>& p = ps_[i];
      if (p.fixed) continue;
      kahan(p.v, p.rv, (R(0.5) * p.a) * dt);
      kahan(p.x, p.rx, p.v * dt);
      R inertia = R(0.4) * p.m * p.r * p.r;
      V3<R> aa = p.tau / inertia;
      kahan(p.w, p.rw, (R(0.5) * aa) * dt);
    }

    // 2. reset (simulation.rs:267-271)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      ps_[i].a = g_;
      ps_[i].tau = v3<R>(R(0), R(0), R(0));
    }

    // broad phase, serial insert in index order (simulation.rs:273-277)
    grid_.clear();
    for (int64_t i = 0; i < n; ++i) grid_.insert(ps_[i].x, i);

    const bool adhesive = normal_ == RP_NORMAL_JKR || normal_ == RP_NORMAL_SJKR;
    const Material<R> pm = pm_, wm = wm_;
    const R estar_pp = effective_youngs(pm.E, pm.nu, pm.E, pm.nu);
    const R estar_pw = effective_youngs(pm.E, pm.nu, wm.E, wm.nu);
    const R gamma_pp = (pm.gamma + pm.gamma) * R(0.5);
    const R gamma_pw = (pm.gamma + wm.gamma) * R(0.5);

    ftot_.assign(n, v3<R>(R(0), R(0), R(0)));
    ttot_.assign(n, v3<R>(R(0), R(0), R(0)));
    fscale_.assign(n, R(0));
    if (record_) pairs_.assign(n, {});
    else pairs_.clear();
    int flag = 0;
    uint64_t fails = 0;

    if (fast_ || fix_) build_fast(n);

    // 3. per-particle gather (simulation.rs:317-526)
#pragma omp parallel reduction(| : flag) reduction(+ : fails)
    {
      std::vector<int64_t> nb;
      std::vector<std::pair<int, int64_t>> walk;  // fast path: (27-cell walk rank, j)
      nb.reserve(256);
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < n; ++i) {
        const Particle<R>& p = ps_[i];
        V3<R> f = v3<R>(R(0), R(0), R(0));
        V3<R> t = v3<R>(R(0), R(0), R(0));
        R fs = R(0);
        std::vector<HistEntry<R>> nhist;  // physics-fix history written this step
        // one visit of j by the grid walk (simulation.rs:322-460)
        auto visit = [&](int64_t j) {
          if (i == j) return;
          const Particle<R>& o = ps_[j];
          V3<R> d = p.x - o.x;
          if (per[0]) {
            if (d.x > bsize.x * R(0.5)) d.x -= bsize.x;
            else if (d.x < -bsize.x * R(0.5)) d.x += bsize.x;
          }
          if (per[1]) {
            if (d.y > bsize.y * R(0.5)) d.y -= bsize.y;
            else if (d.y < -bsize.y * R(0.5)) d.y += bsize.y;
          }
          if (per[2]) {
            if (d.z > bsize.z * R(0.5)) d.z -= bsize.z;
            else if (d.z < -bsize.z * R(0.5)) d.z += bsize.z;
          }
          R dsq = length_sq(d);
          R rsum = p.r + o.r;
          R margin = adhesive ? rsum * R(0.5) : R(0);
          R cutoff = rsum + margin;
          if (!(dsq < cutoff * cutoff)) return;
          R dist = std::sqrt(dsq);
          R overlap = rsum - dist;
          if (!flag) {
            R minr = std::fmin(p.r, o.r);
            if (overlap > R(0.2) * minr) flag = 1;
          }
          V3<R> nrm = dist > R(1e-6) ? d / dist : v3<R>(R(0), R(1), R(0));
          V3<R> ra = (-nrm) * p.r;
          V3<R> rb = nrm * o.r;
          V3<R> va = p.v + cross(p.w, ra);
          V3<R> vb = o.v + cross(o.w, rb);
          V3<R> vrel = va - vb;
          R rstar = effective_radius(p.r, o.r);
          R mstar = effective_mass(p.m, o.m);
          V3<R> fn = normal_force(overlap, nrm, vrel, estar_pp, rstar, mstar, gamma_pp, false, &fails);
          R fnmag = length(fn);
          V3<R> ft;
          if (fix_ && tangential_ != RP_TANGENTIAL_COULOMB) {
            // physics-fix history of the directed contact (i, j), keyed by j's stable id
            V3<R> xi = v3<R>(R(0), R(0), R(0));
            auto it = std::lower_bound(p.hist.begin(), p.hist.end(), o.sid,
                                       [](const HistEntry<R>& h, uint32_t k) { return h.key < k; });
            if (it != p.hist.end() && it->key == o.sid) xi = it->xi;
            const R kt = tangential_ == RP_TANGENTIAL_MINDLIN ? mindlin_kt_fix(gstar_pp_, rstar, overlap)
                                                              : gstar_pp_ * rstar;
            ft = spring_coulomb_fix(fnmag, vrel, nrm, xi, kt, pm.mu, dt);
            if (kt > R(0) && (xi.x != R(0) || xi.y != R(0) || xi.z != R(0))) nhist.push_back({o.sid, xi});
          } else {
            ft = tangential(fnmag, overlap, vrel, nrm, rstar);
          }
          V3<R> fc = fn + ft;
          f = f + fc;
          t = t + cross(ra, fc);
          fs += length(fc);
          if (record_) pairs_[i].push_back(j);
        };
        // one triangle (simulation.rs:465-523)
        auto visit_tri = [&](size_t ti) {
          const Triangle<R>& tr = tris_[ti];
          R pen;
          V3<R> nrm;
          if (!(fix_ ? intersect_sphere_fix(tr, p.x, p.r, &pen, &nrm) : intersect_sphere(tr, p.x, p.r, &pen, &nrm)))
            return;
          if (!flag && pen > R(0.2) * p.r) flag = 1;
          V3<R> ra = (-nrm) * p.r;
          V3<R> vrel = p.v + cross(p.w, ra);
          const std::array<double, 6>& mo = motion_[tri_mesh_[ti]];
          if (mo[3] != 0.0 || mo[4] != 0.0 || mo[5] != 0.0) {
            // a moving wall's surface velocity at the contact (closest) point
            const V3<R> cp = fix_ ? closest_point_fix(tr, p.x) : closest_point(tr, p.x);
            const V3<R> org = v3<R>(R(mo[0]), R(mo[1]), R(mo[2]));
            vrel = vrel - cross(v3<R>(R(mo[3]), R(mo[4]), R(mo[5])), cp - org);
          }
          V3<R> fn = normal_force(pen, nrm, vrel, estar_pw, p.r, p.m, gamma_pw, true, &fails);
          R fnmag = length(fn);
          V3<R> ft = coulomb_friction(fnmag, vrel, nrm, pm.mu);
          V3<R> fc = fn + ft;
          f = f + fc;
          t = t + cross(ra, fc);
          fs += length(fc);
        };
        if (fix_) {
          // physics-fix mode: every pair inside the cutoff once, ascending index
          fix_walk(i, nb);
          for (int64_t j : nb) visit(j);
          for (int32_t ti : fast_tris(i)) visit_tri((size_t)ti);
          ps_[i].hist.swap(nhist);
        } else if (!fast_) {
          grid_.neighbours(p.x, per, nb);
          for (int64_t j : nb) visit(j);
          for (size_t ti = 0; ti < tris_.size(); ++ti) visit_tri(ti);
        } else {
          fast_walk(i, walk);
          for (const auto& e : walk) visit(e.second);
          for (int32_t ti : fast_tris(i)) visit_tri((size_t)ti);
        }
        ftot_[i] = f;
        ttot_[i] = t;
        fscale_[i] = fs;
      }
    }
    overlap_flag_ = flag;
    jkr_fail_ += fails;

    // 4. apply (simulation.rs:533-536) and second half kick (:543-561)
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      Particle<R>& p = ps_[i];
      p.a = p.a + ftot_[i] / p.m;
      p.tau = p.tau + ttot_[i];
      if (p.fixed) continue;
      kahan(p.v, p.rv, (R(0.5) * p.a) * dt);
      R inertia = R(0.4) * p.m * p.r * p.r;
      V3<R> aa = p.tau / inertia;
      kahan(p.w, p.rw, (R(0.5) * aa) * dt);
    }
  }

  // ---- fast candidate enumeration --------------------------------------------
  // An optimisation of the oracle only, off by default: the same visits in
  // the same order as the literal walk, pinned to it bit for bit by
  // tests/test_oracle_fast.py.  The literal walk appends every particle of
  // the 27 reference cells (grid.rs:48-87) and tests each; a particle j can
  // pass the pair test only within the largest cutoff of i, so the fast path
  // takes j from a fine cell list (cells >= that cutoff, minimum image on
  // periodic axes), finds the offsets of the reference walk that reach j's
  // reference cell (grid.rs:58-83: wrap once per periodic axis, range check)
  // and visits j once per such offset, in the walk's order (offset rank, then
  // index, as cells hold indices in insertion order).  Triangles: a triangle
  // can touch particle i only if i lies within r_max of the set where
  // closest_point can return its point (the triangle, v0 + [0, max(1, a, c)] e0
  // and v0 + [0, max(1, c)] e1: the clamps of mesh.rs:67-105), so each fine
  // cell lists the triangles whose grown box of that set meets it, in file
  // order.
  bool fast_{false};
  bool fix_{false};   // RP_MODE_PHYSICS_FIX
  R gstar_pp_{};      // physics-fix effective shear modulus, particle-particle
  double fcs_[3]{}, forg_[3]{};
  int64_t fn_[3]{1, 1, 1};
  bool fall_[3]{};
  std::vector<int64_t> fkey_, fidx_;  // fine-cell key and particle index, sorted by (key, index)
  std::vector<int64_t> fpkey_;        // fine-cell key of particle i
  std::unordered_map<int64_t, std::vector<int32_t>> ftri_;
  double bsz_[3]{};  // box size (minimum image of the pre-test)
  bool adh_{false};   // adhesive cutoff margin (JKR, sJKR)

 public:
  void set_fast(bool on) override { fast_ = on; }
  // Seconds the reference's per-particle gather (simulation.rs:317-461: the
  // grid walk of grid.rs:48-87 over cells of size `cell`, the pair test and
  // the normal law of each contact) takes for nsample particles spread evenly
  // over the current state, OpenMP over the sample; the state is unchanged.
  // With the reference's 1 m cell each walk appends every particle of up to
  // 27 cells, so a step costs O(N^2) and a sample is all that can be timed.
  double time_gather(size_t nsample, double cell) const override {
    const int64_t n = (int64_t)ps_.size();
    if (n == 0 || nsample == 0) return 0.0;
    Grid<R> g;
    g.init(R(cell), bmin_, bmax_);
    for (int64_t i = 0; i < n; ++i) g.insert(ps_[i].x, i);
    const V3<R> bsize = bmax_ - bmin_;
    const bool adhesive = normal_ == RP_NORMAL_JKR || normal_ == RP_NORMAL_SJKR;
    const R estar = effective_youngs(pm_.E, pm_.nu, pm_.E, pm_.nu);
    const R gamma = (pm_.gamma + pm_.gamma) * R(0.5);
    const int64_t stride = std::max<int64_t>(1, n / (int64_t)nsample);
    const int64_t ns = std::min<int64_t>((int64_t)nsample, n);
    double sink = 0.0;
    const double t0 = omp_get_wtime();
#pragma omp parallel reduction(+ : sink)
    {
      std::vector<int64_t> nb;
      uint64_t fails = 0;
#pragma omp for schedule(dynamic, 1)
      for (int64_t q = 0; q < ns; ++q) {
        const int64_t i = q * stride;
        const Particle<R>& p = ps_[i];
        g.neighbours(p.x, periodic_, nb);
        V3<R> f = v3<R>(R(0), R(0), R(0));
        for (int64_t j : nb) {
          if (i == j) continue;
          const Particle<R>& o = ps_[j];
          V3<R> d = p.x - o.x;
          if (periodic_[0]) {
            if (d.x > bsize.x * R(0.5)) d.x -= bsize.x;
            else if (d.x < -bsize.x * R(0.5)) d.x += bsize.x;
          }
          if (periodic_[1]) {
            if (d.y > bsize.y * R(0.5)) d.y -= bsize.y;
            else if (d.y < -bsize.y * R(0.5)) d.y += bsize.y;
          }
          if (periodic_[2]) {
            if (d.z > bsize.z * R(0.5)) d.z -= bsize.z;
            else if (d.z < -bsize.z * R(0.5)) d.z += bsize.z;
          }
          const R rsum = p.r + o.r;
          const R cutoff = rsum + (adhesive ? rsum * R(0.5) : R(0));
          const R dsq = length_sq(d);
          if (!(dsq < cutoff * cutoff)) continue;
          const R dist = std::sqrt(dsq);
          const V3<R> nrm = dist > R(1e-6) ? d / dist : v3<R>(R(0), R(1), R(0));
          f = f + normal_force(rsum - dist, nrm, p.v - o.v, estar, effective_radius(p.r, o.r),
                               effective_mass(p.m, o.m), gamma, false, &fails);
        }
        sink += (double)f.x + (double)nb.size();
      }
    }
    const double t = omp_get_wtime() - t0;
    return sink == -1.0 ? -t : t;  // keeps the loop's result observable
  }

  size_t history_total() const override {
    size_t t = 0;
    for (const auto& q : ps_) t += q.hist.size();
    return t;
  }
  // Physics-fix history in API order: per particle its entries (partner's
  // API index, displacement), by ascending partner.
  void history(int64_t* offsets, int64_t* partner, double* xi3) const override {
    std::vector<std::pair<uint32_t, int64_t>> idx(ps_.size());
    for (size_t i = 0; i < ps_.size(); ++i) idx[i] = {ps_[i].sid, (int64_t)i};
    std::sort(idx.begin(), idx.end());
    int64_t o = 0;
    for (size_t i = 0; i < ps_.size(); ++i) {
      offsets[i] = o;
      for (const auto& h : ps_[i].hist) {
        auto it = std::lower_bound(idx.begin(), idx.end(), std::make_pair(h.key, (int64_t)-1));
        partner[o] = (it != idx.end() && it->first == h.key) ? it->second : -1;
        st(xi3, (size_t)o, h.xi);
        ++o;
      }
    }
    offsets[ps_.size()] = o;
  }
  // grid.rs:30-40 for every particle of the current state (-1: outside the grid)
  void cell_ids(int64_t* out) const override {
    for (size_t i = 0; i < ps_.size(); ++i) {
      size_t idx;
      out[i] = grid_.cell_index(ps_[i].x, &idx) ? (int64_t)idx : -1;
    }
  }

 private:
  int64_t fcoord(double x, int k) const {
    if (!std::isfinite(x)) return 0;
    double q = std::floor((x - forg_[k]) / fcs_[k]);
    if (q < -1e15) q = -1e15;
    if (q > 1e15) q = 1e15;
    int64_t c = (int64_t)q;
    if (periodic_[k]) {
      c %= fn_[k];
      if (c < 0) c += fn_[k];
      return c;
    }
    return std::min<int64_t>(std::max<int64_t>(c, 0), fn_[k] - 1);
  }
  int64_t fkey(int64_t a, int64_t b, int64_t c) const { return (a * fn_[1] + b) * fn_[2] + c; }

  void build_fast(int64_t n) {
    double rmax = 0.0;
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int64_t i = 0; i < n; ++i) {
      rmax = std::max(rmax, (double)ps_[i].r);
      const double x[3] = {(double)ps_[i].x.x, (double)ps_[i].x.y, (double)ps_[i].x.z};
      for (int k = 0; k < 3; ++k)
        if (std::isfinite(x[k])) {
          lo[k] = std::min(lo[k], x[k]);
          hi[k] = std::max(hi[k], x[k]);
        }
    }
    const bool adhesive = normal_ == RP_NORMAL_JKR || normal_ == RP_NORMAL_SJKR;
    const double cs = (adhesive ? 1.5 : 1.0) * 2.0 * rmax * 1.001 + 1e-12;
    const double bsz[3] = {(double)(bmax_.x - bmin_.x), (double)(bmax_.y - bmin_.y), (double)(bmax_.z - bmin_.z)};
    for (int k = 0; k < 3; ++k) bsz_[k] = bsz[k];
    adh_ = adhesive;
    const double bmn[3] = {(double)bmin_.x, (double)bmin_.y, (double)bmin_.z};
    for (int k = 0; k < 3; ++k) {
      if (periodic_[k]) {
        int64_t m = (int64_t)std::floor(bsz[k] / cs);
        m = std::max<int64_t>(1, std::min<int64_t>(m, 1 << 20));
        fn_[k] = m;
        fcs_[k] = bsz[k] / (double)m;
        forg_[k] = bmn[k];
        fall_[k] = m < 3;
      } else {
        if (!(lo[k] <= hi[k])) lo[k] = hi[k] = 0.0;
        double c = cs;
        if ((hi[k] - lo[k]) / c > (double)(1 << 20)) c = (hi[k] - lo[k]) / (double)(1 << 20);
        fcs_[k] = c;
        forg_[k] = lo[k] - c;
        fn_[k] = (int64_t)std::ceil((hi[k] - lo[k]) / c) + 3;
        fall_[k] = false;
      }
    }
    fpkey_.resize(n);
    fidx_.resize(n);
    for (int64_t i = 0; i < n; ++i) {
      fpkey_[i] = fkey(fcoord((double)ps_[i].x.x, 0), fcoord((double)ps_[i].x.y, 1), fcoord((double)ps_[i].x.z, 2));
      fidx_[i] = i;
    }
    std::stable_sort(fidx_.begin(), fidx_.end(), [&](int64_t a, int64_t b) { return fpkey_[a] < fpkey_[b]; });
    fkey_.resize(n);
    for (int64_t q = 0; q < n; ++q) fkey_[q] = fpkey_[fidx_[q]];
    ftri_.clear();
    const double grow = rmax * 1.001 + 1e-9;
    for (size_t ti = 0; ti < tris_.size(); ++ti) {
      const Triangle<R>& tr = tris_[ti];
      const double v[3][3] = {{(double)tr.v0.x, (double)tr.v0.y, (double)tr.v0.z},
                              {(double)tr.v1.x, (double)tr.v1.y, (double)tr.v1.z},
                              {(double)tr.v2.x, (double)tr.v2.y, (double)tr.v2.z}};
      double e0[3], e1[3], a = 0.0, c = 0.0;
      for (int k = 0; k < 3; ++k) {
        e0[k] = v[1][k] - v[0][k];
        e1[k] = v[2][k] - v[0][k];
        a += e0[k] * e0[k];
        c += e1[k] * e1[k];
      }
      const double s0 = std::max(1.0, std::max(a, c)) * 1.001, t0 = std::max(1.0, c) * 1.001;
      int64_t clo[3], chi[3];
      bool ok = true;
      for (int k = 0; k < 3; ++k) {
        const double pts[5] = {v[0][k], v[1][k], v[2][k], v[0][k] + s0 * e0[k], v[0][k] + t0 * e1[k]};
        double mn = HUGE_VAL, mx = -HUGE_VAL;
        for (double q : pts) {
          mn = std::min(mn, q);
          mx = std::max(mx, q);
        }
        if (!std::isfinite(mn) || !std::isfinite(mx)) ok = false;
        mn -= grow;
        mx += grow;
        const double q0 = std::floor((mn - forg_[k]) / fcs_[k]), q1 = std::floor((mx - forg_[k]) / fcs_[k]);
        if (periodic_[k]) {
          if (q1 - q0 + 1.0 >= (double)fn_[k]) {
            clo[k] = 0;
            chi[k] = fn_[k] - 1;
            continue;
          }
          clo[k] = (int64_t)q0;
          chi[k] = (int64_t)q1;
        } else {
          clo[k] = (int64_t)std::max(0.0, std::min(q0, (double)(fn_[k] - 1)));
          chi[k] = (int64_t)std::max(0.0, std::min(q1, (double)(fn_[k] - 1)));
        }
      }
      if (!ok) continue;  // a non-finite vertex never passes intersect_sphere
      auto wrapc = [&](int64_t q, int k) {
        if (!periodic_[k]) return q;
        q %= fn_[k];
        return q < 0 ? q + fn_[k] : q;
      };
      for (int64_t x = clo[0]; x <= chi[0]; ++x)
        for (int64_t y = clo[1]; y <= chi[1]; ++y)
          for (int64_t z = clo[2]; z <= chi[2]; ++z) {
            const int64_t key = fkey(wrapc(x, 0), wrapc(y, 1), wrapc(z, 2));
            if (!std::binary_search(fkey_.begin(), fkey_.end(), key)) continue;  // no particle there
            std::vector<int32_t>& lst = ftri_[key];
            if (lst.empty() || lst.back() != (int32_t)ti) lst.push_back((int32_t)ti);
          }
    }
  }

  // The walk's visits of particle i: (rank oz*9 + oy*3 + ox of the offset, j).
  void fast_walk(int64_t i, std::vector<std::pair<int, int64_t>>& walk) const {
    walk.clear();
    const Particle<R>& p = ps_[i];
    int32_t ci[3];
    grid_.coords(p.x, ci);
    const double x[3] = {(double)p.x.x, (double)p.x.y, (double)p.x.z};
    int64_t cand[3][3];
    int ncand[3];
    std::vector<int64_t> all[3];
    for (int k = 0; k < 3; ++k) {
      const int64_t c = fcoord(x[k], k);
      ncand[k] = 0;
      if (fall_[k]) {
        for (int64_t q = 0; q < fn_[k]; ++q) all[k].push_back(q);
        continue;
      }
      for (int o = -1; o <= 1; ++o) {
        int64_t q = c + o;
        if (periodic_[k]) q = (q + fn_[k]) % fn_[k];
        else if (q < 0 || q >= fn_[k]) continue;
        cand[k][ncand[k]++] = q;
      }
      for (int m = 0; m < ncand[k]; ++m) all[k].push_back(cand[k][m]);
    }
    for (int64_t a : all[0])
      for (int64_t b : all[1])
        for (int64_t c : all[2]) {
          const int64_t key = fkey(a, b, c);
          auto rg = std::equal_range(fkey_.begin(), fkey_.end(), key);
          for (auto it = rg.first; it != rg.second; ++it) {
            const int64_t j = fidx_[it - fkey_.begin()];
            if (j == i) continue;
            // conservative pre-test: minimum-image distance against the pair's
            // cutoff widened by 1e-6 (a superset of the pair test in either
            // precision); the visit itself evaluates the exact predicate
            {
              const double xj[3] = {(double)ps_[j].x.x, (double)ps_[j].x.y, (double)ps_[j].x.z};
              double d2 = 0.0;
              for (int k = 0; k < 3; ++k) {
                double d = x[k] - xj[k];
                if (periodic_[k]) {
                  const double L = bsz_[k];
                  if (d > 0.5 * L) d -= L;
                  else if (d < -0.5 * L) d += L;
                }
                d2 += d * d;
              }
              const double cut = (adh_ ? 1.5 : 1.0) * ((double)p.r + (double)ps_[j].r) * (1.0 + 1e-6) + 1e-300;
              if (!(d2 < cut * cut) && std::isfinite(d2)) continue;
            }
            int32_t cj[3];
            grid_.coords(ps_[j].x, cj);
            bool vis = true;
            for (int k = 0; k < 3; ++k) vis = vis && cj[k] >= 0 && cj[k] < grid_.dims[k];
            if (!vis) continue;  // not in the reference grid (grid.rs:30-46)
            int mask[3];
            for (int k = 0; k < 3; ++k) {
              mask[k] = 0;
              for (int o = -1; o <= 1; ++o) {
                int32_t nn = Grid<R>::wrap_add(ci[k], o);
                if (periodic_[k]) {
                  if (nn < 0) nn += grid_.dims[k];
                  else if (nn >= grid_.dims[k]) nn -= grid_.dims[k];
                }
                if (nn == cj[k]) mask[k] |= 1 << (o + 1);
              }
            }
            for (int oz = 0; oz < 3; ++oz)
              for (int oy = 0; oy < 3; ++oy)
                for (int ox = 0; ox < 3; ++ox)
                  if ((mask[2] >> oz & 1) && (mask[1] >> oy & 1) && (mask[0] >> ox & 1))
                    walk.emplace_back(oz * 9 + oy * 3 + ox, j);
          }
        }
    std::sort(walk.begin(), walk.end());
  }

  // Physics-fix mode: the indices j != i whose minimum-image distance passes
  // the pre-test, ascending (the pair test itself runs in the visit).
  void fix_walk(int64_t i, std::vector<int64_t>& out) const {
    out.clear();
    const Particle<R>& p = ps_[i];
    const double x[3] = {(double)p.x.x, (double)p.x.y, (double)p.x.z};
    std::vector<int64_t> all[3];
    for (int k = 0; k < 3; ++k) {
      const int64_t c = fcoord(x[k], k);
      if (fall_[k]) {
        for (int64_t q = 0; q < fn_[k]; ++q) all[k].push_back(q);
        continue;
      }
      for (int o = -1; o <= 1; ++o) {
        int64_t q = c + o;
        if (periodic_[k]) q = (q + fn_[k]) % fn_[k];
        else if (q < 0 || q >= fn_[k]) continue;
        all[k].push_back(q);
      }
    }
    for (int64_t a : all[0])
      for (int64_t b : all[1])
        for (int64_t c : all[2]) {
          auto rg = std::equal_range(fkey_.begin(), fkey_.end(), fkey(a, b, c));
          for (auto it = rg.first; it != rg.second; ++it) {
            const int64_t j = fidx_[it - fkey_.begin()];
            if (j == i) continue;
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
              double d = x[k] - (double)(k == 0 ? ps_[j].x.x : (k == 1 ? ps_[j].x.y : ps_[j].x.z));
              if (periodic_[k]) {
                if (d > 0.5 * bsz_[k]) d -= bsz_[k];
                else if (d < -0.5 * bsz_[k]) d += bsz_[k];
              }
              d2 += d * d;
            }
            const double cut = (adh_ ? 1.5 : 1.0) * ((double)p.r + (double)ps_[j].r) * (1.0 + 1e-6) + 1e-300;
            if (!(d2 < cut * cut) && std::isfinite(d2)) continue;
            out.push_back(j);
          }
        }
    std::sort(out.begin(), out.end());
  }

  const std::vector<int32_t>& fast_tris(int64_t i) const {
    static const std::vector<int32_t> none;
    auto it = ftri_.find(fpkey_[i]);
    return it == ftri_.end() ? none : it->second;
  }

  R dt_{};
  V3<R> bmin_{}, bmax_{}, g_{};
  int periodic_[3]{0, 0, 0};
  int normal_{RP_NORMAL_HERTZ}, tangential_{RP_TANGENTIAL_MINDLIN};
  Material<R> pm_{}, wm_{};
  Grid<R> grid_;
  std::vector<Particle<R>> ps_;
  std::vector<Triangle<R>> tris_;
  std::vector<int> tri_mesh_;                  // mesh of each triangle
  std::vector<std::array<double, 6>> motion_;  // per mesh: origin, omega
  std::vector<V3<R>> ftot_, ttot_;
  std::vector<R> fscale_;
  std::vector<std::vector<int64_t>> pairs_;
  bool record_{false};
  int overlap_flag_{0};
  uint64_t jkr_fail_{0};
};

}  // namespace rpo

using rpo::SimBase;

extern "C" {

void* rpo_create(const rp_params* p, int bits) {
  if (!p) return nullptr;
  if (bits == 32) return static_cast<SimBase*>(new rpo::Sim<float>(*p));
  return static_cast<SimBase*>(new rpo::Sim<double>(*p));
}
void rpo_destroy(void* s) { delete static_cast<SimBase*>(s); }
void rpo_set_params(void* s, const rp_params* p) { static_cast<SimBase*>(s)->set_params(*p); }
void rpo_set_cell_size(void* s, double cs) { static_cast<SimBase*>(s)->set_cell_size(cs); }
void rpo_set_state(void* s, size_t n, const double* x, const double* v, const double* w, const double* a,
                   const double* tau, const double* rx, const double* rv, const double* rw, const double* r,
                   const double* m, const uint8_t* fixed) {
  static_cast<SimBase*>(s)->set_state(n, x, v, w, a, tau, rx, rv, rw, r, m, fixed);
}
void rpo_get_state(void* s, double* x, double* v, double* w, double* a, double* tau, double* rx, double* rv,
                   double* rw) {
  static_cast<SimBase*>(s)->get_state(x, v, w, a, tau, rx, rv, rw);
}
size_t rpo_count(void* s) { return static_cast<SimBase*>(s)->count(); }
void rpo_add_mesh(void* s, size_t ntri, const float* v9, const float* n3) {
  static_cast<SimBase*>(s)->add_mesh(ntri, v9, n3);
}
int rpo_set_mesh_motion(void* s, int mesh, const double* origin, const double* omega) {
  return static_cast<SimBase*>(s)->set_mesh_motion(mesh, origin, omega);
}
void rpo_step(void* s, uint64_t nsteps) { static_cast<SimBase*>(s)->step(nsteps); }
void rpo_set_record(void* s, int on) { static_cast<SimBase*>(s)->set_record(on); }
size_t rpo_pair_total(void* s) { return static_cast<SimBase*>(s)->pair_total(); }
void rpo_get_pairs(void* s, int64_t* offsets, int64_t* js) { static_cast<SimBase*>(s)->get_pairs(offsets, js); }
void rpo_get_forces(void* s, double* f3, double* t3, double* fscale) {
  static_cast<SimBase*>(s)->get_forces(f3, t3, fscale);
}
int rpo_overlap_flag(void* s) { return static_cast<SimBase*>(s)->overlap_flag(); }
uint64_t rpo_jkr_failures(void* s) { return static_cast<SimBase*>(s)->jkr_failures(); }
void rpo_cell_ids(void* s, int64_t* out) { static_cast<SimBase*>(s)->cell_ids(out); }
size_t rpo_history_total(void* s) { return static_cast<SimBase*>(s)->history_total(); }
double rpo_time_gather(void* s, size_t nsample, double cell) {
  return static_cast<SimBase*>(s)->time_gather(nsample, cell);
}
void rpo_history(void* s, int64_t* offsets, int64_t* partner, double* xi3) {
  static_cast<SimBase*>(s)->history(offsets, partner, xi3);
}
void rpo_set_fast(void* s, int on) { static_cast<SimBase*>(s)->set_fast(on != 0); }
void rpo_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int rpo_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

// ---- unit-level access to the force laws and the mesh routine (both precisions)
template <class R>
static void normal_law(int model, double overlap, const double* n, const double* vrel, double E, double rstar,
                       double mstar, double rest, double gamma, double* out, double* a_out, int* conv) {
  rpo::V3<R> nn = rpo::v3<R>(R(n[0]), R(n[1]), R(n[2]));
  rpo::V3<R> vv = rpo::v3<R>(R(vrel[0]), R(vrel[1]), R(vrel[2]));
  rpo::V3<R> f{};
  R a = R(0);
  rpo::JkrInfo info{0, 1};
  switch (model) {
    case RP_NORMAL_HERTZ: f = rpo::hertzian_contact<R>(R(overlap), nn, vv, R(E), R(rstar), R(mstar), R(rest)); break;
    case RP_NORMAL_LINEAR: {
      R kn = R(E) * R(rstar);
      R gn = R(0.5) * std::sqrt(R(mstar) * kn);
      f = rpo::linear_spring_dashpot<R>(R(overlap), nn, vv, kn, gn);
      break;
    }
    case RP_NORMAL_HYSTERETIC:
      f = rpo::linear_spring_dashpot<R>(R(overlap), nn, rpo::v3<R>(R(0), R(0), R(0)), R(E) * R(rstar), R(0));
      break;
    case RP_NORMAL_JKR: f = rpo::jkr_force<R>(R(overlap), nn, vv, R(E), R(rstar), R(gamma), &info, &a); break;
    default: f = rpo::sjkr_force<R>(R(overlap), nn, vv, R(E), R(rstar), R(gamma)); break;
  }
  out[0] = (double)f.x;
  out[1] = (double)f.y;
  out[2] = (double)f.z;
  if (a_out) *a_out = (double)a;
  if (conv) *conv = info.converged;
}

// The reference's JKR solve (physics.rs:132-219) in FP64 on n contacts:
// contact radius, iterations of its Newton loop (physics.rs:171-189; the
// iteration that stops on a small step counts) and whether the residual check
// passed.  E is the effective modulus E*.
extern "C" void rpo_jkr_solve(size_t n, const double* overlap, const double* E, const double* rstar,
                              const double* gamma, double* a_out, int* iterations, int* converged) {
  const rpo::V3<double> nn = rpo::v3<double>(0.0, 0.0, 1.0);
  for (size_t k = 0; k < n; ++k) {
    rpo::JkrInfo info{0, 1};
    double a = 0.0;
    (void)rpo::jkr_force<double>(overlap[k], nn, nn, E[k], rstar[k], gamma[k], &info, &a);
    a_out[k] = a;
    iterations[k] = info.iterations;
    converged[k] = info.converged;
  }
}

extern "C" void rpo_normal_law(int bits, int model, double overlap, const double* n, const double* vrel, double E,
                    double rstar, double mstar, double rest, double gamma, double* out, double* a_out,
                    int* converged) {
  if (bits == 32)
    normal_law<float>(model, overlap, n, vrel, E, rstar, mstar, rest, gamma, out, a_out, converged);
  else
    normal_law<double>(model, overlap, n, vrel, E, rstar, mstar, rest, gamma, out, a_out, converged);
}

template <class R>
static void coulomb_law(double fn, const double* vrel, const double* n, double mu, double* out) {
  rpo::V3<R> f = rpo::coulomb_friction<R>(R(fn), rpo::v3<R>(R(vrel[0]), R(vrel[1]), R(vrel[2])),
                                          rpo::v3<R>(R(n[0]), R(n[1]), R(n[2])), R(mu));
  out[0] = (double)f.x;
  out[1] = (double)f.y;
  out[2] = (double)f.z;
}
extern "C" void rpo_coulomb(int bits, double fn, const double* vrel, const double* n, double mu, double* out) {
  if (bits == 32) coulomb_law<float>(fn, vrel, n, mu, out);
  else coulomb_law<double>(fn, vrel, n, mu, out);
}

// linear_spring_coulomb with explicit history xi (in/out), for the spring law's own tests.
template <class R>
static void lsc_law(double fn, const double* vrel, const double* n, double* xi, double kt, double mu, double dt,
                    double* out) {
  rpo::V3<R> x = rpo::v3<R>(R(xi[0]), R(xi[1]), R(xi[2]));
  rpo::V3<R> f = rpo::linear_spring_coulomb<R>(R(fn), rpo::v3<R>(R(vrel[0]), R(vrel[1]), R(vrel[2])),
                                               rpo::v3<R>(R(n[0]), R(n[1]), R(n[2])), x, R(kt), R(mu), R(dt));
  xi[0] = (double)x.x;
  xi[1] = (double)x.y;
  xi[2] = (double)x.z;
  out[0] = (double)f.x;
  out[1] = (double)f.y;
  out[2] = (double)f.z;
}
extern "C" void rpo_lsc(int bits, double fn, const double* vrel, const double* n, double* xi, double kt, double mu, double dt,
             double* out) {
  if (bits == 32) lsc_law<float>(fn, vrel, n, xi, kt, mu, dt, out);
  else lsc_law<double>(fn, vrel, n, xi, kt, mu, dt, out);
}

template <class R>
static int tri_law(const double* tri12, const double* p, double radius, double* cp, double* pen, double* nrm) {
  rpo::Triangle<R> tr;
  tr.v0 = rpo::v3<R>(R(tri12[0]), R(tri12[1]), R(tri12[2]));
  tr.v1 = rpo::v3<R>(R(tri12[3]), R(tri12[4]), R(tri12[5]));
  tr.v2 = rpo::v3<R>(R(tri12[6]), R(tri12[7]), R(tri12[8]));
  tr.n = rpo::v3<R>(R(tri12[9]), R(tri12[10]), R(tri12[11]));
  rpo::V3<R> pp = rpo::v3<R>(R(p[0]), R(p[1]), R(p[2]));
  rpo::V3<R> c = rpo::closest_point(tr, pp);
  cp[0] = (double)c.x;
  cp[1] = (double)c.y;
  cp[2] = (double)c.z;
  R pe = R(0);
  rpo::V3<R> nn{};
  int hit = rpo::intersect_sphere(tr, pp, R(radius), &pe, &nn) ? 1 : 0;
  *pen = (double)pe;
  nrm[0] = (double)nn.x;
  nrm[1] = (double)nn.y;
  nrm[2] = (double)nn.z;
  return hit;
}
extern "C" int rpo_triangle(int bits, const double* tri12, const double* p, double radius, double* cp, double* pen,
                 double* nrm) {
  if (bits == 32) return tri_law<float>(tri12, p, radius, cp, pen, nrm);
  return tri_law<double>(tri12, p, radius, cp, pen, nrm);
}


// ---- unit-level access to the physics-fix laws ------------------------------
template <class R>
static void spring_fix_law(double fn, const double* vrel, const double* n, double* xi, double kt, double mu,
                           double dt, double* out) {
  rpo::V3<R> x = rpo::v3<R>(R(xi[0]), R(xi[1]), R(xi[2]));
  rpo::V3<R> f = rpo::spring_coulomb_fix<R>(R(fn), rpo::v3<R>(R(vrel[0]), R(vrel[1]), R(vrel[2])),
                                            rpo::v3<R>(R(n[0]), R(n[1]), R(n[2])), x, R(kt), R(mu), R(dt));
  xi[0] = (double)x.x;
  xi[1] = (double)x.y;
  xi[2] = (double)x.z;
  out[0] = (double)f.x;
  out[1] = (double)f.y;
  out[2] = (double)f.z;
}
extern "C" void rpo_spring_fix(int bits, double fn, const double* vrel, const double* n, double* xi, double kt,
                               double mu, double dt, double* out) {
  if (bits == 32) spring_fix_law<float>(fn, vrel, n, xi, kt, mu, dt, out);
  else spring_fix_law<double>(fn, vrel, n, xi, kt, mu, dt, out);
}
extern "C" double rpo_gstar_fix(double e1, double nu1, double e2, double nu2) {
  return rpo::effective_shear_fix<double>(e1, nu1, e2, nu2);
}
extern "C" double rpo_mindlin_kt_fix(double gstar, double rstar, double overlap) {
  return rpo::mindlin_kt_fix<double>(gstar, rstar, overlap);
}
extern "C" void rpo_closest_point_fix(const double* tri12, const double* p, double* cp) {
  rpo::Triangle<double> tr;
  tr.v0 = rpo::v3<double>(tri12[0], tri12[1], tri12[2]);
  tr.v1 = rpo::v3<double>(tri12[3], tri12[4], tri12[5]);
  tr.v2 = rpo::v3<double>(tri12[6], tri12[7], tri12[8]);
  tr.n = rpo::v3<double>(tri12[9], tri12[10], tri12[11]);
  rpo::V3<double> c = rpo::closest_point_fix(tr, rpo::v3<double>(p[0], p[1], p[2]));
  cp[0] = c.x;
  cp[1] = c.y;
  cp[2] = c.z;
}
