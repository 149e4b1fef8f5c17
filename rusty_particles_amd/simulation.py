"""Python surface of rusty_particles (src/lib.rs:15-277) over the HIP engine.

Method names, argument order, defaults and exception types follow the pyo3
class `Simulation` of the reference.  Differences, all on purpose:

* The time step runs on the GPU engine (librp_hip.so).  The first step()
  creates the engine if enable_gpu() was not called; with no HIP device it
  raises RuntimeError, the error enable_gpu raises in the reference
  (src/lib.rs:146-148).  There is no CPU time step in this package.
* State is FP64.  Setter values are kept as given (the reference rounds them
  to f32); run() still computes its step count in f32 as lib.rs:183 does.
* Setters and add_particle after the engine exists take effect (the
  reference's GPU path snapshots them at enable_gpu, src/gpu.rs:92-115), and
  getters return the engine's current state.
"""
from __future__ import annotations

import math
import sys

import numpy as np

from . import abi
from . import native

_threads_set = False


def set_num_threads(num_threads: int) -> None:
    """src/lib.rs:15-22.  The reference sizes rayon's global pool, which can be
    built once; a second call raises RuntimeError.  The GPU step has no host
    thread pool, so only that contract is kept."""
    global _threads_set
    if _threads_set:
        raise RuntimeError("Failed to set num threads: The global thread pool has already been initialized.")
    if int(num_threads) < 0:
        raise OverflowError("can't convert negative int to unsigned")
    _threads_set = True


def hello() -> str:
    """src/lib.rs:28-31."""
    return "Hello from Rust!"


_FIELDS3 = ("x", "v", "w", "a", "tau", "rx", "rv", "rw")

# meshes whose surface may move at once (rp_engine_set_mesh_motion)
MAX_MOVING_MESHES = 4

# src/simulation.rs:529
OVERLAP_WARNING = "Warning: Excessive particle overlap detected (>20% of radius). Simulation may be unstable."


class Simulation:
    """The reference's `Simulation` class (src/lib.rs:41-277)."""

    def __init__(self, dt: float, bmin, bmax, particle_count: int = 0):
        self._params = abi.default_params(float(dt), bmin, bmax)
        self._dt32 = np.float32(dt)
        n0 = 0
        self._h = {k: np.zeros((n0, 3)) for k in _FIELDS3}
        self._h["radius"] = np.zeros(n0)
        self._h["mass"] = np.zeros(n0)
        self._h["fixed"] = np.zeros(n0, dtype=np.uint8)
        self._host_valid = True      # host arrays hold the latest state
        self._engine_valid = False   # engine holds the latest state
        self._params_dirty = True
        self._meshes: list[tuple[np.ndarray, np.ndarray]] = []
        self._motions: dict[int, tuple[tuple, tuple]] = {}  # mesh -> (origin, omega)
        self._engine: native.Engine | None = None
        self._record = False
        self._jkr = ("exact", None, False)  # solver, surrogate, check
        self._ovl_seen = 0  # overlap_steps of the engine already warned about
        # physics-fix tangential history read with the host state (offsets,
        # partner, xi), loaded again after a re-upload
        self._hist: tuple[np.ndarray, np.ndarray, np.ndarray] | None = None
        # bumped by every change of state, parameters or meshes made through
        # this object (an Ensemble re-uploads a member whose version moved)
        self._version = 0

    # ---- construction (src/lib.rs:57-80) ---------------------------------------
    @staticmethod
    def create(dt, min_x, min_y, min_z, max_x, max_y, max_z, particle_count):
        return Simulation(dt, (min_x, min_y, min_z), (max_x, max_y, max_z), particle_count)

    def add_particle(self, x, y, z, radius, mass, fixed=False):
        self._sync_host()
        h = self._h
        h["x"] = np.vstack([h["x"], [[float(x), float(y), float(z)]]])
        for k in ("v", "w", "a", "tau", "rx", "rv", "rw"):
            h[k] = np.vstack([h[k], np.zeros((1, 3))])
        h["radius"] = np.append(h["radius"], float(radius))
        h["mass"] = np.append(h["mass"], float(mass))
        h["fixed"] = np.append(h["fixed"], np.uint8(1 if fixed else 0))
        self._extend_history(1)
        self._engine_valid = False
        self._version += 1

    def add_particles(self, positions, radii, masses, fixed=None):
        """Bulk form of add_particle (same semantics, one call for many particles)."""
        self._sync_host()
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        h = self._h
        h["x"] = np.vstack([h["x"], pos])
        for k in ("v", "w", "a", "tau", "rx", "rv", "rw"):
            h[k] = np.vstack([h[k], np.zeros((n, 3))])
        h["radius"] = np.concatenate([h["radius"], np.broadcast_to(np.asarray(radii, dtype=np.float64), (n,))])
        h["mass"] = np.concatenate([h["mass"], np.broadcast_to(np.asarray(masses, dtype=np.float64), (n,))])
        f = np.zeros(n, dtype=np.uint8) if fixed is None else np.asarray(fixed, dtype=np.uint8).reshape(n)
        h["fixed"] = np.concatenate([h["fixed"], f])
        self._extend_history(n)
        self._engine_valid = False
        self._version += 1

    def add_mesh(self, path):
        """src/lib.rs:87-91 (IOError when the file cannot be read)."""
        v9, n3 = native.read_stl(str(path))
        self.add_mesh_arrays(v9, n3)

    def add_mesh_arrays(self, v9, n3):
        """Appends triangles given as (ntri, 9) vertices and (ntri, 3) normals."""
        v9 = np.asarray(v9, dtype=np.float32).reshape(-1, 9)
        n3 = np.asarray(n3, dtype=np.float32).reshape(-1, 3)
        self._meshes.append((v9, n3))
        self._version += 1
        if self._engine is not None:
            self._engine.add_mesh(v9, n3)

    def set_mesh_motion(self, index, origin, angular_velocity):
        """Extension (the reference's walls are at rest, src/simulation.rs:475-479):
        mesh `index` (0-based, in the order added) moves as the rigid rotation
        `angular_velocity` (rad/s) about the point `origin`.  Only the surface
        velocity enters the wall contacts; the triangles stay where they are, which
        is exact for surfaces of revolution about the axis (a rotating drum)."""
        index = int(index)
        if not 0 <= index < len(self._meshes):
            raise IndexError("mesh index out of range")
        o = tuple(float(c) for c in origin)
        w = tuple(float(c) for c in angular_velocity)
        if len(o) != 3 or len(w) != 3:
            raise ValueError("origin and angular_velocity are 3-vectors")
        if not all(math.isfinite(c) for c in (*o, *w)):
            raise ValueError("origin and angular_velocity must be finite")
        moving = {k for k, (_, ww) in self._motions.items() if any(ww)}
        moving.discard(index)
        if any(w) and len(moving) >= MAX_MOVING_MESHES:
            raise ValueError(f"at most {MAX_MOVING_MESHES} meshes move at once")
        if self._engine is not None:
            self._engine.set_mesh_motion(index, o, w)
        self._motions[index] = (o, w)
        self._version += 1

    # ---- stepping (src/lib.rs:94-96, 182-195) --------------------------------------
    def step(self):
        self._steps(1)

    def run(self, duration):
        steps = int(math.ceil(float(np.float32(duration) / self._dt32)))
        if steps > 0:
            self._steps(steps)

    def _steps(self, n: int) -> None:
        eng = self._prepare_engine()
        eng.step(n)
        self._host_valid = False
        self._version += 1
        self._warn_overlap(eng)

    def _warn_overlap(self, eng: native.Engine) -> None:
        """The reference prints this line once for every step in which a pair
        or a wall contact overlaps by more than 20 % of the smaller radius
        (src/simulation.rs:361-370, 469-473, 528-530)."""
        k = int(eng.stats()["overlap_steps"])
        for _ in range(k - self._ovl_seen):
            print(OVERLAP_WARNING, file=sys.stderr)
        self._ovl_seen = k

    def _prepare_engine(self) -> native.Engine:
        eng = self._ensure_engine()
        if self._params_dirty:
            eng.set_params(self._params)
            self._params_dirty = False
        if not self._engine_valid:
            h = self._h
            eng.upload(h["x"], h["radius"], h["mass"], h["fixed"], v=h["v"], w=h["w"], a=h["a"], tau=h["tau"],
                       rx=h["rx"], rv=h["rv"], rw=h["rw"])
            if self._keeps_history() and self._hist is not None and self._hist[1].size:
                eng.set_history(*self._hist)
            self._engine_valid = True
            self._ovl_seen = 0  # the upload starts the engine's counters afresh
        return eng

    # ---- materials, models, boundaries (src/lib.rs:98-180) --------------------------
    def set_particle_material(self, youngs_modulus, poissons_ratio, density, friction_coefficient,
                              restitution_coefficient, surface_energy=0.0):
        self._params.particle = abi.RpMaterial(youngs_modulus, poissons_ratio, density, friction_coefficient,
                                               restitution_coefficient, surface_energy)
        self._params_changed()

    def set_wall_material(self, youngs_modulus, poissons_ratio, density, friction_coefficient,
                          restitution_coefficient, surface_energy=0.0):
        self._params.wall = abi.RpMaterial(youngs_modulus, poissons_ratio, density, friction_coefficient,
                                           restitution_coefficient, surface_energy)
        self._params_changed()

    def set_periodic_boundary(self, x, y, z):
        for k, flag in enumerate((x, y, z)):
            self._params.periodic[k] = 1 if flag else 0
        self._params_changed()

    def set_contact_models(self, normal, tangential):
        nm = abi.NORMAL_MODELS.get(str(normal).lower())
        if nm is None:
            raise ValueError("Invalid normal force model")
        tm = abi.TANGENTIAL_MODELS.get(str(tangential).lower())
        if tm is None:
            raise ValueError("Invalid tangential force model")
        self._params.normal_model = nm
        self._params.tangential_model = tm
        self._params_changed()

    def set_gravity(self, gx, gy, gz):
        """Extension: the reference hard-codes (0, -9.81, 0) (src/simulation.rs:267)."""
        self._params.gravity[0], self._params.gravity[1], self._params.gravity[2] = gx, gy, gz
        self._params_changed()

    def set_engine_tuning(self, cell_size=0.0, skin=0.0):
        """Extension: broad-phase cell size and Verlet skin (0 = automatic)."""
        self._params.cell_size = float(cell_size)
        self._params.skin = float(skin)
        self._params_changed()

    def set_physics_fix(self, on=True):
        """Extension (no counterpart in the reference): the physics-fix mode
        (RP_MODE_PHYSICS_FIX) -- the shear modulus from E and nu, Mindlin's
        stiffness only under compression, a tangential history per contact
        that persists across steps and re-uploads, Eberly's closest point,
        every pair inside the cutoff once.  Off (the default): the
        reference's arithmetic, quirks included."""
        self._params.mode = abi.RP_MODE_PHYSICS_FIX if on else abi.RP_MODE_PARITY
        self._params_changed()

    def set_jkr_solver(self, solver="exact", surrogate=None, check=False):
        """Extension: 'exact' (the reference's Newton solve, default),
        'shortcut' (the same results; contacts that cannot have a root skip
        the iterations) or 'surrogate' (a trained surrogate.JkrSurrogate,
        FP64-polished, exact fallback).  check=True also runs the exact solve
        for every served contact and records the force errors (jkr_stats)."""
        if solver not in abi.JKR_SOLVERS:
            raise ValueError("Invalid JKR solver")
        if solver == "surrogate" and surrogate is None:
            raise ValueError("the surrogate solver needs a trained network")
        self._jkr = (solver, surrogate, bool(check))
        self._version += 1
        if self._engine is not None:
            self._engine.set_jkr_solver(solver, surrogate, check=check)

    def jkr_stats(self) -> dict:
        """Extension: surrogate counters of the last step (served, fallback, errors)."""
        if self._engine is None:
            return {}
        return self._engine.jkr_stats()

    def autotune(self, steps=20):
        """Extension: picks the fastest Verlet skin and cell size for the current
        state (the state is unchanged; results never depend on them)."""
        self._prepare_engine()
        skin, cell = self._engine.autotune(steps)
        self._params.skin, self._params.cell_size = skin, cell
        return skin, cell

    def _params_changed(self) -> None:
        self._params_dirty = True
        self._version += 1

    def enable_gpu(self):
        """src/lib.rs:146-148: RuntimeError when no device can be used."""
        self._ensure_engine()

    def init_mpi(self):
        """src/lib.rs:150-152.  One GPU holds the whole domain: no slab decomposition."""
        print("init_mpi: single-GPU engine, the whole domain stays on one device", file=sys.stderr)

    # ---- output and getters (src/lib.rs:202-276) --------------------------------------
    def write_vtk(self, filename):
        """Legacy ASCII VTK POLYDATA, layout of src/vtk.rs:5-44."""
        self._sync_host()
        h = self._h
        n = h["x"].shape[0]
        f = lambda v: np.format_float_positional(np.float32(v), unique=True, trim="-")
        try:
            with open(filename, "w") as fh:
                fh.write("# vtk DataFile Version 3.0\nDEM Simulation Data\nASCII\nDATASET POLYDATA\n")
                fh.write(f"POINTS {n} float\n")
                for p in h["x"]:
                    fh.write(f"{f(p[0])} {f(p[1])} {f(p[2])}\n")
                fh.write(f"POINT_DATA {n}\nVECTORS velocity float\n")
                for p in h["v"]:
                    fh.write(f"{f(p[0])} {f(p[1])} {f(p[2])}\n")
                fh.write("SCALARS radius float 1\nLOOKUP_TABLE default\n")
                for r in h["radius"]:
                    fh.write(f"{f(r)}\n")
                fh.write("SCALARS id int 1\nLOOKUP_TABLE default\n")
                for i in range(n):
                    fh.write(f"{i}\n")
        except OSError as e:
            raise OSError(str(e)) from e

    def get_particle_count(self):
        if not self._host_valid and self._engine is not None:
            return self._engine.count()
        return int(self._h["x"].shape[0])

    def _get3(self, key, idx):
        self._sync_host()
        idx = int(idx)
        if idx < 0 or idx >= self._h["x"].shape[0]:
            raise IndexError("Particle index out of bounds")
        p = self._h[key][idx]
        return (float(p[0]), float(p[1]), float(p[2]))

    def get_particle_position(self, id):
        return self._get3("x", id)

    def get_particle_velocity(self, id):
        return self._get3("v", id)

    def get_particle_angular_velocity(self, index):
        return self._get3("w", index)

    def _set3(self, key, idx, a, b, c):
        self._sync_host()
        idx = int(idx)
        if idx < 0 or idx >= self._h["x"].shape[0]:
            raise IndexError("Particle index out of bounds")
        self._h[key][idx] = (float(a), float(b), float(c))
        self._engine_valid = False
        self._version += 1

    def set_particle_velocity(self, id, vx, vy, vz):
        self._set3("v", id, vx, vy, vz)

    def set_particle_angular_velocity(self, index, x, y, z):
        self._set3("w", index, x, y, z)

    # ---- extensions for bulk access and diagnostics --------------------------------------
    def state(self) -> dict:
        """Copy of the full FP64 state in API order (x, v, w, a, tau, residuals, radius, mass, fixed)."""
        self._sync_host()
        return {k: np.array(v) for k, v in self._h.items()}

    def set_record_pairs(self, on=True):
        self._record = bool(on)
        if self._engine is not None:
            self._engine.set_record(self._record)

    def contact_pairs(self):
        """(offsets, js) pair list of the last step (needs set_record_pairs(True) before stepping)."""
        if self._engine is None:
            raise RuntimeError("no step has run on the engine")
        return self._engine.pairs()

    def stats(self) -> dict:
        if self._engine is None:
            return {}
        return self._engine.stats()

    # ---- internals ---------------------------------------------------------------------
    def _ensure_engine(self) -> native.Engine:
        if self._engine is None:
            try:
                eng = native.Engine(self._params)
            except native.EngineError as e:
                raise RuntimeError(str(e)) from e
            # meshes, motions and solver first: the engine is published only
            # once all of them applied (a failure leaves no engine behind)
            try:
                for v9, n3 in self._meshes:
                    eng.add_mesh(v9, n3)
                for k, (o, w) in self._motions.items():
                    eng.set_mesh_motion(k, o, w)
                eng.set_record(self._record)
                if self._jkr[0] != "exact":
                    eng.set_jkr_solver(*self._jkr[:2], check=self._jkr[2])
            except Exception:
                eng.close()
                raise
            self._engine = eng
            self._params_dirty = False
            self._engine_valid = False
        return self._engine

    def _sync_host(self) -> None:
        if self._host_valid:
            return
        st = self._engine.download()
        self._h = {k: st[k] for k in (*_FIELDS3, "radius", "mass", "fixed")}
        self._hist = self._engine.history() if self._keeps_history() else None
        self._host_valid = True

    def _keeps_history(self) -> bool:
        p = self._params
        return p.mode == abi.RP_MODE_PHYSICS_FIX and p.tangential_model != abi.RP_TANGENTIAL_COULOMB

    def _extend_history(self, n: int) -> None:
        """n particles appended: they start without history."""
        if self._hist is not None:
            offs, part, xi = self._hist
            self._hist = (np.concatenate([offs, np.full(n, offs[-1], dtype=np.int64)]), part, xi)


class Ensemble:
    """Extension (no counterpart in the reference): steps several Simulation
    objects together on one batched engine -- one set of kernel launches per
    step over all of them.  The simulations must agree on everything but
    their particles and materials (dt, bounds, periodic axes, contact models,
    gravity, meshes and their motions, engine tuning, JKR solver); each
    evolves exactly as it would alone (with the JKR surrogate: as it would
    alone with the surrogate, to its force-error bound).  After run()/step()
    every member's getters return its new state.  A member changed between
    two ensemble steps (a setter, add_particle, or a step of its own) makes
    the next ensemble step upload every member's current state again.  In the
    physics-fix mode each member's tangential history goes with its state,
    into the ensemble and back."""

    def __init__(self, sims):
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("an ensemble needs at least one simulation")
        self._check_members()
        self._engine: native.Engine | None = None
        self._versions: list[int] | None = None
        self._dt32 = self.sims[0]._dt32

    def _check_members(self) -> None:
        p0, m0 = self.sims[0]._params, self.sims[0]._meshes
        for s in self.sims[1:]:
            p = s._params
            same = (p.dt == p0.dt and list(p.bounds_min) == list(p0.bounds_min)
                    and list(p.bounds_max) == list(p0.bounds_max) and list(p.periodic) == list(p0.periodic)
                    and p.normal_model == p0.normal_model and p.tangential_model == p0.tangential_model
                    and p.mode == p0.mode and list(p.gravity) == list(p0.gravity)
                    and p.ref_cell_size == p0.ref_cell_size and p.cell_size == p0.cell_size
                    and p.skin == p0.skin and p.cell_order == p0.cell_order)
            if not same:
                raise ValueError("ensemble members must share dt, bounds, periodic axes, models, mode, gravity and "
                                 "tuning")
            if len(s._meshes) != len(m0) or any(not (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]))
                                                for a, b in zip(s._meshes, m0)):
                raise ValueError("ensemble members must share their meshes")
            if s._motions != self.sims[0]._motions:
                raise ValueError("ensemble members must share their meshes' motions")
            j0, j = self.sims[0]._jkr, s._jkr
            if j[0] != j0[0] or j[1] is not j0[1] or j[2] != j0[2]:
                raise ValueError("ensemble members must share their JKR solver (and surrogate network)")

    def step(self):
        self._steps(1)

    def run(self, duration):
        steps = int(math.ceil(float(np.float32(duration) / self._dt32)))
        if steps > 0:
            self._steps(steps)

    def _upload(self) -> None:
        self._check_members()
        if self._engine is not None:
            self._engine.close()
            self._engine = None
        try:
            eng = native.Engine(abi.RpParams.from_buffer_copy(self.sims[0]._params))
        except native.EngineError as e:
            raise RuntimeError(str(e)) from e
        for s in self.sims:
            s._sync_host()
        hs = [s._h for s in self.sims]
        cat = lambda k: np.concatenate([h[k] for h in hs])
        eng.upload_batch([h["x"].shape[0] for h in hs], [s._params.particle for s in self.sims], cat("x"),
                         cat("radius"), cat("mass"), cat("fixed"), v=cat("v"), w=cat("w"), a=cat("a"),
                         tau=cat("tau"), rx=cat("rx"), rv=cat("rv"), rw=cat("rw"),
                         wall_materials=[s._params.wall for s in self.sims])
        for v9, n3 in self.sims[0]._meshes:
            eng.add_mesh(v9, n3)
        for k, (o, w) in self.sims[0]._motions.items():
            eng.set_mesh_motion(k, o, w)
        solver, net, check = self.sims[0]._jkr
        if solver != "exact":
            eng.set_jkr_solver(solver, net, check=check)
        if self.sims[0]._keeps_history():
            # the members' tables, partners shifted to the batch's indices
            offs, part, xi = [np.zeros(1, dtype=np.int64)], [], []
            ne = npart = 0
            for s, h in zip(self.sims, hs):
                n = h["x"].shape[0]
                if s._hist is not None:
                    o_, p_, x_ = s._hist
                    offs.append(o_[1:] + ne)
                    part.append(np.where(p_ >= 0, p_ + npart, -1))
                    xi.append(x_)
                    ne += int(o_[-1])
                else:
                    offs.append(np.full(n, ne, dtype=np.int64))
                npart += n
            if ne:
                eng.set_history(np.concatenate(offs), np.concatenate(part), np.concatenate(xi))
        self._engine = eng

    def _steps(self, n: int) -> None:
        if self._engine is None or self._versions != [s._version for s in self.sims]:
            self._upload()
        self._engine.step(n)
        # members see their new state through their own getters
        counts = self._engine.instance_counts()
        st = self._engine.download()
        hist = self._engine.history() if self.sims[0]._keeps_history() else None
        o = 0
        for s, c in zip(self.sims, counts):
            c = int(c)
            s._h = {k: np.array(st[k][o:o + c]) for k in (*_FIELDS3, "radius", "mass", "fixed")}
            s._hist = None
            if hist is not None:
                a, b = int(hist[0][o]), int(hist[0][o + c])
                p_ = hist[1][a:b]
                s._hist = (hist[0][o:o + c + 1] - a, np.where(p_ >= 0, p_ - o, -1), np.array(hist[2][a:b]))
            s._host_valid = True
            s._engine_valid = False
            o += c
        self._versions = [s._version for s in self.sims]
