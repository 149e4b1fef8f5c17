"""Loader and thin object wrapper for librp_hip.so (the C ABI in include/rp_abi.h).

The library is the product: there is no fallback.  If it is missing or a
call fails, a Python exception says so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librp_hip.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)
_i64p = C.POINTER(C.c_int64)
_eng = C.c_void_p


def lib() -> C.CDLL:
    """Load librp_hip.so (once).  Raises ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(
            f"{_LIB_PATH} is missing: build it with `python -m rusty_particles_amd.build` "
            "(the HIP engine has no CPU fallback)")
    L = C.CDLL(_LIB_PATH)
    L.rp_params_default.argtypes = [C.POINTER(abi.RpParams), C.c_double, _dp, _dp]
    L.rp_params_default.restype = None
    L.rp_device_count.argtypes = []
    L.rp_device_count.restype = C.c_int32
    L.rp_engine_create.argtypes = [C.POINTER(abi.RpParams), C.POINTER(_eng)]
    L.rp_engine_set_params.argtypes = [_eng, C.POINTER(abi.RpParams)]
    L.rp_engine_upload.argtypes = [_eng, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _u8p]
    L.rp_engine_upload_full.argtypes = [_eng, C.c_size_t, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        _u8p]
    L.rp_engine_add_mesh.argtypes = [_eng, C.c_size_t, _fp, _fp]
    L.rp_engine_add_mesh_stl.argtypes = [_eng, C.c_char_p]
    L.rp_engine_set_mesh_motion.argtypes = [_eng, C.c_int32, _dp, _dp]
    L.rp_engine_step.argtypes = [_eng, C.c_uint64]
    L.rp_engine_synchronize.argtypes = [_eng]
    L.rp_engine_count.argtypes = [_eng, C.POINTER(C.c_size_t)]
    L.rp_engine_download.argtypes = [_eng, _dp, _dp, _dp, _dp, _dp]
    L.rp_engine_download_aux.argtypes = [_eng, _dp, _dp, _dp, _dp, _dp, _u8p]
    L.rp_engine_set_record.argtypes = [_eng, C.c_int]
    L.rp_stl_read.argtypes = [C.c_char_p, C.POINTER(C.c_size_t), _fp, _fp]
    L.rp_engine_pairs.argtypes = [_eng, _i64p, _i64p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.rp_engine_cells.argtypes = [_eng, C.c_int32, _i64p, C.POINTER(C.c_uint32), _i64p, _dp]
    L.rp_engine_history.argtypes = [_eng, _i64p, _i64p, _dp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.rp_engine_set_history.argtypes = [_eng, _i64p, _i64p, _dp, C.c_size_t]
    L.rp_engine_stats.argtypes = [_eng, C.POINTER(abi.RpStats)]
    L.rp_engine_set_timing.argtypes = [_eng, C.c_int]
    L.rp_engine_timing.argtypes = [_eng, _dp, C.POINTER(C.c_uint64)]
    L.rp_probe_copy.argtypes = [C.c_size_t, _dp]
    L.rp_engine_set_jkr_solver.argtypes = [_eng, C.c_int32, C.POINTER(abi.RpJkrEnvelope), _fp, C.c_int32]
    L.rp_engine_jkr_stats.argtypes = [_eng, C.POINTER(abi.RpJkrStats)]
    L.rp_engine_upload_batch.argtypes = [_eng, C.c_int32, C.POINTER(C.c_size_t), C.POINTER(abi.RpMaterial),
                                         C.POINTER(abi.RpMaterial), _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                         _dp, _u8p]
    L.rp_engine_instance_counts.argtypes = [_eng, C.POINTER(C.c_size_t)]
    L.rp_engine_set_graph.argtypes = [_eng, C.c_int]
    L.rp_engine_set_tiling.argtypes = [_eng, C.c_int32]
    L.rp_engine_get_tiling.argtypes = [_eng, C.POINTER(C.c_int32)]
    L.rp_engine_autotune.argtypes = [_eng, C.c_uint32, _dp, _dp]
    L.rp_engine_destroy.argtypes = [_eng]
    L.rp_engine_destroy.restype = None
    L.rp_last_error.argtypes = []
    L.rp_last_error.restype = C.c_char_p
    for name in ("rp_engine_create", "rp_engine_set_params", "rp_engine_upload", "rp_engine_upload_full",
                 "rp_engine_add_mesh", "rp_engine_add_mesh_stl", "rp_engine_set_mesh_motion", "rp_engine_step", "rp_engine_synchronize",
                 "rp_engine_count", "rp_engine_download", "rp_engine_download_aux", "rp_engine_set_record",
                 "rp_stl_read", "rp_engine_pairs", "rp_engine_cells", "rp_engine_history", "rp_engine_set_history", "rp_engine_stats", "rp_engine_set_timing",
                 "rp_engine_timing", "rp_probe_copy", "rp_engine_set_jkr_solver", "rp_engine_jkr_stats",
                 "rp_engine_upload_batch", "rp_engine_instance_counts", "rp_engine_set_graph",
                 "rp_engine_autotune", "rp_engine_set_tiling", "rp_engine_get_tiling"):
        getattr(L, name).restype = C.c_int32
    _lib = L
    return L


class EngineError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rp error {code}: {msg}")
        self.code = code


def _check(code: int) -> None:
    if code != abi.RP_OK:
        msg = lib().rp_last_error()
        raise EngineError(code, msg.decode() if msg else "")


def _ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def _vec(a, n: int) -> np.ndarray | None:
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n, 3))


def device_count() -> int:
    return int(lib().rp_device_count())


def probe_copy(nbytes: int) -> float:
    """Runs the library's plain copy kernel over nbytes; returns milliseconds."""
    ms = C.c_double(0.0)
    _check(lib().rp_probe_copy(int(nbytes), C.byref(ms)))
    return ms.value


def read_stl(path: str) -> tuple[np.ndarray, np.ndarray]:
    """Triangles of an STL file: (ntri, 9) vertices and (ntri, 3) normals, float32.

    Raises OSError like the reference's add_mesh (src/lib.rs:87-91)."""
    L = lib()
    n = C.c_size_t(0)
    code = L.rp_stl_read(path.encode(), C.byref(n), None, None)
    if code == abi.RP_ERR_IO:
        raise OSError(f"cannot read STL file: {path}")
    _check(code)
    v9 = np.zeros((n.value, 9), dtype=np.float32)
    n3 = np.zeros((n.value, 3), dtype=np.float32)
    if n.value:
        _check(L.rp_stl_read(path.encode(), C.byref(n), _ptr(v9, C.c_float), _ptr(n3, C.c_float)))
    return v9, n3


class Engine:
    """One rp_engine (one domain on the current HIP device)."""

    def __init__(self, params: abi.RpParams):
        self._L = lib()
        h = _eng()
        _check(self._L.rp_engine_create(C.byref(params), C.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.rp_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: abi.RpParams) -> None:
        _check(self._L.rp_engine_set_params(self._h, C.byref(params)))

    def upload(self, x, radius, mass, fixed=None, v=None, w=None, a=None, tau=None, rx=None, rv=None,
               rw=None) -> None:
        x = np.asarray(x, dtype=np.float64)
        n = int(x.reshape(-1, 3).shape[0])
        xs = _vec(x, n)
        r = np.ascontiguousarray(np.asarray(radius, dtype=np.float64).reshape(n))
        m = np.ascontiguousarray(np.asarray(mass, dtype=np.float64).reshape(n))
        f = None if fixed is None else np.ascontiguousarray(np.asarray(fixed, dtype=np.uint8).reshape(n))
        arrs = [_vec(q, n) for q in (v, w, a, tau, rx, rv, rw)]
        _check(self._L.rp_engine_upload_full(self._h, n, _ptr(xs), *[_ptr(q) for q in arrs], _ptr(r), _ptr(m),
                                             _ptr(f, C.c_uint8)))
        self._ninst = 1

    def upload_batch(self, counts, materials, x, radius, mass, fixed=None, v=None, w=None,
                     wall_materials=None, a=None, tau=None, rx=None, rv=None, rw=None) -> None:
        """Independent instances in one engine: counts[k] particles of
        instance k (arrays concatenated in instance order), each with its own
        particle material (abi.RpMaterial) and optionally wall material."""
        counts = [int(c) for c in counts]
        k = len(counts)
        n = sum(counts)
        xs = _vec(np.asarray(x, dtype=np.float64), n)
        r = np.ascontiguousarray(np.asarray(radius, dtype=np.float64).reshape(n))
        m = np.ascontiguousarray(np.asarray(mass, dtype=np.float64).reshape(n))
        f = None if fixed is None else np.ascontiguousarray(np.asarray(fixed, dtype=np.uint8).reshape(n))
        arrs = [_vec(q, n) for q in (v, w, a, tau, rx, rv, rw)]
        cnt = (C.c_size_t * k)(*counts)
        pm = (abi.RpMaterial * k)(*materials)
        wm = None if wall_materials is None else (abi.RpMaterial * k)(*wall_materials)
        _check(self._L.rp_engine_upload_batch(self._h, k, cnt, pm, wm, _ptr(xs), *[_ptr(q) for q in arrs],
                                              _ptr(r), _ptr(m), _ptr(f, C.c_uint8)))
        self._ninst = k

    def instance_counts(self) -> np.ndarray:
        k = getattr(self, "_ninst", 1)
        cnt = (C.c_size_t * k)()
        _check(self._L.rp_engine_instance_counts(self._h, cnt))
        return np.array(list(cnt), dtype=np.int64)

    def add_mesh(self, v9, n3) -> None:
        v9 = np.ascontiguousarray(np.asarray(v9, dtype=np.float32).reshape(-1, 9))
        n3 = np.ascontiguousarray(np.asarray(n3, dtype=np.float32).reshape(-1, 3))
        _check(self._L.rp_engine_add_mesh(self._h, v9.shape[0], _ptr(v9, C.c_float), _ptr(n3, C.c_float)))

    def set_mesh_motion(self, mesh: int, origin, omega) -> None:
        """Surface velocity of mesh `mesh` (0-based, in the order added): the
        rigid rotation omega (rad/s) about origin; the triangles stay put."""
        o = np.ascontiguousarray(origin, dtype=np.float64).reshape(3)
        w = np.ascontiguousarray(omega, dtype=np.float64).reshape(3)
        _check(self._L.rp_engine_set_mesh_motion(self._h, int(mesh), _ptr(o, C.c_double), _ptr(w, C.c_double)))

    def step(self, nsteps: int = 1) -> None:
        _check(self._L.rp_engine_step(self._h, int(nsteps)))

    def synchronize(self) -> None:
        _check(self._L.rp_engine_synchronize(self._h))

    def count(self) -> int:
        n = C.c_size_t(0)
        _check(self._L.rp_engine_count(self._h, C.byref(n)))
        return int(n.value)

    def download(self) -> dict:
        n = self.count()
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "w", "a", "tau", "rx", "rv", "rw")}
        out["radius"] = np.zeros(n)
        out["mass"] = np.zeros(n)
        out["fixed"] = np.zeros(n, dtype=np.uint8)
        _check(self._L.rp_engine_download(self._h, *[_ptr(out[k]) for k in ("x", "v", "w", "a", "tau")]))
        _check(self._L.rp_engine_download_aux(self._h, _ptr(out["rx"]), _ptr(out["rv"]), _ptr(out["rw"]),
                                              _ptr(out["radius"]), _ptr(out["mass"]),
                                              _ptr(out["fixed"], C.c_uint8)))
        return out

    def set_record(self, on: bool) -> None:
        _check(self._L.rp_engine_set_record(self._h, 1 if on else 0))

    def pairs(self) -> tuple[np.ndarray, np.ndarray]:
        """(offsets[n+1], js[total]) of the last recorded step, in the reference's neighbour order."""
        tot = C.c_size_t(0)
        _check(self._L.rp_engine_pairs(self._h, None, None, 0, C.byref(tot)))
        n = self.count()
        offs = np.zeros(n + 1, dtype=np.int64)
        js = np.zeros(max(tot.value, 1), dtype=np.int64)
        _check(self._L.rp_engine_pairs(self._h, _ptr(offs, C.c_int64), _ptr(js, C.c_int64), js.size,
                                       C.byref(tot)))
        return offs, js[: tot.value]

    def history(self):
        """Physics-fix tangential history (rp_engine_history): (offsets[n+1],
        partner API index, displacement[total, 3]), per particle by ascending partner."""
        tot = C.c_size_t(0)
        _check(self._L.rp_engine_history(self._h, None, None, None, 0, C.byref(tot)))
        n = self.count()
        offs = np.zeros(n + 1, dtype=np.int64)
        part = np.zeros(max(tot.value, 1), dtype=np.int64)
        xi = np.zeros((max(tot.value, 1), 3))
        _check(self._L.rp_engine_history(self._h, _ptr(offs, C.c_int64), _ptr(part, C.c_int64), _ptr(xi),
                                         part.size, C.byref(tot)))
        return offs, part[:tot.value], xi[:tot.value]

    def set_history(self, offsets, partner, xi) -> None:
        """Loads a tangential history in the layout of history()
        (rp_engine_set_history), e.g. one read before a re-upload."""
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        part = np.ascontiguousarray(partner, dtype=np.int64).reshape(-1)
        x = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
        if offs.size != self.count() + 1:
            raise ValueError("offsets needs one entry per particle plus one")
        if x.shape[0] != part.size:
            raise ValueError("partner and xi differ in length")
        _check(self._L.rp_engine_set_history(self._h, _ptr(offs, C.c_int64), _ptr(part, C.c_int64), _ptr(x),
                                             part.size))

    def cells(self, rebuild: bool = True) -> dict:
        """Broad-phase cells (rp_engine_cells): per particle in API order the
        reference-grid cell index (-1 outside) and the engine's fine-cell key;
        the slot -> API permutation of the last rebuild (forced first when
        rebuild); the fine grid's geometry."""
        n = self.count()
        ref = np.zeros(n, dtype=np.int64)
        key = np.zeros(n, dtype=np.uint32)
        perm = np.zeros(n, dtype=np.int64)
        grid = np.zeros(22)
        _check(self._L.rp_engine_cells(self._h, 1 if rebuild else 0, _ptr(ref, C.c_int64), _ptr(key, C.c_uint32),
                                       _ptr(perm, C.c_int64), _ptr(grid)))
        g = {"origin": grid[0:3], "inv_cell": grid[3:6], "cells": grid[6:9].astype(int),
             "shift": grid[9:12].astype(int), "hi": grid[12:15].astype(int), "lo": grid[15:18].astype(int),
             "periodic": grid[18:21].astype(int), "ncells": int(grid[21])}
        return {"ref_cell": ref, "fine_key": key, "slot_api": perm, "grid": g}

    def autotune(self, steps: int = 20) -> tuple[float, float]:
        """Pick the fastest Verlet skin and cell size for the current state,
        then the faster tiling of the force kernel's zero-force pass (the
        state itself is restored); returns (skin, cell_size), 0 = automatic;
        tiling() gives the tiling."""
        skin, cell = C.c_double(0.0), C.c_double(0.0)
        _check(self._L.rp_engine_autotune(self._h, int(steps), C.byref(skin), C.byref(cell)))
        return skin.value, cell.value

    def set_tiling(self, tiling: str) -> None:
        """'wave' (default): a wavefront deals its particles' list entries over
        its lanes in the zero-force pass; 'lane': one lane per particle."""
        _check(self._L.rp_engine_set_tiling(self._h, abi.TILINGS[tiling]))

    def tiling(self) -> str:
        t = C.c_int32(0)
        _check(self._L.rp_engine_get_tiling(self._h, C.byref(t)))
        return {v: k for k, v in abi.TILINGS.items()}[t.value]

    def set_graph(self, on: bool) -> None:
        """Replay captured 10-launch step loops (default on)."""
        _check(self._L.rp_engine_set_graph(self._h, 1 if on else 0))

    def set_timing(self, on: bool) -> None:
        _check(self._L.rp_engine_set_timing(self._h, 1 if on else 0))

    def timing(self) -> dict:
        """Summed HIP-event times (ms) and launch counts since the last call."""
        ms = np.zeros(abi.RP_TIMING_KINDS)
        cnt = (C.c_uint64 * abi.RP_TIMING_KINDS)()
        _check(self._L.rp_engine_timing(self._h, _ptr(ms), cnt))
        names = ("force", "rebuild", "begin", "collect", "network")
        return {names[k]: (float(ms[k]), int(cnt[k])) for k in range(abi.RP_TIMING_KINDS)}

    def set_jkr_solver(self, solver: str, surrogate=None, check: bool = False) -> None:
        """'exact' (the reference's Newton solve), 'shortcut' (identical
        results, no iterations for contacts that cannot have a root) or
        'surrogate' (needs a trained surrogate.JkrSurrogate)."""
        code = abi.JKR_SOLVERS[solver]
        env = params = None
        if code == abi.RP_JKR_SURROGATE:
            if surrogate is None:
                raise ValueError("the surrogate solver needs a trained network")
            env = C.byref(surrogate.env)
            self._surr_params = np.ascontiguousarray(surrogate.params, dtype=np.float32)
            params = self._surr_params.ctypes.data_as(C.POINTER(C.c_float))
        _check(self._L.rp_engine_set_jkr_solver(self._h, code, env, params, 1 if check else 0))

    def jkr_stats(self) -> dict:
        s = abi.RpJkrStats()
        _check(self._L.rp_engine_jkr_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in abi.RpJkrStats._fields_}

    def stats(self) -> dict:
        s = abi.RpStats()
        _check(self._L.rp_engine_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in abi.RpStats._fields_ if name != "reserved"}
