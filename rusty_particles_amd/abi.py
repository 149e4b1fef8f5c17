"""ctypes mirror of include/rp_abi.h (types and constants only, no library load)."""
from __future__ import annotations

import ctypes as C

RP_OK = 0
RP_ERR_INVALID = -1
RP_ERR_DEVICE = -2
RP_ERR_NOMEM = -3
RP_ERR_STATE = -4
RP_ERR_IO = -5

# Enum discriminants of src/physics.rs:6-28 (an ABI already in src/gpu.rs:108).
RP_NORMAL_LINEAR = 0
RP_NORMAL_HERTZ = 1
RP_NORMAL_HYSTERETIC = 2
RP_NORMAL_JKR = 3
RP_NORMAL_SJKR = 4
RP_TANGENTIAL_COULOMB = 0
RP_TANGENTIAL_LSC = 1
RP_TANGENTIAL_MINDLIN = 2
RP_MODE_PARITY = 0
RP_MODE_PHYSICS_FIX = 1
RP_CELL_ORDER_AUTO, RP_CELL_ORDER_ROW, RP_CELL_ORDER_PENCIL = 0, 1, 2

# Model names accepted by set_contact_models (src/lib.rs:160-180), lower-cased.
NORMAL_MODELS = {
    "hertz": RP_NORMAL_HERTZ,
    "hertzian": RP_NORMAL_HERTZ,
    "linear": RP_NORMAL_LINEAR,
    "spring_dashpot": RP_NORMAL_LINEAR,
    "hysteretic": RP_NORMAL_HYSTERETIC,
    "jkr": RP_NORMAL_JKR,
    "sjkr": RP_NORMAL_SJKR,
    "simplified_jkr": RP_NORMAL_SJKR,
}
TANGENTIAL_MODELS = {
    "mindlin": RP_TANGENTIAL_MINDLIN,
    "coulomb": RP_TANGENTIAL_COULOMB,
    "linear": RP_TANGENTIAL_LSC,
    "spring_coulomb": RP_TANGENTIAL_LSC,
}


class RpMaterial(C.Structure):
    _fields_ = [
        ("youngs_modulus", C.c_double),
        ("poissons_ratio", C.c_double),
        ("density", C.c_double),
        ("friction_coefficient", C.c_double),
        ("restitution_coefficient", C.c_double),
        ("surface_energy", C.c_double),
    ]


class RpParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("bounds_min", C.c_double * 3),
        ("bounds_max", C.c_double * 3),
        ("periodic", C.c_int32 * 3),
        ("normal_model", C.c_int32),
        ("tangential_model", C.c_int32),
        ("mode", C.c_int32),
        ("particle", RpMaterial),
        ("wall", RpMaterial),
        ("gravity", C.c_double * 3),
        ("ref_cell_size", C.c_double),
        ("cell_size", C.c_double),
        ("skin", C.c_double),
        ("cell_order", C.c_int32),
        ("reserved0", C.c_int32),
    ]


RP_MLP_PARAMS = 4610
RP_TIMING_KINDS = 5
RP_JKR_EXACT, RP_JKR_SHORTCUT, RP_JKR_SURROGATE = 0, 1, 2
RP_TILING_LANE, RP_TILING_WAVE = 0, 1
TILINGS = {"lane": RP_TILING_LANE, "wave": RP_TILING_WAVE}
JKR_SOLVERS = {"exact": RP_JKR_EXACT, "shortcut": RP_JKR_SHORTCUT, "surrogate": RP_JKR_SURROGATE}


class RpJkrStats(C.Structure):
    _fields_ = [
        ("served", C.c_uint64),
        ("fallback", C.c_uint64),
        ("rows_max", C.c_uint64),
        ("rows_capacity", C.c_uint64),
        ("err_max", C.c_double),
        ("err_mean", C.c_double),
        ("raw_max", C.c_double),
        ("raw_mean", C.c_double),
    ]


class RpJkrEnvelope(C.Structure):
    _fields_ = [
        ("d_lo", C.c_double), ("d_hi", C.c_double),
        ("rstar_lo", C.c_double), ("rstar_hi", C.c_double),
        ("estar_lo", C.c_double), ("estar_hi", C.c_double),
        ("gamma_lo", C.c_double), ("gamma_hi", C.c_double),
    ]


class RpStats(C.Structure):
    _fields_ = [
        ("steps", C.c_uint64),
        ("rebuilds", C.c_uint64),
        ("contacts", C.c_uint64),
        ("wall_contacts", C.c_uint64),
        ("removed", C.c_uint64),
        ("jkr_failures", C.c_uint64),
        ("excessive_overlap", C.c_int32),
        ("neighbour_capacity", C.c_int32),
        ("max_neighbours", C.c_int32),
        ("reserved", C.c_int32),
        ("overlap_steps", C.c_uint64),
        ("history_contacts", C.c_uint64),
    ]


def default_params(dt: float, bmin, bmax) -> RpParams:
    """rp_params_default in Python (defaults of Simulation::new, src/simulation.rs:62-91)."""
    p = RpParams()
    p.dt = dt
    for k in range(3):
        p.bounds_min[k] = bmin[k]
        p.bounds_max[k] = bmax[k]
        p.periodic[k] = 0
    p.normal_model = RP_NORMAL_HERTZ
    p.tangential_model = RP_TANGENTIAL_MINDLIN
    p.mode = RP_MODE_PARITY
    p.particle = RpMaterial(1e7, 0.3, 2500.0, 0.5, 0.5, 0.0)
    p.wall = RpMaterial(1e9, 0.3, 7800.0, 0.5, 0.5, 0.0)
    p.gravity[0], p.gravity[1], p.gravity[2] = 0.0, -9.81, 0.0
    p.ref_cell_size = 1.0
    p.cell_size = 0.0
    p.skin = 0.0
    return p


# Every exported symbol of librp_hip.so declared in include/rp_abi.h.
EXPORTS = (
    "rp_params_default",
    "rp_device_count",
    "rp_engine_create",
    "rp_engine_set_params",
    "rp_engine_upload",
    "rp_engine_upload_full",
    "rp_engine_add_mesh",
    "rp_engine_add_mesh_stl",
    "rp_engine_set_mesh_motion",
    "rp_engine_step",
    "rp_engine_synchronize",
    "rp_engine_count",
    "rp_engine_download",
    "rp_engine_download_aux",
    "rp_engine_set_record",
    "rp_stl_read",
    "rp_engine_pairs",
    "rp_engine_cells",
    "rp_engine_history",
    "rp_engine_set_history",
    "rp_engine_stats",
    "rp_engine_set_timing",
    "rp_engine_timing",
    "rp_probe_copy",
    "rp_jkr_surrogate_train",
    "rp_jkr_samples",
    "rp_jkr_envelope_scale",
    "rp_mlp_infer",
    "rp_mlp_infer_noalloc",
    "rp_mlp_gradient",
    "rp_mfma_probe",
    "rp_surrogate_last_error",
    "rp_engine_set_jkr_solver",
    "rp_engine_upload_batch",
    "rp_engine_set_graph",
    "rp_engine_autotune",
    "rp_engine_set_tiling",
    "rp_engine_get_tiling",
    "rp_engine_instance_counts",
    "rp_engine_jkr_stats",
    "rp_engine_destroy",
    "rp_last_error",
)
