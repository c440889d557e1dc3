"""ctypes binding of librunko_amd.so (the C ABI declared in include/runko_amd.h).

The HIP library is the only compute path: if it is missing or no GPU is
visible, every compute call raises.  There is no CPU fallback.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librunko_amd.so")

RK_MAX_SPECIES = 4

RK_OK, RK_EINVAL, RK_EHIP, RK_EOVERFLOW, RK_ESTATE = 0, 1, 2, 3, 4


class RkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class WorldDesc(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int32 * 3),
        ("n_tiles", ctypes.c_int32 * 3),
        ("tile_offset", ctypes.c_int32 * 3),
        ("grid_tiles", ctypes.c_int32 * 3),
        ("brick", ctypes.c_int32 * 3),
        ("periodic", ctypes.c_int32),
        ("n_species", ctypes.c_int32),
        ("current_filter", ctypes.c_int32),
        ("cfl", ctypes.c_double),
        ("charge", ctypes.c_double * RK_MAX_SPECIES),
        ("mass", ctypes.c_double * RK_MAX_SPECIES),
    ]


class ReflectorWall(ctypes.Structure):
    """rk_reflector_wall"""
    _fields_ = [("walloc", ctypes.c_float), ("betawall", ctypes.c_float), ("gammawall", ctypes.c_float)]


class EdgeBC(ctypes.Structure):
    """rk_edge_bc"""
    _fields_ = [("direction", ctypes.c_int32), ("side", ctypes.c_int32), ("position", ctypes.c_float),
                ("E", ctypes.c_float * 3), ("B", ctypes.c_float * 3), ("J", ctypes.c_float * 3),
                ("E_components", ctypes.c_int32), ("B_components", ctypes.c_int32),
                ("J_components", ctypes.c_int32)]


class DiffDesc(ctypes.Structure):
    """rk_diff_desc"""
    _fields_ = [
        ("n", ctypes.c_int32 * 3),
        ("n_species", ctypes.c_int32),
        ("count", ctypes.c_int64 * RK_MAX_SPECIES),
        ("charge", ctypes.c_double * RK_MAX_SPECIES),
        ("mass", ctypes.c_double * RK_MAX_SPECIES),
        ("cfl", ctypes.c_double),
        ("filter_passes", ctypes.c_int32),
        ("precision", ctypes.c_int32),
    ]


class DiffState(ctypes.Structure):
    """rk_diff_state"""
    _fields_ = [("E", ctypes.c_void_p * 3), ("B", ctypes.c_void_p * 3), ("x", ctypes.c_void_p * 3),
                ("u", ctypes.c_void_p * 3)]


_lib = None

P3 = ctypes.c_void_p * 3


def lib():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"runko_amd: HIP library not built ({LIB_PATH} missing); run __graft_entry__.build() "
            "or python -m runko_amd._build")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    L.rk_last_error.restype = ctypes.c_char_p
    L.rk_version.restype = ctypes.c_int
    L.rk_world_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(WorldDesc)]
    L.rk_world_destroy.argtypes = [vp]
    L.rk_world_field_nodes.argtypes = [vp]
    L.rk_world_field_nodes.restype = i64
    L.rk_world_num_bricks.argtypes = [vp]
    L.rk_world_num_bricks.restype = i32
    L.rk_world_set_fields.argtypes = [vp, P3, P3, P3, P3]
    L.rk_world_j_pointer.argtypes = [vp, i32]
    L.rk_world_j_pointer.restype = vp
    L.rk_world_set_species.argtypes = [vp, i32, i64, P3, P3, vp, vp, vp, vp]
    L.rk_world_set_migration.argtypes = [vp, i32, i64, P3, P3, vp, vp]
    L.rk_world_set_mailbox.argtypes = [vp, i64, P3, P3, vp, vp]
    L.rk_world_set_current_staging.argtypes = [vp, vp, i64]
    L.rk_reflect_particles.argtypes = [vp, i32, vp, i32, vp]
    L.rk_apply_edge_bc.argtypes = [vp, i32, ctypes.POINTER(EdgeBC), i32, vp]
    for name in ["rk_push_half_b", "rk_push_e", "rk_add_current", "rk_filter_current", "rk_push_particles",
                 "rk_deposit_current", "rk_pack_outgoing_particles", "rk_sort_particles"]:
        f = getattr(L, name)
        f.argtypes = [vp, i32, vp]
    L.rk_particle_step_fused.argtypes = [vp, vp]
    L.rk_filter_current_passes.argtypes = [vp, i32, vp]
    L.rk_particle_step_fused_walls.argtypes = [vp, vp, i32, vp]
    L.rk_world_set_fused_launch.argtypes = [vp, i32]
    L.rk_world_set_fused_workgroup.argtypes = [vp, i32]
    L.rk_world_set_walls.argtypes = [vp, vp, i32]
    L.rk_particle_step_fused_device_walls.argtypes = [vp, vp]
    L.rk_advance_walls.argtypes = [vp, vp]
    L.rk_read_walls.argtypes = [vp, vp, i32, vp]
    L.rk_push_e_add_current.argtypes = [vp, i32, vp]
    L.rk_profile_enable.argtypes = [vp, i32]
    L.rk_profile_read.argtypes = [vp, vp, vp]
    L.rk_comm_local.argtypes = [vp, i32, vp]
    L.rk_inject.argtypes = [vp, i32, i32, i64, P3, P3, vp, vp]
    L.rk_pending_overflow.argtypes = [vp, ctypes.POINTER(i64), vp]
    L.rk_repack.argtypes = [vp, i32, i64, P3, P3, vp, vp, vp, vp]
    L.rk_retry_append.argtypes = [vp, vp]
    L.rk_move_bins.argtypes = [vp, i32, i32, vp, vp, vp, vp]
    L.rk_relayout.argtypes = [vp, i32, vp, vp, i64, vp, vp]
    L.rk_tile_particle_count.argtypes = [vp, i32, i32, ctypes.POINTER(i64), vp]
    L.rk_tile_particles.argtypes = [vp, i32, i32, P3, P3, vp, i64, ctypes.POINTER(i64), vp]
    L.rk_outgoing.argtypes = [vp, P3, P3, vp, vp, vp, vp, i64, ctypes.POINTER(i64), vp]
    L.rk_energies.argtypes = [vp, i32, vp, vp]
    L.rk_energies_device.argtypes = [vp, i32, vp, vp]
    L.rk_gauss_residual.argtypes = [vp, i32, vp, vp]
    L.rk_error_counters.argtypes = [vp, vp, vp]
    L.rk_diff_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(DiffDesc)]
    L.rk_diff_create_ensemble.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(DiffDesc), i32]
    L.rk_diff_destroy.argtypes = [vp]
    L.rk_diff_step.argtypes = [vp, ctypes.POINTER(DiffState), ctypes.POINTER(DiffState), vp]
    L.rk_diff_step_adjoint.argtypes = [vp, ctypes.POINTER(DiffState), ctypes.POINTER(DiffState),
                                       ctypes.POINTER(DiffState), vp]
    L.rk_diff_error_count.argtypes = [vp, ctypes.POINTER(i64), vp]
    L.rk_snapshot_fields.argtypes = [vp, i32, i32, vp, vp]
    L.rk_snapshot_spectra.argtypes = [vp, i32, i32, ctypes.c_float, ctypes.c_float, i32, vp, vp]
    L.rk_interpolate_at.argtypes = [vp, i32, i64, P3, P3, P3, vp]
    L.rk_world_set_schemes.argtypes = [vp, i32, vp]
    _lib = L
    return _lib


def check(rc):
    if rc != RK_OK:
        raise RkError(rc, lib().rk_last_error().decode())
    return rc


def p3(tensors):
    return P3(*[t.data_ptr() for t in tensors])


# exported symbols declared in include/runko_amd.h (checked by the CPU tests)
EXPORTED = [
    "rk_last_error", "rk_version", "rk_world_create", "rk_world_destroy", "rk_world_field_nodes",
    "rk_world_num_bricks", "rk_world_set_fields", "rk_world_j_pointer", "rk_world_set_species",
    "rk_world_set_migration", "rk_push_half_b", "rk_push_e", "rk_add_current", "rk_filter_current",
    "rk_push_particles", "rk_deposit_current", "rk_pack_outgoing_particles", "rk_sort_particles",
    "rk_filter_current_passes", "rk_particle_step_fused", "rk_particle_step_fused_walls", "rk_world_set_fused_launch", "rk_world_set_fused_workgroup", "rk_world_set_walls",
    "rk_particle_step_fused_device_walls", "rk_advance_walls", "rk_read_walls", "rk_comm_local", "rk_inject", "rk_pending_overflow", "rk_bin_demand",
    "rk_repack", "rk_retry_append", "rk_move_bins", "rk_relayout", "rk_tile_particle_count", "rk_tile_particles", "rk_outgoing",
    "rk_energies", "rk_energies_device", "rk_gauss_residual", "rk_error_counters", "rk_push_e_add_current", "rk_profile_enable",
    "rk_profile_read", "rk_world_set_mailbox", "rk_world_set_current_staging", "rk_reflect_particles",
    "rk_apply_edge_bc", "rk_diff_create", "rk_diff_create_ensemble", "rk_diff_destroy", "rk_diff_step", "rk_diff_step_adjoint",
    "rk_diff_error_count", "rk_snapshot_fields", "rk_snapshot_spectra", "rk_interpolate_at",
    "rk_world_set_schemes",
]

PROF_CATEGORIES = ["fused_particle_step", "push", "deposit", "pack", "append", "fdtd", "filter", "ghost", "other"]
