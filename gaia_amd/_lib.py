"""ctypes binding of include/gaia_vbd.h (libgaia_vbd.so).

The library must exist in-tree (gaia_amd/libgaia_vbd.so, built by
gaia_amd/build.py); a missing library is an error, never a fallback.
"""
import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libgaia_vbd.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "gaia_vbd.h")

GVBD_OK = 0
GVBD_COLOR_MCS = 0
GVBD_COLOR_GREEDY = 1

c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
c_float_p = ctypes.POINTER(ctypes.c_float)


class PhysicsParams(ctypes.Structure):
    """gvbd_physics_params."""

    _fields_ = [
        ("gravity", ctypes.c_float * 3),
        ("time_step", ctypes.c_float),
        ("num_substeps", ctypes.c_int32),
        ("iterations", ctypes.c_int32),
        ("world_lower", ctypes.c_float * 3),
        ("world_upper", ctypes.c_float * 3),
        ("use_plane_ground", ctypes.c_int32),
        ("boundary_collision_stiffness", ctypes.c_float),
        ("boundary_friction_dynamic", ctypes.c_float),
        ("boundary_friction_eps_v", ctypes.c_float),
        ("step_size", ctypes.c_float),
        ("step_size_gd", ctypes.c_float),
        ("friction_start_iter", ctypes.c_int32),
        ("associate_gravity_with_inertia", ctypes.c_int32),
        ("cloth_boundary_contact_stiffness", ctypes.c_float),
        ("cloth_boundary_friction_eps", ctypes.c_float),
        ("degenerate_triangle_thres", ctypes.c_float),
        ("cloth_step_size", ctypes.c_float),
        ("handle_collision", ctypes.c_int32),
        ("contact_stiffness", ctypes.c_float),
        ("contact_radius", ctypes.c_float),
        ("thickness", ctypes.c_float),
        ("conservative_step_relaxation", ctypes.c_float),
        ("contact_energy_type", ctypes.c_int32),
        ("contact_query_radius", ctypes.c_float),
        ("contact_bvh_rebuild_steps", ctypes.c_int32),
        ("use_accelerator", ctypes.c_int32),
        ("accelerator_rho", ctypes.c_float),
        ("tet_contact_model", ctypes.c_int32),
        ("allow_dcd", ctypes.c_int32),
        ("allow_ccd", ctypes.c_int32),
        ("handle_self_collision", ctypes.c_int32),
        ("collision_stiffness", ctypes.c_float),
        ("collision_off_height", ctypes.c_float),
        ("check_feasible_region", ctypes.c_int32),
        ("check_tet_traverse", ctypes.c_int32),
        ("tet_traverse_non_self", ctypes.c_int32),
        ("shift_query_point_to_center", ctypes.c_int32),
        ("stop_traversing_after_passing", ctypes.c_int32),
        ("feasible_region_epsilon", ctypes.c_float),
        ("ray_tri_intersection_epsilon", ctypes.c_float),
        ("max_search_distance_multiplier", ctypes.c_float),
        ("center_shift_level", ctypes.c_float),
        ("dcd_bvh_rebuild_steps", ctypes.c_int32),
    ]


class TetMeshDesc(ctypes.Structure):
    """gvbd_tet_mesh_desc."""

    _fields_ = [
        ("num_vertices", ctypes.c_int32),
        ("positions", c_float_p),
        ("velocities", c_float_p),
        ("initial_velocity", ctypes.c_float * 3),
        ("num_tets", ctypes.c_int32),
        ("tets", c_int32_p),
        ("density", ctypes.c_float),
        ("mu", ctypes.c_float),
        ("lmbd", ctypes.c_float),
        ("damping_hydrostatic", ctypes.c_float),
        ("damping_deviatoric", ctypes.c_float),
        ("exponential_vel_damping", ctypes.c_float),
        ("constant_vel_damping", ctypes.c_float),
        ("initialization_type", ctypes.c_int32),
        ("init_ratio_g", ctypes.c_float),
        ("num_fixed", ctypes.c_int32),
        ("fixed_points", c_int32_p),
        ("vertex_colors", c_int32_p),
        ("coloring_algorithm", ctypes.c_int32),
        ("friction_dynamic", ctypes.c_float),
        ("friction_eps_v", ctypes.c_float),
        ("has_no_grav_zone", ctypes.c_int32),
        ("no_grav_zone_threshold", ctypes.c_float),
        ("max_velocity_magnitude", ctypes.c_float),
    ]


class ClothMeshDesc(ctypes.Structure):
    """gvbd_cloth_mesh_desc."""

    _fields_ = [
        ("num_vertices", ctypes.c_int32),
        ("positions", c_float_p),
        ("velocities", c_float_p),
        ("initial_velocity", ctypes.c_float * 3),
        ("num_faces", ctypes.c_int32),
        ("faces", c_int32_p),
        ("density", ctypes.c_float),
        ("mu", ctypes.c_float),
        ("lmbd", ctypes.c_float),
        ("bending_stiffness", ctypes.c_float),
        ("damping_stvk", ctypes.c_float),
        ("damping_bending", ctypes.c_float),
        ("exponential_vel_damping", ctypes.c_float),
        ("constant_vel_damping", ctypes.c_float),
        ("initialization_type", ctypes.c_int32),
        ("num_fixed", ctypes.c_int32),
        ("fixed_points", c_int32_p),
        ("vertex_colors", c_int32_p),
        ("coloring_algorithm", ctypes.c_int32),
        ("friction_dynamic", ctypes.c_float),
        ("friction_eps_v", ctypes.c_float),
        ("has_no_grav_zone", ctypes.c_int32),
        ("no_grav_zone_threshold", ctypes.c_float),
        ("max_velocity_magnitude", ctypes.c_float),
    ]


MODE_GROUP_LAUNCHES = 0
MODE_WORK_QUEUE = 1


class Schedule(ctypes.Structure):
    """gvbd_schedule."""
    _fields_ = [
        ("lanes_per_vertex", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("vertices_per_entry", ctypes.c_int32),
        ("blocks_per_cu", ctypes.c_int32),
        ("spin_limit", ctypes.c_int32),
        ("stage_records", ctypes.c_int32),
        ("xcd_blocks", ctypes.c_int32),
    ]


class ContactStats(ctypes.Structure):
    """gvbd_contact_stats."""
    _fields_ = [
        ("frames", ctypes.c_int64),
        ("iterations", ctypes.c_int64),
        ("detections", ctypes.c_int64),
        ("rebuilds", ctypes.c_int64),
        ("detect_ms", ctypes.c_double),
        ("vf_pairs", ctypes.c_int64),
        ("ee_pairs", ctypes.c_int64),
    ]


class DeformerDesc(ctypes.Structure):
    """gvbd_deformer_desc."""
    _fields_ = [
        ("kind", ctypes.c_int32),
        ("center", ctypes.c_float * 3),
        ("axis", ctypes.c_float * 3),
        ("speed", ctypes.c_float * 3),
        ("start_time", ctypes.c_float),
        ("angular_velocity", ctypes.c_double),
        ("end_time", ctypes.c_double),
        ("num_meshes", ctypes.c_int32),
        ("meshes", ctypes.POINTER(ctypes.c_int32)),
        ("num_vertices", ctypes.POINTER(ctypes.c_int32)),
        ("vertices", ctypes.POINTER(ctypes.c_int32)),
    ]


class AttachmentDesc(ctypes.Structure):
    """gvbd_attachment_desc."""
    _fields_ = [
        ("mesh", ctypes.c_int32),
        ("vertex", ctypes.c_int32),
        ("target_mesh", ctypes.c_int32),
        ("face", ctypes.c_int32 * 3),
        ("bary", ctypes.c_float * 3),
        ("stiffness", ctypes.c_float),
    ]


class ServerStats(ctypes.Structure):
    """gvbd_server_stats."""
    _fields_ = [
        ("budget_bytes", ctypes.c_int64),
        ("resident_bytes", ctypes.c_int64),
        ("vertex_solves", ctypes.c_int64),
        ("served", ctypes.c_int64),
        ("admissions", ctypes.c_int64),
        ("evictions", ctypes.c_int64),
        ("ticks", ctypes.c_int64),
        ("resident", ctypes.c_int32),
        ("pending", ctypes.c_int32),
        ("waiting", ctypes.c_int32),
        ("last_stepped", ctypes.c_int32),
    ]


class ServedStep(ctypes.Structure):
    """gvbd_served_step."""
    _fields_ = [
        ("tag", ctypes.c_int64),
        ("scene", ctypes.c_int32),
        ("frame", ctypes.c_int32),
        ("latency_ms", ctypes.c_double),
        ("queue_ms", ctypes.c_double),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int32
_SCHED = ctypes.POINTER(Schedule)
# name -> argtypes (all return int unless listed in _RESTYPES)
_SIGNATURES = {
    "gvbd_abi_struct_sizes": [c_int32_p, c_int32_p, c_int32_p, c_int32_p],
    "gvbd_context_create": [_I, ctypes.POINTER(_P)],
    "gvbd_context_destroy": [_P],
    "gvbd_last_error": [_P],
    "gvbd_synchronize": [_P],
    "gvbd_context_memory": [_P, c_int64_p],
    "gvbd_scene_device_bytes": [_P, _I, c_int64_p, c_int64_p],
    "gvbd_scene_create": [_P, ctypes.POINTER(PhysicsParams), c_int32_p],
    "gvbd_scene_destroy": [_P, _I],
    "gvbd_scene_add_tet_mesh": [_P, _I, ctypes.POINTER(TetMeshDesc), c_int32_p],
    "gvbd_scene_add_cloth_mesh": [_P, _I, ctypes.POINTER(ClothMeshDesc), c_int32_p],
    "gvbd_scene_initialize": [_P, _I],
    "gvbd_scene_load_json": [_P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, c_int32_p],
    "gvbd_scene_num_meshes": [_P, _I, c_int32_p],
    "gvbd_scene_mesh_info": [_P, _I, _I, c_int32_p, c_int32_p],
    "gvbd_scene_num_parallel_groups": [_P, _I, c_int32_p],
    "gvbd_scene_parallel_group": [_P, _I, _I, c_int32_p, c_int32_p],
    "gvbd_scene_mesh_constants": [_P, _I, _I, c_float_p, c_float_p, c_float_p],
    "gvbd_scene_set_schedule": [_P, _I, _I, _I],
    "gvbd_scene_set_schedule_ex": [_P, _I, _SCHED],
    "gvbd_scene_get_schedule": [_P, _I, _SCHED],
    "gvbd_scene_run_step": [_P, _I],
    "gvbd_scene_run_steps": [_P, _I, _I],
    "gvbd_scene_apply_initial_step": [_P, _I],
    "gvbd_scene_solve_parallel_group": [_P, _I, _I, _I],
    "gvbd_scene_run_iteration": [_P, _I, _I],
    "gvbd_scene_update_velocities": [_P, _I],
    "gvbd_scene_collision_detection": [_P, _I],
    "gvbd_scene_contact_stats": [_P, _I, ctypes.POINTER(ContactStats)],
    "gvbd_scene_solve_failures": [_P, _I, _I, ctypes.POINTER(ctypes.c_int64)],
    "gvbd_scene_add_deformer": [_P, _I, ctypes.POINTER(DeformerDesc)],
    "gvbd_scene_add_attachments": [_P, _I, ctypes.POINTER(AttachmentDesc), _I],
    "gvbd_scene_get_time": [_P, _I, ctypes.POINTER(ctypes.c_float)],
    "gvbd_scene_set_time": [_P, _I, ctypes.c_float],
    "gvbd_scene_evaluate_convergence": [_P, _I] + [ctypes.POINTER(ctypes.c_double)] * 4,
    "gvbd_scene_evaluate_convergence_noalloc": [_P, _I] + [ctypes.POINTER(ctypes.c_double)] * 4,
    "gvbd_scene_penetration_results": [_P, _I, c_int32_p, c_int32_p, ctypes.POINTER(ctypes.c_uint8), c_int32_p],
    "gvbd_scene_record_graph": [_P, _I, _I],
    "gvbd_scene_get_mesh_state": [_P, _I, _I, c_float_p, c_float_p],
    "gvbd_scene_set_mesh_state": [_P, _I, _I, c_float_p, c_float_p],
    "gvbd_scene_save_state": [_P, _I, ctypes.c_char_p, _I, ctypes.c_double, _I],
    "gvbd_scene_load_state": [_P, _I, ctypes.c_char_p, c_int32_p, ctypes.POINTER(ctypes.c_double)],
    "gvbd_scene_write_ply": [_P, _I, ctypes.c_char_p, _I],
    "gvbd_scene_write_binary_frame": [_P, _I, ctypes.c_char_p],
    "gvbd_state_write": [ctypes.c_char_p, _I, ctypes.c_double, _I, _I, c_int32_p, c_float_p, c_float_p],
    "gvbd_state_read": [ctypes.c_char_p, _I, c_int32_p, c_int32_p, ctypes.c_int64, c_float_p, c_float_p, c_int32_p,
                        ctypes.POINTER(ctypes.c_double)],
    "gvbd_write_ply": [ctypes.c_char_p, _I, c_float_p, _I, c_int32_p],
    "gvbd_batch_create": [_P, c_int32_p, _I, c_int32_p],
    "gvbd_batch_destroy": [_P, _I],
    "gvbd_batch_set_schedule": [_P, _I, _I, _I],
    "gvbd_batch_set_schedule_ex": [_P, _I, _SCHED],
    "gvbd_batch_get_schedule": [_P, _I, _SCHED],
    "gvbd_batch_queue_info": [_P, _I, c_int32_p, c_int32_p],
    "gvbd_batch_record_graph": [_P, _I, _I],
    "gvbd_batch_run_steps": [_P, _I, _I],
    "gvbd_batch_apply_initial_step": [_P, _I],
    "gvbd_batch_run_iteration": [_P, _I, _I],
    "gvbd_batch_update_velocities": [_P, _I],
    "gvbd_batch_solve_parallel_group": [_P, _I, _I, _I],
    "gvbd_batch_num_parallel_groups": [_P, _I, c_int32_p],
    "gvbd_batch_group_work": [_P, _I, _I, c_int64_p, c_int64_p],
    "gvbd_scene_clone": [_P, _I, c_int32_p],
    "gvbd_server_create": [_P, ctypes.c_int64, c_int32_p],
    "gvbd_server_destroy": [_P, _I],
    "gvbd_server_set_schedule": [_P, _I, _SCHED],
    "gvbd_server_add_scene": [_P, _I, _I],
    "gvbd_server_remove_scene": [_P, _I, _I],
    "gvbd_server_request_step": [_P, _I, _I, ctypes.c_int64],
    "gvbd_server_tick": [_P, _I, ctypes.POINTER(ServerStats)],
    "gvbd_server_poll": [_P, _I, ctypes.POINTER(ServedStep), _I, c_int32_p],
    "gvbd_server_get_stats": [_P, _I, ctypes.POINTER(ServerStats)],
    "gvbd_server_scene_resident": [_P, _I, _I, c_int32_p],
    "gvbd_cd_create": [_P, ctypes.c_float, _I, _I, c_int32_p],
    "gvbd_cd_destroy": [_P, _I],
    "gvbd_cd_add_mesh": [_P, _I, _I, c_float_p, _I, c_int32_p, c_int32_p],
    "gvbd_cd_finalize": [_P, _I],
    "gvbd_cd_sizes": [_P, _I, c_int32_p, c_int32_p, c_int32_p],
    "gvbd_cd_mesh_offsets": [_P, _I, _I, c_int32_p, c_int32_p, c_int32_p],
    "gvbd_cd_topology": [_P, _I, c_int32_p, c_int32_p],
    "gvbd_cd_set_positions": [_P, _I, _I, c_float_p],
    "gvbd_cd_get_positions": [_P, _I, _I, c_float_p],
    "gvbd_cd_perturb": [_P, _I, ctypes.c_uint64, _I, ctypes.c_float, c_float_p],
    "gvbd_cd_update_bvh": [_P, _I, _I],
    "gvbd_cd_query": [_P, _I, _I],
    "gvbd_cd_bounds": [_P, _I, ctypes.c_float, ctypes.c_float, c_float_p],
    "gvbd_cd_counts": [_P, _I, c_int64_p],
    "gvbd_cd_vf_pairs": [_P, _I, c_int32_p, c_float_p],
    "gvbd_cd_vf_contacts": [_P, _I, c_int32_p, c_float_p],
    "gvbd_cd_ee_contacts": [_P, _I, c_int32_p, c_float_p, c_float_p],
    "gvbd_cd_vf_results": [_P, _I, c_int32_p, _P],
    "gvbd_cd_ee_results": [_P, _I, c_int32_p, _P],
    "gvbd_cd_min_distances": [_P, _I, c_float_p, c_float_p, c_float_p],
    "gvbd_stream_copy": [_P, ctypes.c_int64, _I, c_float_p],
    "gvbd_copy_bandwidth": [_P, ctypes.c_int64, _I, _I, c_float_p],
    "gvbd_event_record": [_P, _I],
    "gvbd_event_elapsed_ms": [_P, _I, _I, c_float_p],
    "gvbd_tet_mesh_constants": [_I, c_float_p, _I, c_int32_p, ctypes.c_float, c_float_p, c_float_p, c_float_p],
    "gvbd_color_graph": [_I, _I, c_int32_p, _I, c_int32_p, c_int32_p],
    "gvbd_color_tet_mesh": [_I, _I, c_int32_p, _I, c_int32_p, c_int32_p],
    "gvbd_cloth_mesh_constants": [_I, c_float_p, _I, c_int32_p, ctypes.c_float, c_float_p, c_float_p, c_float_p,
                                  c_int32_p, c_int32_p, c_float_p],
    "gvbd_color_tri_mesh": [_I, _I, c_int32_p, _I, _I, c_int32_p, c_int32_p],
}
_RESTYPES = {"gvbd_last_error": ctypes.c_char_p}

_lib = None


class VBDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gaia_vbd error {code}: {msg}")
        self.code = code


def header_functions():
    """Names of every function declared in include/gaia_vbd.h."""
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:GVBD_API\s+)?(?:int|const char\*)\s+(gvbd_\w+)\s*\(", text, re.M)))


def lib():
    """Loads libgaia_vbd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def check(rc, ctx=None):
    if rc != GVBD_OK:
        msg = lib().gvbd_last_error(ctx)
        raise VBDError(rc, msg.decode() if msg else "")
    return rc


def fptr(a):
    return a.ctypes.data_as(c_float_p)


def iptr(a):
    return a.ctypes.data_as(c_int32_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
