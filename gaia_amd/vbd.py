"""Python host interface over the C ABI, mirroring the reference's VBD framework API.

`VBDPhysics` follows GAIA::VBDPhysics (Simulator/Modules/VBD/VBDPhysics.h:36-254):
loadRunningparameters -> initialize -> runStep, with the per-iteration units
(applyInitialStep, solveParallelGroup, updateVelocities) and
recordVBDSolveGraph. Every call goes through libgaia_vbd.so; the device work
runs in hand-written HIP kernels.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import (MODE_GROUP_LAUNCHES, MODE_WORK_QUEUE, ClothMeshDesc, ContactStats, PhysicsParams, Schedule,
                   ServedStep, ServerStats, TetMeshDesc, check, fptr, f32, i32, iptr, lib)

# Reference defaults (VBD/VBDPhysicsParameters.h:8-81, Parameters/PhysicsParameters.h:18-120).
DEFAULT_PHYSICS = dict(
    gravity=(0.0, -10.0, 0.0),
    time_step=1.0 / 60.0,
    num_substeps=20,
    iterations=1,
    world_lower=(-1e7, -1e7, -1e7),
    world_upper=(1e7, 1e7, 1e7),
    use_plane_ground=1,
    boundary_collision_stiffness=1e5,
    boundary_friction_dynamic=0.1,
    boundary_friction_eps_v=1.0,
    step_size=1.0,
    step_size_gd=1e-3,
    friction_start_iter=0,
    associate_gravity_with_inertia=1,
    # cloth (VBDCloth/VBDClothPhysicsParameters.h:15-81); handle_collision defaults to 0 here (the
    # reference's JSON default is 1, which gvbd_scene_load_json keeps)
    cloth_boundary_contact_stiffness=1e4,
    cloth_boundary_friction_eps=1e-2,
    degenerate_triangle_thres=1e-6,
    cloth_step_size=1.0,
    handle_collision=0,
    contact_stiffness=1e4,
    contact_radius=1.0,
    thickness=0.0,
    conservative_step_relaxation=0.45,
    contact_energy_type=1,
    contact_query_radius=0.4,
    contact_bvh_rebuild_steps=32,
    # Chebyshev acceleration of the tet iterations (VBD/VBDPhysicsParameters.h:64-65)
    use_accelerator=0,
    accelerator_rho=0.93,
    # tet meshes with handle_collision: 0 cloth-model contact on their boundary surface, 1 the
    # reference tet solver's penetration contact (CollisionDetertionParameters.h:43-80,
    # VBD/VBDPhysicsParameters.h:36,70; PhysicsParameters.h:81)
    tet_contact_model=0,
    allow_dcd=1,
    allow_ccd=1,
    handle_self_collision=1,
    collision_stiffness=1e5,
    collision_off_height=1e7,
    check_feasible_region=1,
    check_tet_traverse=1,
    tet_traverse_non_self=1,
    shift_query_point_to_center=1,
    stop_traversing_after_passing=1,
    feasible_region_epsilon=1e-2,
    ray_tri_intersection_epsilon=1e-10,
    max_search_distance_multiplier=1.8,
    center_shift_level=0.01,
    dcd_bvh_rebuild_steps=16,
)

# VBDObjectParamsTriMeshStVk defaults (VBDCloth/VBDTriMeshStVK.h:17-20, VBDCloth/VBDBaseTriMesh.h:12-16,
# TriMesh/TriMesh.h:33-36).
DEFAULT_CLOTH_MATERIAL = dict(
    density=1.0,
    mu=1e4,
    lmbd=1e4,
    bending_stiffness=0.0,
    damping_stvk=0.0,
    damping_bending=0.0,
    exponential_vel_damping=1.0,
    constant_vel_damping=0.0,
    initialization_type=2,
    friction_dynamic=0.1,
    friction_eps_v=0.1,
    # no-gravity zone (Materials/Materials.h:34-35): speed capped instead of damped while the
    # velocity's y exceeds the threshold (needs max_velocity_magnitude > 0)
    has_no_grav_zone=0,
    no_grav_zone_threshold=0.0,
    max_velocity_magnitude=-1.0,
)

# ObjectParametersVBDNeoHookean defaults (VBD/VBD_NeoHookean.h:25-29, VBD/VBD_BaseMaterial.h:29-45,
# Materials/Materials.h:32).
DEFAULT_TET_MATERIAL = dict(
    density=1.0,
    mu=1e4,
    lmbd=1e4,
    damping_hydrostatic=0.0,
    damping_deviatoric=0.0,
    exponential_vel_damping=1.0,
    constant_vel_damping=0.0,
    initialization_type=0,
    init_ratio_g=0.5,
    # boundary-surface contact friction in scenes with collision handling (as TriMesh.h:41-42)
    friction_dynamic=0.1,
    friction_eps_v=0.1,
    # no-gravity zone (Materials/Materials.h:34-35): speed capped instead of damped while the
    # position's y exceeds the threshold (needs max_velocity_magnitude > 0)
    has_no_grav_zone=0,
    no_grav_zone_threshold=0.0,
    max_velocity_magnitude=-1.0,
)


def physics_params(**overrides):
    d = dict(DEFAULT_PHYSICS)
    unknown = set(overrides) - set(d)
    if unknown:
        raise KeyError(f"unknown physics parameters {sorted(unknown)}")
    d.update(overrides)
    p = PhysicsParams()
    for k, v in d.items():
        if isinstance(v, (tuple, list)):
            getattr(p, k)[:] = [float(x) for x in v]
        else:
            setattr(p, k, v)
    return p


def physics_dict(**overrides):
    d = dict(DEFAULT_PHYSICS)
    d.update(overrides)
    return d


class Context:
    """One device context (one HIP stream)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(lib().gvbd_context_create(device, ctypes.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("context closed")
        return self._h

    def close(self):
        if self._h:
            lib().gvbd_context_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def check(self, rc):
        return check(rc, self._h)

    def synchronize(self):
        self.check(lib().gvbd_synchronize(self.handle))

    def memory_bytes(self):
        b = ctypes.c_int64()
        self.check(lib().gvbd_context_memory(self.handle, ctypes.byref(b)))
        return b.value

    def stream_copy_ms(self, nbytes, reps=5, method=0):
        """Average time of a device copy of nbytes (measured copy bandwidth): method 0 the
        library's streaming kernel, 1 the runtime's device-to-device hipMemcpyAsync."""
        ms = ctypes.c_float()
        self.check(lib().gvbd_copy_bandwidth(self.handle, int(nbytes), reps, int(method), ctypes.byref(ms)))
        return ms.value

    def event_record(self, slot):
        self.check(lib().gvbd_event_record(self.handle, slot))

    def event_elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        self.check(lib().gvbd_event_elapsed_ms(self.handle, a, b, ctypes.byref(ms)))
        return ms.value


def state_write(json_path, positions, velocities=None, frame_id=0, cur_time=0.0, binary=False):
    """Recovery-state file from per-mesh (N, 3) arrays (no device)."""
    counts = i32([len(np.asarray(p).reshape(-1, 3)) for p in positions])
    pos = f32(np.concatenate([np.asarray(p, np.float32).reshape(-1) for p in positions])) if len(positions) \
        else np.zeros(0, np.float32)
    vel = None
    if velocities is not None:
        vel = f32(np.concatenate([np.asarray(v, np.float32).reshape(-1) for v in velocities])) if len(velocities) \
            else np.zeros(0, np.float32)
    check(lib().gvbd_state_write(os.fsencode(json_path), int(frame_id), float(cur_time), 1 if binary else 0,
                                 len(counts), iptr(counts), fptr(pos), fptr(vel) if vel is not None else None))


def state_read(json_path):
    """Reads a recovery-state file: (frameId, curTime, [positions (N, 3)], [velocities (N, 3)])."""
    nm = ctypes.c_int32()
    check(lib().gvbd_state_read(os.fsencode(json_path), 0, ctypes.byref(nm), None, 0, None, None, None, None))
    counts = np.zeros(max(nm.value, 1), np.int32)
    check(lib().gvbd_state_read(os.fsencode(json_path), nm.value, ctypes.byref(nm), iptr(counts), 0, None, None,
                                None, None))
    total = 3 * int(counts[:nm.value].sum())
    pos = np.zeros(max(total, 1), np.float32)
    vel = np.zeros(max(total, 1), np.float32)
    fid = ctypes.c_int32()
    t = ctypes.c_double()
    check(lib().gvbd_state_read(os.fsencode(json_path), nm.value, ctypes.byref(nm), iptr(counts), total, fptr(pos),
                                fptr(vel), ctypes.byref(fid), ctypes.byref(t)))
    splits = np.cumsum(3 * counts[:nm.value])[:-1]
    return (fid.value, t.value, [p.reshape(-1, 3) for p in np.split(pos[:total], splits)],
            [v.reshape(-1, 3) for v in np.split(vel[:total], splits)])


def write_ply(path, positions, faces):
    """ASCII PLY of one surface (GAIA::saveAsPly's format)."""
    pos = f32(positions).reshape(-1)
    fc = i32(faces).reshape(-1)
    check(lib().gvbd_write_ply(os.fsencode(path), len(pos) // 3, fptr(pos), len(fc) // 3, iptr(fc)))


def color_tet_mesh(num_vertices, tets, algorithm="mcs"):
    """Vertex colouring of a tet mesh's vertex graph (GraphColoring mcs/greedy)."""
    tets = i32(tets).reshape(-1)
    colors = np.zeros(num_vertices, dtype=np.int32)
    nc = ctypes.c_int32()
    alg = _lib.GVBD_COLOR_GREEDY if algorithm == "greedy" else _lib.GVBD_COLOR_MCS
    check(lib().gvbd_color_tet_mesh(num_vertices, tets.size // 4, iptr(tets), alg, iptr(colors), ctypes.byref(nc)))
    return colors


def tet_mesh_constants(positions, tets, density):
    """Rest-pose DmInv (T, 9 column-major), signed volume (T,) and lumped mass (N,)."""
    pos = f32(positions).reshape(-1)
    tets = i32(tets).reshape(-1)
    nv, nt = pos.size // 3, tets.size // 4
    dminv = np.zeros(9 * nt, np.float32)
    vol = np.zeros(nt, np.float32)
    mass = np.zeros(nv, np.float32)
    check(lib().gvbd_tet_mesh_constants(nv, fptr(pos), nt, iptr(tets), float(density), fptr(dminv), fptr(vol),
                                        fptr(mass)))
    return dminv.reshape(nt, 9), vol, mass


def cloth_mesh_constants(positions, faces, density):
    """Rest-pose DmInv (F, 4 col-major 2x2), area (F,), mass (N,), edges (E, 6) and Q (E, 4, 4)."""
    pos = f32(positions).reshape(-1)
    faces = i32(faces).reshape(-1)
    nv, nf = pos.size // 3, faces.size // 3
    ne = ctypes.c_int32()
    check(lib().gvbd_cloth_mesh_constants(nv, fptr(pos), nf, iptr(faces), float(density), None, None, None,
                                          ctypes.byref(ne), None, None))
    dminv = np.zeros(4 * nf, np.float32)
    area = np.zeros(nf, np.float32)
    mass = np.zeros(nv, np.float32)
    edges = np.zeros(6 * ne.value, np.int32)
    Q = np.zeros(16 * ne.value, np.float32)
    check(lib().gvbd_cloth_mesh_constants(nv, fptr(pos), nf, iptr(faces), float(density), fptr(dminv), fptr(area),
                                          fptr(mass), ctypes.byref(ne), iptr(edges), fptr(Q)))
    return dminv.reshape(nf, 4), area, mass, edges.reshape(-1, 6), Q.reshape(-1, 4, 4)


def color_tri_mesh(num_vertices, faces, algorithm="mcs", with_bending=True):
    """Vertex colouring of a triangle mesh's vertex graph (optionally with bending cross edges)."""
    faces = i32(faces).reshape(-1)
    colors = np.zeros(num_vertices, dtype=np.int32)
    nc = ctypes.c_int32()
    alg = _lib.GVBD_COLOR_GREEDY if algorithm == "greedy" else _lib.GVBD_COLOR_MCS
    check(lib().gvbd_color_tri_mesh(num_vertices, faces.size // 3, iptr(faces), 1 if with_bending else 0, alg,
                                    iptr(colors), ctypes.byref(nc)))
    return colors


def color_graph(num_vertices, edges, algorithm="mcs"):
    edges = i32(edges).reshape(-1)
    colors = np.zeros(num_vertices, dtype=np.int32)
    nc = ctypes.c_int32()
    alg = _lib.GVBD_COLOR_GREEDY if algorithm == "greedy" else _lib.GVBD_COLOR_MCS
    check(lib().gvbd_color_graph(num_vertices, edges.size // 2, iptr(edges), alg, iptr(colors), ctypes.byref(nc)))
    return colors


class VBDPhysics:
    """One scene on one context (GAIA::VBDPhysics counterpart)."""

    def __init__(self, ctx=None, device=0):
        self.ctx = ctx if ctx is not None else Context(device)
        self.scene = None
        self._pending = []
        self._deformers = []
        self._attachments = []
        self._params = None
        self._json = None

    # ---- set-up (BasePhysicFramework::loadRunningparameters / initialize) ----
    def loadRunningparameters(self, models_json, params_json, out_folder=None, repo_root=None):
        self._json = (os.fspath(models_json), os.fspath(params_json), repo_root)
        self.outFolder = out_folder

    def setPhysicsParams(self, **overrides):
        self._params = physics_params(**overrides)

    def addTetMesh(self, positions, tets, fixed_points=(), vertex_colors=None, coloring="mcs", velocities=None,
                   initial_velocity=(0.0, 0.0, 0.0), **material):
        self._pending.append(("tet", positions, tets, fixed_points, vertex_colors, coloring, velocities,
                              initial_velocity, material))

    def addClothMesh(self, positions, faces, fixed_points=(), vertex_colors=None, coloring="mcs", velocities=None,
                     initial_velocity=(0.0, 0.0, 0.0), **material):
        """StVK + bending triangle mesh (VBDClothSimulationFramework mesh)."""
        self._pending.append(("cloth", positions, faces, fixed_points, vertex_colors, coloring, velocities,
                              initial_velocity, material))

    def addDeformer(self, deformer):
        """A deformer given as the reference's JSON object (loadDeformers, VBD/VBDDeformer.h:181-222):
        DeformerName "Rotator" (center, axis, angularVelocity, rotationEndTime) or "Translator"
        (speed, deformationEndTime, deformationStartTime), selectedMeshes, selectedVertices. Its
        vertices become fixed points moved along the prescribed path."""
        self._deformers.append(dict(deformer))

    def addAttachments(self, params):
        """Vertex-face attachments given as the reference's VertexFaceAttachmentConstraintParams
        (vertexConstraintVertexInfos, vertexConstraintTargetFacesInfos, vertexConstraintBarys,
        vertexConstraintStiffness; VBDCloth/VBDClothPhysics.cpp:148-224)."""
        self._attachments.append(dict(params))

    @property
    def curTime(self):
        """The scene time (float, + dt per substep)."""
        t = ctypes.c_float()
        self.ctx.check(lib().gvbd_scene_get_time(self.ctx.handle, self.scene, ctypes.byref(t)))
        return t.value

    @curTime.setter
    def curTime(self, t):
        self.ctx.check(lib().gvbd_scene_set_time(self.ctx.handle, self.scene, float(t)))

    def _add_deformer(self, d):
        name = d.get("DeformerName")
        if name not in ("Rotator", "Translator"):
            raise ValueError(f"unrecognised deformer {name!r} (Rotator, Translator)")
        desc = _lib.DeformerDesc()
        rot = name == "Rotator"
        desc.kind = 0 if rot else 1
        desc.center[:] = [float(c) for c in d.get("center", (0, 0, 0))]
        desc.axis[:] = [float(c) for c in d.get("axis", (0, 0, 0))]
        desc.speed[:] = [float(c) for c in d.get("speed", (0, 0, 0))]
        desc.start_time = float(d.get("deformationStartTime", 0.0))
        desc.angular_velocity = float(d.get("angularVelocity", 0.0))
        desc.end_time = float(d.get("rotationEndTime" if rot else "deformationEndTime", -1.0))
        meshes = i32(np.asarray(d["selectedMeshes"], dtype=np.int32).reshape(-1))
        sel = [np.asarray(v, dtype=np.int32).reshape(-1) for v in d["selectedVertices"]]
        if len(sel) != len(meshes):
            raise ValueError("selectedVertices needs one list per selected mesh")
        counts = i32(np.array([len(v) for v in sel], dtype=np.int32))
        verts = i32(np.concatenate(sel) if sel else np.zeros(0, np.int32))
        desc.num_meshes = len(meshes)
        desc.meshes = iptr(meshes)
        desc.num_vertices = iptr(counts)
        desc.vertices = iptr(verts)
        self.ctx.check(lib().gvbd_scene_add_deformer(self.ctx.handle, self.scene, ctypes.byref(desc)))

    def _add_attachments(self, params):
        vi = np.asarray(params["vertexConstraintVertexInfos"], dtype=np.int32).reshape(-1, 2)
        tf = np.asarray(params["vertexConstraintTargetFacesInfos"], dtype=np.int32).reshape(-1, 4)
        bary = np.asarray(params["vertexConstraintBarys"], dtype=np.float64).reshape(-1, 3)
        k = np.asarray(params.get("vertexConstraintStiffness", []), dtype=np.float64).reshape(-1)
        n = len(vi)
        if len(tf) != n or len(bary) != n or (len(k) and len(k) != n):
            raise ValueError("attachment arrays disagree in length")
        arr = (_lib.AttachmentDesc * max(n, 1))()
        for i in range(n):
            a = arr[i]
            a.mesh, a.vertex = int(vi[i, 0]), int(vi[i, 1])
            a.target_mesh = int(tf[i, 0])
            a.face[:] = [int(c) for c in tf[i, 1:]]
            a.bary[:] = [float(c) for c in bary[i]]
            a.stiffness = float(k[i]) if len(k) else 1e4
        self.ctx.check(lib().gvbd_scene_add_attachments(self.ctx.handle, self.scene, arr, n))

    def initialize(self, defer_upload=False):
        """Creates the scene and uploads it. With defer_upload the scene is only created with
        its meshes (host side): a Server uploads it when its first step is requested."""
        h = self.ctx.handle
        sid = ctypes.c_int32()
        if self._json is not None:
            models, params, root = self._json
            self.ctx.check(lib().gvbd_scene_load_json(h, models.encode(), params.encode(),
                                                      (root or "").encode(), ctypes.byref(sid)))
            self.scene = sid.value
            return
        params = self._params if self._params is not None else physics_params()
        self.ctx.check(lib().gvbd_scene_create(h, ctypes.byref(params), ctypes.byref(sid)))
        self.scene = sid.value
        try:
            for (kind, pos, elems, fixed, colors, coloring, vel, v0, material) in self._pending:
                if kind == "tet":
                    self._add_tet(pos, elems, fixed, colors, coloring, vel, v0, material)
                else:
                    self._add_cloth(pos, elems, fixed, colors, coloring, vel, v0, material)
            for d in self._deformers:
                self._add_deformer(d)
            for a in self._attachments:
                self._add_attachments(a)
            if not defer_upload:
                self.ctx.check(lib().gvbd_scene_initialize(h, self.scene))
        except Exception:
            lib().gvbd_scene_destroy(h, self.scene)
            self.scene = None
            raise
        self._pending = []

    def _add_tet(self, pos, tets, fixed, colors, coloring, vel, v0, material):
        mat = dict(DEFAULT_TET_MATERIAL)
        unknown = set(material) - set(mat)
        if unknown:
            raise KeyError(f"unknown material parameters {sorted(unknown)}")
        mat.update(material)
        pos = f32(pos).reshape(-1)
        tets = i32(tets).reshape(-1)
        fixed = i32(np.asarray(fixed, dtype=np.int32).reshape(-1))
        d = TetMeshDesc()
        d.num_vertices = pos.size // 3
        d.positions = fptr(pos)
        keep = [pos, tets, fixed]
        if vel is not None:
            vel = f32(vel).reshape(-1)
            keep.append(vel)
            d.velocities = fptr(vel)
        d.initial_velocity[:] = [float(x) for x in v0]
        d.num_tets = tets.size // 4
        d.tets = iptr(tets)
        for k in ("density", "mu", "lmbd", "damping_hydrostatic", "damping_deviatoric", "exponential_vel_damping",
                  "constant_vel_damping", "init_ratio_g", "friction_dynamic", "friction_eps_v",
                  "no_grav_zone_threshold", "max_velocity_magnitude"):
            setattr(d, k, float(mat[k]))
        d.initialization_type = int(mat["initialization_type"])
        d.has_no_grav_zone = int(bool(mat["has_no_grav_zone"]))
        d.num_fixed = fixed.size
        d.fixed_points = iptr(fixed) if fixed.size else None
        if colors is not None:
            colors = i32(colors).reshape(-1)
            keep.append(colors)
            d.vertex_colors = iptr(colors)
        d.coloring_algorithm = _lib.GVBD_COLOR_GREEDY if coloring == "greedy" else _lib.GVBD_COLOR_MCS
        mid = ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_add_tet_mesh(self.ctx.handle, self.scene, ctypes.byref(d), ctypes.byref(mid)))
        del keep
        return mid.value

    def _add_cloth(self, pos, faces, fixed, colors, coloring, vel, v0, material):
        mat = dict(DEFAULT_CLOTH_MATERIAL)
        unknown = set(material) - set(mat)
        if unknown:
            raise KeyError(f"unknown cloth material parameters {sorted(unknown)}")
        mat.update(material)
        pos = f32(pos).reshape(-1)
        faces = i32(faces).reshape(-1)
        fixed = i32(np.asarray(fixed, dtype=np.int32).reshape(-1))
        d = ClothMeshDesc()
        d.num_vertices = pos.size // 3
        d.positions = fptr(pos)
        keep = [pos, faces, fixed]
        if vel is not None:
            vel = f32(vel).reshape(-1)
            keep.append(vel)
            d.velocities = fptr(vel)
        d.initial_velocity[:] = [float(x) for x in v0]
        d.num_faces = faces.size // 3
        d.faces = iptr(faces)
        for k in ("density", "mu", "lmbd", "bending_stiffness", "damping_stvk", "damping_bending",
                  "exponential_vel_damping", "constant_vel_damping", "friction_dynamic", "friction_eps_v",
                  "no_grav_zone_threshold", "max_velocity_magnitude"):
            setattr(d, k, float(mat[k]))
        d.initialization_type = int(mat["initialization_type"])
        d.has_no_grav_zone = int(bool(mat["has_no_grav_zone"]))
        d.num_fixed = fixed.size
        d.fixed_points = iptr(fixed) if fixed.size else None
        if colors is not None:
            colors = i32(colors).reshape(-1)
            keep.append(colors)
            d.vertex_colors = iptr(colors)
        d.coloring_algorithm = _lib.GVBD_COLOR_GREEDY if coloring == "greedy" else _lib.GVBD_COLOR_MCS
        mid = ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_add_cloth_mesh(self.ctx.handle, self.scene, ctypes.byref(d),
                                                       ctypes.byref(mid)))
        del keep
        return mid.value

    # ---- stepping (VBDPhysics::runStep and its per-iteration units) ----
    def runStep(self):
        self.ctx.check(lib().gvbd_scene_run_step(self.ctx.handle, self.scene))

    def runSteps(self, n):
        self.ctx.check(lib().gvbd_scene_run_steps(self.ctx.handle, self.scene, n))

    def applyInitialStep(self):
        self.ctx.check(lib().gvbd_scene_apply_initial_step(self.ctx.handle, self.scene))

    def solveParallelGroup(self, group, iIter=0):
        self.ctx.check(lib().gvbd_scene_solve_parallel_group(self.ctx.handle, self.scene, group, iIter))

    def updateVelocities(self):
        self.ctx.check(lib().gvbd_scene_update_velocities(self.ctx.handle, self.scene))

    def applyCollisionDetection(self):
        """VBDClothSimulationFramework::applyCollisionDetection (cloth scenes with handle_collision)."""
        self.ctx.check(lib().gvbd_scene_collision_detection(self.ctx.handle, self.scene))

    def contactStats(self):
        """Counters of a scene with handle_collision: frames, iterations, detection passes, BVH
        rebuilds, device time of detection (ms), and the candidate counts of the last detection."""
        st = ContactStats()
        self.ctx.check(lib().gvbd_scene_contact_stats(self.ctx.handle, self.scene, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in ContactStats._fields_}

    def deviceBytes(self):
        """Device bytes of the scene as a Server's budget counts them: {"arena": ..., "detector": ...}
        (the arena is 0 while the scene is not uploaded or is copied out)."""
        a = ctypes.c_int64()
        d = ctypes.c_int64()
        self.ctx.check(lib().gvbd_scene_device_bytes(self.ctx.handle, self.scene, ctypes.byref(a), ctypes.byref(d)))
        return {"arena": a.value, "detector": d.value}

    def penetrationResults(self):
        """The last detection of a scene with the penetration contact model (tet_contact_model 1), one
        row per scene vertex in mesh order: (mesh, vertex, hit mesh, hit face, contained, from CCD),
        hit face -1 when none (the rows of VBDPhysics::collisionResultsAll, VBD/VBDPhysics.h:154), and
        the counts (DCD hits, contained vertices without a face, CCD hits, walk stack overflows, closest-face
        searches past the nearest 8 faces)."""
        sizes = [self.meshInfo(m)[0] for m in range(self.numMeshes())]
        nv = sum(sizes)
        hm = np.zeros(nv, np.int32)
        hf = np.zeros(nv, np.int32)
        fl = np.zeros(nv, np.uint8)
        counts = np.zeros(5, np.int32)
        self.ctx.check(lib().gvbd_scene_penetration_results(
            self.ctx.handle, self.scene, iptr(hm), iptr(hf), fl.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
            iptr(counts)))
        mesh = np.concatenate([np.full(n, m, np.int32) for m, n in enumerate(sizes)])
        vert = np.concatenate([np.arange(n, dtype=np.int32) for n in sizes])
        rows = np.stack([mesh, vert, hm, hf, (fl & 1).astype(np.int32), (fl >> 1 & 1).astype(np.int32)], axis=1)
        return rows, counts

    def solveFailures(self, reset=False):
        """Vertex solves whose 3x3 block had no solution (gradient step for tet vertices, no move for
        cloth vertices) since initialisation or the last reset."""
        n = ctypes.c_int64()
        self.ctx.check(lib().gvbd_scene_solve_failures(self.ctx.handle, self.scene, 1 if reset else 0,
                                                       ctypes.byref(n)))
        return n.value

    def evaluateConvergence(self):
        """VBDPhysics::evaluateConvergenceGPU: merit energy, its inertia and elastic parts, and the
        average force norm over the tet vertices, at the current positions."""
        out = [ctypes.c_double() for _ in range(4)]
        self.ctx.check(lib().gvbd_scene_evaluate_convergence(self.ctx.handle, self.scene, *[ctypes.byref(o) for o in out]))
        return dict(zip(("energy", "inertia_energy", "elastic_energy", "avg_force_norm"), (o.value for o in out)))

    def evaluateConvergenceNoAlloc(self):
        """The same evaluation without device allocation (the library's fixed per-GPU scratch);
        evaluateConvergence must have run once on this scene to make its energy records resident."""
        out = [ctypes.c_double() for _ in range(4)]
        self.ctx.check(lib().gvbd_scene_evaluate_convergence_noalloc(self.ctx.handle, self.scene,
                                                                     *[ctypes.byref(o) for o in out]))
        return dict(zip(("energy", "inertia_energy", "elastic_energy", "avg_force_norm"), (o.value for o in out)))

    def recordVBDSolveGraph(self, enable=True):
        self.ctx.check(lib().gvbd_scene_record_graph(self.ctx.handle, self.scene, 1 if enable else 0))

    def setSchedule(self, lanes_per_vertex, block, mode=None, vertices_per_entry=0, blocks_per_cu=0, spin_limit=0,
                    stage_records=0, xcd_blocks=0):
        """Vertex-solve schedule; `mode` "group" (one launch per parallel group) or "queue" (one
        persistent work-queue launch per iteration); None keeps the current mode. stage_records
        (group launches): incidence records staged through LDS per chunk (0: none); xcd_blocks
        (group launches): each XCD's blocks take one contiguous range of the work items."""
        if mode is None and not stage_records and not xcd_blocks:
            self.ctx.check(lib().gvbd_scene_set_schedule(self.ctx.handle, self.scene, lanes_per_vertex, block))
            return
        sc = make_schedule(lanes_per_vertex, block, mode or "group", vertices_per_entry, blocks_per_cu, spin_limit,
                           stage_records, xcd_blocks)
        self.ctx.check(lib().gvbd_scene_set_schedule_ex(self.ctx.handle, self.scene, ctypes.byref(sc)))

    def getSchedule(self):
        sc = Schedule()
        self.ctx.check(lib().gvbd_scene_get_schedule(self.ctx.handle, self.scene, ctypes.byref(sc)))
        return schedule_dict(sc)

    def runIteration(self, iIter):
        """One VBD iteration (every parallel group in order) after applyInitialStep."""
        self.ctx.check(lib().gvbd_scene_run_iteration(self.ctx.handle, self.scene, iIter))

    def synchronize(self):
        self.ctx.synchronize()

    # ---- queries / state ----
    def numMeshes(self):
        n = ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_num_meshes(self.ctx.handle, self.scene, ctypes.byref(n)))
        return n.value

    def meshInfo(self, mesh):
        nv, nt = ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_mesh_info(self.ctx.handle, self.scene, mesh, ctypes.byref(nv), ctypes.byref(nt)))
        return nv.value, nt.value

    def parallelGroups(self):
        n = ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_num_parallel_groups(self.ctx.handle, self.scene, ctypes.byref(n)))
        out = []
        for g in range(n.value):
            cnt = ctypes.c_int32()
            self.ctx.check(lib().gvbd_scene_parallel_group(self.ctx.handle, self.scene, g, None, ctypes.byref(cnt)))
            buf = np.zeros(2 * cnt.value, dtype=np.int32)
            self.ctx.check(lib().gvbd_scene_parallel_group(self.ctx.handle, self.scene, g, iptr(buf), ctypes.byref(cnt)))
            out.append(buf.reshape(-1, 2))
        return out

    def meshConstants(self, mesh):
        nv, nt = self.meshInfo(mesh)
        dminv = np.zeros(9 * nt, dtype=np.float32)
        vol = np.zeros(nt, dtype=np.float32)
        mass = np.zeros(nv, dtype=np.float32)
        self.ctx.check(lib().gvbd_scene_mesh_constants(self.ctx.handle, self.scene, mesh, fptr(dminv), fptr(vol),
                                                       fptr(mass)))
        return dminv.reshape(nt, 9), vol, mass

    def getState(self, mesh=0):
        nv, _ = self.meshInfo(mesh)
        x = np.zeros(3 * nv, dtype=np.float32)
        v = np.zeros(3 * nv, dtype=np.float32)
        self.ctx.check(lib().gvbd_scene_get_mesh_state(self.ctx.handle, self.scene, mesh, fptr(x), fptr(v)))
        return x.reshape(nv, 3), v.reshape(nv, 3)

    def setState(self, mesh, positions=None, velocities=None):
        x = f32(positions).reshape(-1) if positions is not None else None
        v = f32(velocities).reshape(-1) if velocities is not None else None
        self.ctx.check(lib().gvbd_scene_set_mesh_state(self.ctx.handle, self.scene, mesh,
                                                       fptr(x) if x is not None else None,
                                                       fptr(v) if v is not None else None))

    # ---- scene files (BasePhysicsFramework.h:145-319, BasePhysicsFramework.cpp:264-281,441-503) ----
    def saveState(self, json_path, frame_id=0, cur_time=None, binary=False):
        """PhysicsState of every mesh (positions and velocities) as a recovery-state JSON; binary
        writes <stem>.pos / <stem>.vel beside it. cur_time defaults to the scene time."""
        if cur_time is None:
            cur_time = self.curTime
        self.ctx.check(lib().gvbd_scene_save_state(self.ctx.handle, self.scene, os.fsencode(json_path), int(frame_id),
                                                   float(cur_time), 1 if binary else 0))

    def recoverFromState(self, state_file):
        """BasePhysicFramework::recoverFromState: loads a recovery state (text or binary) into the
        meshes; returns (frameId, curTime)."""
        fid = ctypes.c_int32()
        t = ctypes.c_double()
        self.ctx.check(lib().gvbd_scene_load_state(self.ctx.handle, self.scene, os.fsencode(state_file),
                                                   ctypes.byref(fid), ctypes.byref(t)))
        self.frameId, self.curTime = fid.value, t.value
        return fid.value, t.value

    def writePLY(self, path, all_in_one=True):
        """Surfaces of all meshes as ASCII PLY (one file, or <stem>_Model_NNNNNN<ext> per mesh)."""
        self.ctx.check(lib().gvbd_scene_write_ply(self.ctx.handle, self.scene, os.fsencode(path), 1 if all_in_one else 0))

    def writeBinaryFrame(self, path):
        """Float32 surface positions of all meshes back to back (the reference's .bin frame)."""
        self.ctx.check(lib().gvbd_scene_write_binary_frame(self.ctx.handle, self.scene, os.fsencode(path)))

    def writeOutputs(self, out_folder, frame_id, output_ext="ply", save_all_models_together=True,
                     output_recovery_state=False, recovery_state_step=1, recovery_state_binary=False,
                     binary_visualization_steps=10, cur_time=0.0):
        """BasePhysicFramework::writeOutputs (BasePhysicsFramework.cpp:441-503): A<frame, 8 digits>.<ext>
        in out_folder ("ply", or "bin" with TemplateMesh/TemplateMesh.ply at frame 0 and a PLY every
        binary_visualization_steps frames), and RecoveryStates/A<frame>.json every
        recovery_state_step frames when asked."""
        num = f"{int(frame_id):08d}"
        os.makedirs(out_folder, exist_ok=True)
        if output_ext == "ply":
            self.writePLY(os.path.join(out_folder, f"A{num}.ply"), save_all_models_together)
        elif output_ext == "bin":
            if frame_id == 0:
                tdir = os.path.join(out_folder, "TemplateMesh")
                os.makedirs(tdir, exist_ok=True)
                self.writePLY(os.path.join(tdir, "TemplateMesh.ply"), all_in_one=False)
            self.writeBinaryFrame(os.path.join(out_folder, f"A{num}.bin"))
            if frame_id % binary_visualization_steps == 0:
                self.writePLY(os.path.join(out_folder, f"A{num}.ply"), save_all_models_together)
        else:
            raise ValueError(f"unsupported output extension {output_ext!r}")
        if output_recovery_state and frame_id % recovery_state_step == 0:
            sdir = os.path.join(out_folder, "RecoveryStates")
            os.makedirs(sdir, exist_ok=True)
            self.saveState(os.path.join(sdir, f"A{num}.json"), frame_id, cur_time, recovery_state_binary)

    def clone(self, defer_upload=True):
        """A new scene with this one's parameters and meshes as added (host preparation copied);
        with defer_upload it is not uploaded yet (see initialize)."""
        out = VBDPhysics(self.ctx)
        sid = ctypes.c_int32()
        self.ctx.check(lib().gvbd_scene_clone(self.ctx.handle, self.scene, ctypes.byref(sid)))
        out.scene = sid.value
        if not defer_upload:
            self.ctx.check(lib().gvbd_scene_initialize(self.ctx.handle, out.scene))
        return out

    def destroy(self):
        if self.scene is not None and self.ctx._h:
            self.ctx.check(lib().gvbd_scene_destroy(self.ctx.handle, self.scene))
        self.scene = None


_MODES = {"group": MODE_GROUP_LAUNCHES, "queue": MODE_WORK_QUEUE}


def make_schedule(lanes_per_vertex, block, mode="group", vertices_per_entry=0, blocks_per_cu=0, spin_limit=0,
                  stage_records=0, xcd_blocks=0):
    if mode not in _MODES:
        raise ValueError(f"mode must be one of {sorted(_MODES)}")
    return Schedule(lanes_per_vertex, block, _MODES[mode], vertices_per_entry, blocks_per_cu, spin_limit,
                    stage_records, int(xcd_blocks))


def schedule_dict(sc):
    mode = {v: k for k, v in _MODES.items()}[sc.mode]
    return dict(lanes_per_vertex=sc.lanes_per_vertex, block=sc.block, mode=mode,
                vertices_per_entry=sc.vertices_per_entry, blocks_per_cu=sc.blocks_per_cu, spin_limit=sc.spin_limit,
                stage_records=sc.stage_records, xcd_blocks=sc.xcd_blocks)


class Batch:
    """Many initialised scenes advanced together: one launch per parallel group, or one
    persistent work-queue launch per iteration."""

    def __init__(self, ctx, scenes):
        self.ctx = ctx
        ids = i32([s.scene if isinstance(s, VBDPhysics) else int(s) for s in scenes])
        b = ctypes.c_int32()
        ctx.check(lib().gvbd_batch_create(ctx.handle, iptr(ids), ids.size, ctypes.byref(b)))
        self.batch = b.value

    def set_schedule(self, lanes_per_vertex, block, mode=None, vertices_per_entry=0, blocks_per_cu=0, spin_limit=0,
                     stage_records=0, xcd_blocks=0):
        if mode is None and not stage_records and not xcd_blocks:
            self.ctx.check(lib().gvbd_batch_set_schedule(self.ctx.handle, self.batch, lanes_per_vertex, block))
            return
        sc = make_schedule(lanes_per_vertex, block, mode or "group", vertices_per_entry, blocks_per_cu, spin_limit,
                           stage_records, xcd_blocks)
        self.ctx.check(lib().gvbd_batch_set_schedule_ex(self.ctx.handle, self.batch, ctypes.byref(sc)))

    def get_schedule(self):
        sc = Schedule()
        self.ctx.check(lib().gvbd_batch_get_schedule(self.ctx.handle, self.batch, ctypes.byref(sc)))
        return schedule_dict(sc)

    def queue_info(self):
        """(entries of one iteration's work queue, persistent grid blocks) of the current schedule."""
        e, g = ctypes.c_int32(), ctypes.c_int32()
        self.ctx.check(lib().gvbd_batch_queue_info(self.ctx.handle, self.batch, ctypes.byref(e), ctypes.byref(g)))
        return e.value, g.value

    def record_graph(self, enable=True):
        self.ctx.check(lib().gvbd_batch_record_graph(self.ctx.handle, self.batch, 1 if enable else 0))

    def run_steps(self, n):
        self.ctx.check(lib().gvbd_batch_run_steps(self.ctx.handle, self.batch, n))

    def apply_initial_step(self):
        self.ctx.check(lib().gvbd_batch_apply_initial_step(self.ctx.handle, self.batch))

    def run_iteration(self, iteration):
        self.ctx.check(lib().gvbd_batch_run_iteration(self.ctx.handle, self.batch, iteration))

    def update_velocities(self):
        self.ctx.check(lib().gvbd_batch_update_velocities(self.ctx.handle, self.batch))

    def solve_group(self, group, iteration=0):
        self.ctx.check(lib().gvbd_batch_solve_parallel_group(self.ctx.handle, self.batch, group, iteration))

    def num_groups(self):
        n = ctypes.c_int32()
        self.ctx.check(lib().gvbd_batch_num_parallel_groups(self.ctx.handle, self.batch, ctypes.byref(n)))
        return n.value

    def group_work(self, group):
        b, v = ctypes.c_int64(), ctypes.c_int64()
        self.ctx.check(lib().gvbd_batch_group_work(self.ctx.handle, self.batch, group, ctypes.byref(b), ctypes.byref(v)))
        return b.value, v.value

    def destroy(self):
        if self.batch is not None and self.ctx._h:
            lib().gvbd_batch_destroy(self.ctx.handle, self.batch)
        self.batch = None


SERVED_STEP_DTYPE = np.dtype([("tag", "<i8"), ("scene", "<i4"), ("frame", "<i4"), ("latency_ms", "<f8"),
                              ("queue_ms", "<f8")])


def _scene_id(s):
    return s.scene if isinstance(s, VBDPhysics) else int(s)


class Server:
    """Serving loop: step requests for many scenes, continuously batched into one frame launch
    sequence per tick, with scenes admitted to HBM under a byte budget (idle ones copied out to
    host memory, least recently stepped first)."""

    def __init__(self, ctx, budget_bytes):
        self.ctx = ctx
        sid = ctypes.c_int32()
        ctx.check(lib().gvbd_server_create(ctx.handle, int(budget_bytes), ctypes.byref(sid)))
        self.server = sid.value

    def set_schedule(self, lanes_per_vertex, block, mode="group", vertices_per_entry=0, blocks_per_cu=0,
                     spin_limit=0, stage_records=0, xcd_blocks=0):
        sc = make_schedule(lanes_per_vertex, block, mode, vertices_per_entry, blocks_per_cu, spin_limit,
                           stage_records, xcd_blocks)
        self.ctx.check(lib().gvbd_server_set_schedule(self.ctx.handle, self.server, ctypes.byref(sc)))

    def add_scene(self, scene):
        self.ctx.check(lib().gvbd_server_add_scene(self.ctx.handle, self.server, _scene_id(scene)))

    def remove_scene(self, scene):
        self.ctx.check(lib().gvbd_server_remove_scene(self.ctx.handle, self.server, _scene_id(scene)))

    def request_step(self, scene, tag=0):
        self.ctx.check(lib().gvbd_server_request_step(self.ctx.handle, self.server, _scene_id(scene), int(tag)))

    def tick(self):
        st = ServerStats()
        self.ctx.check(lib().gvbd_server_tick(self.ctx.handle, self.server, ctypes.byref(st)))
        return _stats_dict(st)

    def poll(self, capacity=4096):
        """Completed requests (structured array: tag, scene, frame, latency_ms, queue_ms)."""
        out = np.zeros(capacity, dtype=SERVED_STEP_DTYPE)
        n = ctypes.c_int32()
        self.ctx.check(lib().gvbd_server_poll(self.ctx.handle, self.server,
                                              out.ctypes.data_as(ctypes.POINTER(ServedStep)), capacity,
                                              ctypes.byref(n)))
        return out[: n.value]

    def stats(self):
        st = ServerStats()
        self.ctx.check(lib().gvbd_server_get_stats(self.ctx.handle, self.server, ctypes.byref(st)))
        return _stats_dict(st)

    def resident(self, scene):
        r = ctypes.c_int32()
        self.ctx.check(lib().gvbd_server_scene_resident(self.ctx.handle, self.server, _scene_id(scene),
                                                        ctypes.byref(r)))
        return bool(r.value)

    def destroy(self):
        if self.server is not None and self.ctx._h:
            self.ctx.check(lib().gvbd_server_destroy(self.ctx.handle, self.server))
        self.server = None


def _stats_dict(st):
    return {name: getattr(st, name) for name, _ in ServerStats._fields_}
