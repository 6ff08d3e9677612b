"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

# flat parameter indices (oracle.h)
OP = dict(GRAVITY_X=0, GRAVITY_Y=1, GRAVITY_Z=2, DT=3, NUM_SUBSTEPS=4, ITERATIONS=5, LOWER_X=6, LOWER_Y=7,
          LOWER_Z=8, UPPER_X=9, UPPER_Y=10, UPPER_Z=11, USE_PLANE_GROUND=12, BOUNDARY_STIFFNESS=13,
          BOUNDARY_FRICTION_MU=14, BOUNDARY_FRICTION_EPS_V=15, STEP_SIZE=16, STEP_SIZE_GD=17,
          FRICTION_START_ITER=18, ASSOCIATE_GRAVITY=19, INIT_RATIO_G=21, USE_ACCELERATOR=32, ACCELERATOR_RHO=33)
OP_COUNT = 34
OM = dict(DENSITY=0, MU=1, LAMBDA=2, DAMP_HYDRO=3, DAMP_DEV=4, EXP_VEL_DAMPING=5, CONST_VEL_DAMPING=6, INIT_TYPE=7,
          FRICTION_MU=8, FRICTION_EPS_V=9, NO_GRAV_ZONE=10, NO_GRAV_THRESHOLD=11, MAX_VELOCITY=12)
OM_COUNT = 13

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: build with `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_tet_prepare_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_double, _dp, _dp, _dp]
        L.oracle_tet_prepare_f32.argtypes = [ctypes.c_int, _fp, ctypes.c_int, _ip, ctypes.c_float, _fp, _fp, _fp]
        L.oracle_neohookean_vertex_f64.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_int, _dp, _dp]
        L.oracle_neohookean_vertex_f64.restype = None
        L.oracle_solve3x3_f64.argtypes = [_dp, _dp, _dp]
        L.oracle_solve3x3_f32.argtypes = [_fp, _fp, _fp]
        L.oracle_color_tet_mesh.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _ip]
        L.oracle_color_graph.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, _ip]
        L.oracle_tet_scene_create.argtypes = [ctypes.c_int, _dp]
        L.oracle_tet_scene_create.restype = ctypes.c_void_p
        L.oracle_tet_scene_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_tet_scene_destroy.restype = None
        L.oracle_tet_scene_add_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _dp,
                                                ctypes.c_int, _ip, _ip]
        L.oracle_tet_scene_finalize.argtypes = [ctypes.c_void_p]
        L.oracle_tet_scene_num_groups.argtypes = [ctypes.c_void_p]
        L.oracle_tet_scene_group.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_tet_scene_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_tet_scene_get_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_tet_scene_set_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_tet_scene_evaluate_convergence.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_tet_scene_add_deformer.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, _ip]
        L.oracle_tet_scene_time.argtypes = [ctypes.c_void_p]
        L.oracle_tet_scene_time.restype = ctypes.c_double
        L.oracle_tet_scene_solve_failures.argtypes = [ctypes.c_void_p]
        L.oracle_tet_scene_solve_failures.restype = ctypes.c_longlong
        L.oracle_tet_scene_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oracle_tet_scene_set_time.restype = None
        L.oracle_tet_scene_set_collision.argtypes = [ctypes.c_void_p, _dp]
        L.oracle_tet_scene_set_box_prefilter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oracle_tet_scene_set_box_prefilter.restype = None
        L.oracle_tet_scene_collisions.argtypes = [ctypes.c_void_p, _ip]
        L.oracle_tet_scene_surface.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip]
        L.oracle_point_in_tet_f32.argtypes = [_fp] * 5
        L.oracle_ccd_point_triangle_f64.argtypes = [_dp] * 6
        L.oracle_cubic_roots01_f64.argtypes = [_dp, _dp]
        L.oracle_tet_walk_f64.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, _dp, ctypes.c_double,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_penetration_force_f64.argtypes = [_dp, _dp, ctypes.c_int] + [ctypes.c_double] * 4 + [_dp, _dp]
        L.oracle_penetration_force_f64.restype = None
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def physics_vector(p):
    """Flat oracle parameter vector from a physics dict (gaia_amd.vbd.physics_dict keys)."""
    v = np.zeros(OP_COUNT)
    g = p["gravity"]
    v[0:3] = g
    v[OP["DT"]] = np.float32(np.float32(p["time_step"]) / np.float32(p["num_substeps"]))
    v[OP["NUM_SUBSTEPS"]] = p["num_substeps"]
    v[OP["ITERATIONS"]] = p["iterations"]
    v[6:9] = p["world_lower"]
    v[9:12] = p["world_upper"]
    v[OP["USE_PLANE_GROUND"]] = p["use_plane_ground"]
    v[OP["BOUNDARY_STIFFNESS"]] = p["boundary_collision_stiffness"]
    v[OP["BOUNDARY_FRICTION_MU"]] = p["boundary_friction_dynamic"]
    v[OP["BOUNDARY_FRICTION_EPS_V"]] = p["boundary_friction_eps_v"]
    v[OP["STEP_SIZE"]] = p["step_size"]
    v[OP["STEP_SIZE_GD"]] = p["step_size_gd"]
    v[OP["FRICTION_START_ITER"]] = p["friction_start_iter"]
    v[OP["ASSOCIATE_GRAVITY"]] = p["associate_gravity_with_inertia"]
    v[OP["INIT_RATIO_G"]] = p.get("init_ratio_g", 0.5)
    v[OP["USE_ACCELERATOR"]] = p.get("use_accelerator", 0)
    v[OP["ACCELERATOR_RHO"]] = p.get("accelerator_rho", 0.93)
    return v


# penetration contact of tet meshes (oracle.h OracleCollisionParam); defaults of
# CollisionDetertionParameters.h:43-80 and VBDPhysicsParameters.h:36,70
OC_KEYS = ("allow_dcd", "allow_ccd", "self_collision", "feasible_region", "tet_traverse", "traverse_non_self",
           "shift_to_center", "stop_after_passing", "stiffness", "off_height", "feasible_eps", "ray_eps",
           "max_search_mult", "center_shift")
OC_DEFAULTS = dict(allow_dcd=1, allow_ccd=1, self_collision=1, feasible_region=1, tet_traverse=1, traverse_non_self=1,
                   shift_to_center=1, stop_after_passing=1, stiffness=1e5, off_height=1e7, feasible_eps=1e-2,
                   ray_eps=1e-10, max_search_mult=1.8, center_shift=0.01)


def collision_vector(c):
    full = dict(OC_DEFAULTS, **(c or {}))
    return np.array([float(full[k]) for k in OC_KEYS])


def point_in_tet(p, a, b, c, d):
    """CollisionGeometry.h:56-88 in float: strictly inside a tet of |6V| >= 1e-5."""
    arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (p, a, b, c, d)]
    return bool(lib().oracle_point_in_tet_f32(*[x.ctypes.data_as(_fp) for x in arrs]))


def ccd_point_triangle(f0, f1, p0, p1):
    """Earliest hit of the point moving p0 -> p1 with the triangle moving f0 -> f1 (3x3 rows):
    (t, barycentrics) or None."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (f0, f1, p0, p1)]
    t = np.zeros(1)
    bary = np.zeros(3)
    hit = lib().oracle_ccd_point_triangle_f64(*[x.ctypes.data_as(_dp) for x in arrs], t.ctypes.data_as(_dp),
                                              bary.ctypes.data_as(_dp))
    return (float(t[0]), bary) if hit else None


def cubic_roots01(coef):
    """Roots in [0, 1] where c0 + c1 x + c2 x^2 + c3 x^3 changes sign, ascending."""
    c = np.ascontiguousarray(coef, dtype=np.float64)
    r = np.zeros(3)
    n = lib().oracle_cubic_roots01_f64(c.ctypes.data_as(_dp), r.ctypes.data_as(_dp))
    return r[:n]


def tet_walk(pos, tets, origin, direction, max_dis, start_tet, start_local, target):
    """The closest-point check's ray walk through the tets (TetMeshFEM::tetrahedralTraverseTo)."""
    x, px = _d(np.asarray(pos).reshape(-1))
    t, pt = _i(np.asarray(tets).reshape(-1))
    o, po = _d(origin)
    d, pd = _d(direction)
    return bool(lib().oracle_tet_walk_f64(len(x) // 3, px, len(t) // 4, pt, po, pd, float(max_dis), int(start_tet),
                                          int(start_local), int(target)))


def penetration_force(x4, xp4, order, k, mu=0.0, eps_v=1.0, dt=1.0 / 60.0):
    """Force and Hessian of one penetration contact on the vertex of `order` (3: the colliding
    vertex, 0..2: a vertex of its face); x4 rows = vertex, a, b, c."""
    x, px = _d(np.asarray(x4).reshape(-1))
    xp, pxp = _d(np.asarray(xp4).reshape(-1))
    f = np.zeros(3)
    h = np.zeros(9)
    lib().oracle_penetration_force_f64(px, pxp, int(order), float(k), float(mu), float(eps_v), float(dt),
                                       f.ctypes.data_as(_dp), h.ctypes.data_as(_dp))
    return f, h.reshape(3, 3).T


def material_vector(m):
    v = np.zeros(OM_COUNT)
    v[OM["DENSITY"]] = m.get("density", 1.0)
    v[OM["MU"]] = m.get("mu", 1e4)
    v[OM["LAMBDA"]] = m.get("lmbd", 1e4)
    v[OM["DAMP_HYDRO"]] = m.get("damping_hydrostatic", 0.0)
    v[OM["DAMP_DEV"]] = m.get("damping_deviatoric", 0.0)
    v[OM["EXP_VEL_DAMPING"]] = m.get("exponential_vel_damping", 1.0)
    v[OM["CONST_VEL_DAMPING"]] = m.get("constant_vel_damping", 0.0)
    v[OM["INIT_TYPE"]] = m.get("initialization_type", 0)
    v[OM["FRICTION_MU"]] = m.get("friction_dynamic", 0.1)
    v[OM["FRICTION_EPS_V"]] = m.get("friction_eps_v", 0.1)
    v[OM["NO_GRAV_ZONE"]] = m.get("has_no_grav_zone", 0)
    v[OM["NO_GRAV_THRESHOLD"]] = m.get("no_grav_zone_threshold", 0.0)
    v[OM["MAX_VELOCITY"]] = m.get("max_velocity_magnitude", -1.0)
    return v


def tet_prepare(pos, tets, density, precision=64):
    pos = np.asarray(pos).reshape(-1, 3)
    tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
    nv, nt = len(pos), len(tets)
    if precision == 64:
        p = np.ascontiguousarray(pos, dtype=np.float64)
        dminv, vol, mass = np.zeros(9 * nt), np.zeros(nt), np.zeros(nv)
        rc = lib().oracle_tet_prepare_f64(nv, p.ctypes.data_as(_dp), nt, tets.ctypes.data_as(_ip), density,
                                          dminv.ctypes.data_as(_dp), vol.ctypes.data_as(_dp), mass.ctypes.data_as(_dp))
    else:
        p = np.ascontiguousarray(pos, dtype=np.float32)
        dminv = np.zeros(9 * nt, np.float32)
        vol = np.zeros(nt, np.float32)
        mass = np.zeros(nv, np.float32)
        rc = lib().oracle_tet_prepare_f32(nv, p.ctypes.data_as(_fp), nt, tets.ctypes.data_as(_ip), density,
                                          dminv.ctypes.data_as(_fp), vol.ctypes.data_as(_fp), mass.ctypes.data_as(_fp))
    if rc != 0:
        raise ValueError("degenerate tet")
    return dminv.reshape(nt, 9), vol, mass


def neohookean_vertex(x4, dminv, vol, mu, lmbd, order):
    x4, px = _d(np.asarray(x4).reshape(-1))
    dm, pd = _d(np.asarray(dminv).reshape(-1))
    f = np.zeros(3)
    h = np.zeros(9)
    lib().oracle_neohookean_vertex_f64(px, pd, vol, mu, lmbd, order, f.ctypes.data_as(_dp), h.ctypes.data_as(_dp))
    return f, h.reshape(3, 3, order="F")


def solve3x3(h, b, precision=64):
    if precision == 64:
        h = np.asfortranarray(h, dtype=np.float64).reshape(-1, order="F").copy()
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(3)
        ok = lib().oracle_solve3x3_f64(h.ctypes.data_as(_dp), b.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
    else:
        h = np.asarray(h, dtype=np.float32).reshape(-1, order="F").copy()
        b = np.ascontiguousarray(b, dtype=np.float32)
        x = np.zeros(3, np.float32)
        ok = lib().oracle_solve3x3_f32(h.ctypes.data_as(_fp), b.ctypes.data_as(_fp), x.ctypes.data_as(_fp))
    return bool(ok), x


def color_tet_mesh(nv, tets, algorithm="mcs"):
    tets, pt = _i(np.asarray(tets).reshape(-1))
    colors = np.zeros(nv, np.int32)
    lib().oracle_color_tet_mesh(nv, tets.size // 4, pt, 0 if algorithm == "mcs" else 1, colors.ctypes.data_as(_ip))
    return colors


def color_graph(nv, edges, algorithm="mcs"):
    edges, pe = _i(np.asarray(edges).reshape(-1))
    colors = np.zeros(nv, np.int32)
    lib().oracle_color_graph(nv, edges.size // 2, pe, 0 if algorithm == "mcs" else 1, colors.ctypes.data_as(_ip))
    return colors


def deformer_args(d):
    """(kind, params, meshes, counts, vertices) of a deformer given as the reference's JSON object
    (VBD/VBDDeformer.h:181-222): DeformerName "Rotator" (center, axis, angularVelocity,
    rotationEndTime) or "Translator" (speed, deformationEndTime, deformationStartTime), with
    selectedMeshes and selectedVertices."""
    name = d["DeformerName"]
    if name not in ("Rotator", "Translator"):
        raise ValueError(f"unrecognised deformer {name!r}")
    rot = name == "Rotator"
    end = d.get("rotationEndTime" if rot else "deformationEndTime", -1)
    prm = np.array([*d.get("center", (0, 0, 0)), *d.get("axis", (0, 0, 0)), d.get("angularVelocity", 0.0), end,
                    *d.get("speed", (0, 0, 0)), d.get("deformationStartTime", 0.0)], dtype=np.float64)
    meshes = np.asarray(d["selectedMeshes"], dtype=np.int32)
    sel = [np.asarray(v, dtype=np.int32).reshape(-1) for v in d["selectedVertices"]]
    if len(sel) != len(meshes):
        raise ValueError("selectedVertices needs one list per selected mesh")
    counts = np.array([len(v) for v in sel], dtype=np.int32)
    verts = np.concatenate(sel) if sel else np.zeros(0, np.int32)
    return 0 if rot else 1, prm, meshes, counts, verts.astype(np.int32)


def _add_deformer(fn, h, d):
    kind, prm, meshes, counts, verts = deformer_args(d)
    a, pa = _d(prm)
    m, pm = _i(meshes)
    c, pc = _i(counts)
    v, pv = _i(verts)
    if fn(h, kind, pa, len(meshes), pm, pc, pv) < 0:
        raise ValueError("oracle: bad deformer")


class TetScene:
    """Oracle tet VBD scene (VBDPhysics::runStep_serialCollisionHandling restated)."""

    def __init__(self, physics, precision=32):
        pv, pp = _d(physics_vector(physics))
        self._h = lib().oracle_tet_scene_create(precision, pp)
        self.nmesh = 0
        self.nv = []

    def add_mesh(self, pos, tets, material, colors, fixed=(), vel=None):
        pos = np.asarray(pos).reshape(-1, 3)
        nv = len(pos)
        p, pp = _d(pos.reshape(-1))
        v = None
        pv = None
        if vel is not None:
            v, pv = _d(np.asarray(vel).reshape(-1))
        t, pt = _i(np.asarray(tets).reshape(-1))
        m, pm = _d(material_vector(material))
        f, pf = _i(np.asarray(fixed, dtype=np.int32).reshape(-1))
        c, pc = _i(colors)
        rc = lib().oracle_tet_scene_add_mesh(self._h, nv, pp, pv, t.size // 4, pt, pm, f.size, pf, pc)
        if rc < 0:
            raise ValueError("oracle: bad mesh")
        self.nmesh += 1
        self.nv.append(nv)
        return rc

    def finalize(self):
        lib().oracle_tet_scene_finalize(self._h)

    def set_collision(self, params=None):
        """Penetration contact (DCD + CCD, runStepGPU's order) with the given parameters over the
        reference's defaults (keys of OC_KEYS); before finalize."""
        v, pv = _d(collision_vector(params))
        lib().oracle_tet_scene_set_collision(self._h, pv)

    def set_box_prefilter(self, on):
        """Box prefilter of the brute-force DCD / CCD (default on; the results are the same)."""
        lib().oracle_tet_scene_set_box_prefilter(self._h, int(bool(on)))

    def collisions(self):
        """The last detection, one row per surface vertex: (mesh, vertex, hit mesh, hit face,
        penetrated, from CCD); hit face -1 = none."""
        n = lib().oracle_tet_scene_collisions(self._h, None)
        out = np.zeros((n, 6), np.int32)
        lib().oracle_tet_scene_collisions(self._h, out.ctypes.data_as(_ip))
        return out

    def surface(self, mesh=0):
        """(faces (F, 3), owner (F, 2): owning tet and local face) of the collision model."""
        nf = lib().oracle_tet_scene_surface(self._h, mesh, None, None)
        faces = np.zeros((nf, 3), np.int32)
        owner = np.zeros((nf, 2), np.int32)
        lib().oracle_tet_scene_surface(self._h, mesh, faces.ctypes.data_as(_ip), owner.ctypes.data_as(_ip))
        return faces, owner

    def groups(self):
        out = []
        for g in range(lib().oracle_tet_scene_num_groups(self._h)):
            n = lib().oracle_tet_scene_group(self._h, g, None)
            buf = np.zeros(2 * n, np.int32)
            lib().oracle_tet_scene_group(self._h, g, buf.ctypes.data_as(_ip))
            out.append(buf.reshape(-1, 2))
        return out

    def step(self, n=1, threads=1):
        lib().oracle_tet_scene_step(self._h, n, threads)

    def add_deformer(self, d):
        """A deformer as the reference's JSON object (before finalize)."""
        _add_deformer(lib().oracle_tet_scene_add_deformer, self._h, d)

    @property
    def time(self):
        return lib().oracle_tet_scene_time(self._h)

    def solve_failures(self):
        """Vertex solves whose 3x3 block had no solution so far."""
        return lib().oracle_tet_scene_solve_failures(self._h)

    @time.setter
    def time(self, t):
        lib().oracle_tet_scene_set_time(self._h, float(t))

    def evaluate_convergence(self):
        out = np.zeros(4)
        lib().oracle_tet_scene_evaluate_convergence(self._h, out.ctypes.data_as(_dp))
        return dict(zip(("energy", "inertia_energy", "elastic_energy", "avg_force_norm"), out.tolist()))

    def get_state(self, mesh=0):
        nv = self.nv[mesh]
        x = np.zeros(3 * nv)
        v = np.zeros(3 * nv)
        lib().oracle_tet_scene_get_state(self._h, mesh, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp))
        return x.reshape(nv, 3), v.reshape(nv, 3)

    def set_state(self, mesh, x=None, v=None):
        px = pv = None
        if x is not None:
            xa, px = _d(np.asarray(x).reshape(-1))
        if v is not None:
            va, pv = _d(np.asarray(v).reshape(-1))
        lib().oracle_tet_scene_set_state(self._h, mesh, px, pv)

    def __del__(self):
        try:
            if self._h:
                lib().oracle_tet_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass


# ---------------------------------------------------------------- cloth
OP.update(CLOTH_BOUNDARY_STIFFNESS=22, CLOTH_BOUNDARY_FRICTION_EPS=23, DEGENERATE_THRES=24, HANDLE_COLLISION=25,
          CONTACT_STIFFNESS=26, CONTACT_RADIUS=27, THICKNESS=28, CONSERVATIVE_RELAXATION=29, CONTACT_ENERGY_TYPE=30,
          MAX_QUERY_DIS=31)
OMC = dict(DENSITY=0, MU=1, LAMBDA=2, BENDING=3, DAMP_STVK=4, DAMP_BENDING=5, EXP_VEL_DAMPING=6,
           CONST_VEL_DAMPING=7, INIT_TYPE=8, FRICTION_MU=9, FRICTION_EPS_V=10, NO_GRAV_ZONE=11, NO_GRAV_THRESHOLD=12,
           MAX_VELOCITY=13)
OMC_COUNT = 16


def _cloth_lib():
    L = lib()
    if not getattr(L, "_cloth_bound", False):
        L.oracle_cloth_prepare.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_double, _dp,
                                           _dp, _dp, _ip, _ip, _dp]
        L.oracle_stvk_vertex_f64.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_int, _dp, _dp]
        L.oracle_stvk_vertex_f64.restype = None
        L.oracle_cloth_scene_create.argtypes = [ctypes.c_int, _dp]
        L.oracle_cloth_scene_create.restype = ctypes.c_void_p
        L.oracle_cloth_scene_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_cloth_scene_destroy.restype = None
        L.oracle_cloth_scene_add_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _dp,
                                                  ctypes.c_int, _ip, _ip]
        L.oracle_cloth_scene_add_tet.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, _dp,
                                                 ctypes.c_int, _ip, _ip]
        L.oracle_cloth_scene_surface.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_cloth_scene_finalize.argtypes = [ctypes.c_void_p]
        L.oracle_cloth_scene_num_groups.argtypes = [ctypes.c_void_p]
        L.oracle_cloth_scene_group.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip]
        L.oracle_cloth_scene_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.oracle_cloth_scene_get_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_cloth_scene_set_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.oracle_cloth_scene_add_deformer.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, _ip]
        L.oracle_cloth_scene_time.argtypes = [ctypes.c_void_p]
        L.oracle_cloth_scene_time.restype = ctypes.c_double
        L.oracle_cloth_scene_solve_failures.argtypes = [ctypes.c_void_p]
        L.oracle_cloth_scene_solve_failures.restype = ctypes.c_longlong
        L.oracle_cloth_scene_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oracle_cloth_scene_set_time.restype = None
        L.oracle_cloth_scene_add_attachments.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp]
        L.oracle_color_tri_mesh.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, _ip]
        L._cloth_bound = True
    return L


def cloth_physics_vector(p):
    v = physics_vector(p)
    v[OP["CLOTH_BOUNDARY_STIFFNESS"]] = p.get("cloth_boundary_contact_stiffness", 1e4)
    v[OP["CLOTH_BOUNDARY_FRICTION_EPS"]] = p.get("cloth_boundary_friction_eps", 1e-2)
    v[OP["DEGENERATE_THRES"]] = p.get("degenerate_triangle_thres", 1e-6)
    v[OP["STEP_SIZE"]] = p.get("cloth_step_size", 1.0)
    v[OP["HANDLE_COLLISION"]] = p.get("handle_collision", 0)
    v[OP["CONTACT_STIFFNESS"]] = p.get("contact_stiffness", 1e4)
    v[OP["CONTACT_RADIUS"]] = p.get("contact_radius", 1.0)
    v[OP["THICKNESS"]] = p.get("thickness", 0.0)
    v[OP["CONSERVATIVE_RELAXATION"]] = p.get("conservative_step_relaxation", 0.45)
    v[OP["CONTACT_ENERGY_TYPE"]] = p.get("contact_energy_type", 1)
    v[OP["MAX_QUERY_DIS"]] = p.get("contact_query_radius", 0.4)
    return v


def cloth_material_vector(m):
    v = np.zeros(OMC_COUNT)
    v[OMC["DENSITY"]] = m.get("density", 1.0)
    v[OMC["MU"]] = m.get("mu", 1e4)
    v[OMC["LAMBDA"]] = m.get("lmbd", 1e4)
    v[OMC["BENDING"]] = m.get("bending_stiffness", 0.0)
    v[OMC["DAMP_STVK"]] = m.get("damping_stvk", 0.0)
    v[OMC["DAMP_BENDING"]] = m.get("damping_bending", 0.0)
    v[OMC["EXP_VEL_DAMPING"]] = m.get("exponential_vel_damping", 1.0)
    v[OMC["CONST_VEL_DAMPING"]] = m.get("constant_vel_damping", 0.0)
    v[OMC["INIT_TYPE"]] = m.get("initialization_type", 2)
    v[OMC["FRICTION_MU"]] = m.get("friction_dynamic", 0.1)
    v[OMC["FRICTION_EPS_V"]] = m.get("friction_eps_v", 0.1)
    v[OMC["NO_GRAV_ZONE"]] = m.get("has_no_grav_zone", 0)
    v[OMC["NO_GRAV_THRESHOLD"]] = m.get("no_grav_zone_threshold", 0.0)
    v[OMC["MAX_VELOCITY"]] = m.get("max_velocity_magnitude", -1.0)
    return v


def cloth_prepare(pos, faces, density, precision=64):
    L = _cloth_lib()
    p, pp = _d(np.asarray(pos).reshape(-1))
    f, pf = _i(np.asarray(faces).reshape(-1))
    nv, nf = p.size // 3, f.size // 3
    ne = ctypes.c_int32()
    rc = L.oracle_cloth_prepare(precision, nv, pp, nf, pf, density, None, None, None, ctypes.byref(ne), None, None)
    if rc != 0:
        raise ValueError("oracle: bad cloth mesh")
    dminv, area, mass = np.zeros(4 * nf), np.zeros(nf), np.zeros(nv)
    edges = np.zeros(6 * ne.value, np.int32)
    Q = np.zeros(16 * ne.value)
    L.oracle_cloth_prepare(precision, nv, pp, nf, pf, density, dminv.ctypes.data_as(_dp), area.ctypes.data_as(_dp),
                           mass.ctypes.data_as(_dp), ctypes.byref(ne), edges.ctypes.data_as(_ip),
                           Q.ctypes.data_as(_dp))
    return dminv.reshape(nf, 4), area, mass, edges.reshape(-1, 6), Q.reshape(-1, 4, 4)


def stvk_vertex(x3, dminv4, area, mu, lmbd, order):
    L = _cloth_lib()
    x, px = _d(np.asarray(x3).reshape(-1))
    dm, pd = _d(np.asarray(dminv4).reshape(-1))
    f = np.zeros(3)
    h = np.zeros(9)
    L.oracle_stvk_vertex_f64(px, pd, area, mu, lmbd, order, f.ctypes.data_as(_dp), h.ctypes.data_as(_dp))
    return f, h.reshape(3, 3, order="F")


def color_tri_mesh(nv, faces, algorithm="mcs", with_bending=True):
    L = _cloth_lib()
    f, pf = _i(np.asarray(faces).reshape(-1))
    colors = np.zeros(nv, np.int32)
    L.oracle_color_tri_mesh(nv, f.size // 3, pf, 1 if with_bending else 0, 0 if algorithm == "mcs" else 1,
                            colors.ctypes.data_as(_ip))
    return colors


def attachment_arrays(params):
    """Rows (mesh, vertex, target mesh, f0, f1, f2) and (b0, b1, b2, stiffness) from the
    reference's VertexFaceAttachmentConstraintParams (VBDClothPhysics.cpp:148-224; stiffness 1e4
    when vertexConstraintStiffness is empty)."""
    vi = np.asarray(params["vertexConstraintVertexInfos"], dtype=np.int32).reshape(-1, 2)
    tf = np.asarray(params["vertexConstraintTargetFacesInfos"], dtype=np.int32).reshape(-1, 4)
    bary = np.asarray(params["vertexConstraintBarys"], dtype=np.float64).reshape(-1, 3)
    k = np.asarray(params.get("vertexConstraintStiffness", []), dtype=np.float64).reshape(-1)
    n = len(vi)
    if len(tf) != n or len(bary) != n or (len(k) and len(k) != n):
        raise ValueError("attachment arrays disagree in length")
    if not len(k):
        k = np.full(n, 1e4)
    return np.concatenate([vi, tf], axis=1).astype(np.int32), np.concatenate([bary, k[:, None]], axis=1)


class ClothScene:
    """Oracle cloth VBD scene (VBDClothSimulationFramework::runStep_CDEveryIter, no contact)."""

    def __init__(self, physics, precision=32):
        L = _cloth_lib()
        pv, pp = _d(cloth_physics_vector(physics))
        self._h = L.oracle_cloth_scene_create(precision, pp)
        self.nv = []

    def add_mesh(self, pos, faces, material, colors, fixed=(), vel=None):
        L = _cloth_lib()
        pos = np.asarray(pos).reshape(-1, 3)
        nv = len(pos)
        p, pp = _d(pos.reshape(-1))
        pv = None
        if vel is not None:
            v, pv = _d(np.asarray(vel).reshape(-1))
        f, pf = _i(np.asarray(faces).reshape(-1))
        m, pm = _d(cloth_material_vector(material))
        fx, pfx = _i(np.asarray(fixed, dtype=np.int32).reshape(-1))
        c, pc = _i(colors)
        rc = L.oracle_cloth_scene_add_mesh(self._h, nv, pp, pv, f.size // 3, pf, pm, fx.size, pfx, pc)
        if rc < 0:
            raise ValueError("oracle: bad cloth mesh")
        self.nv.append(nv)
        return rc

    def add_tet_mesh(self, pos, tets, material, colors, fixed=(), vel=None):
        """A neo-Hookean tet mesh whose boundary surface takes part in contact (scenes with
        collision handling): the tet vertex solve plus cloth-model contact on surface vertices."""
        L = _cloth_lib()
        pos = np.asarray(pos).reshape(-1, 3)
        nv = len(pos)
        p, pp = _d(pos.reshape(-1))
        pv = None
        if vel is not None:
            v, pv = _d(np.asarray(vel).reshape(-1))
        t, pt = _i(np.asarray(tets).reshape(-1))
        m, pm = _d(material_vector(material))
        fx, pfx = _i(np.asarray(fixed, dtype=np.int32).reshape(-1))
        c, pc = _i(colors)
        rc = L.oracle_cloth_scene_add_tet(self._h, nv, pp, pv, t.size // 4, pt, pm, fx.size, pfx, pc)
        if rc < 0:
            raise ValueError("oracle: bad tet mesh")
        self.nv.append(nv)
        return rc

    def surface(self, mesh):
        """The mesh's contact faces (a tet mesh's boundary surface, outward oriented)."""
        L = _cloth_lib()
        n = L.oracle_cloth_scene_surface(self._h, mesh, None)
        buf = np.zeros(3 * n, np.int32)
        L.oracle_cloth_scene_surface(self._h, mesh, buf.ctypes.data_as(_ip))
        return buf.reshape(-1, 3)

    def finalize(self):
        if _cloth_lib().oracle_cloth_scene_finalize(self._h) != 0:
            raise ValueError("oracle: contact detector setup failed")

    def groups(self):
        L = _cloth_lib()
        out = []
        for g in range(L.oracle_cloth_scene_num_groups(self._h)):
            n = L.oracle_cloth_scene_group(self._h, g, None)
            buf = np.zeros(2 * n, np.int32)
            L.oracle_cloth_scene_group(self._h, g, buf.ctypes.data_as(_ip))
            out.append(buf.reshape(-1, 2))
        return out

    def step(self, n=1, threads=1):
        _cloth_lib().oracle_cloth_scene_step(self._h, n, threads)

    def add_deformer(self, d):
        """A deformer as the reference's JSON object (before finalize)."""
        _add_deformer(_cloth_lib().oracle_cloth_scene_add_deformer, self._h, d)

    def add_attachments(self, params):
        """Vertex-face attachments given as the reference's VertexFaceAttachmentConstraintParams
        object (VBDCloth/VBDTriMeshConstraints.h:6-29)."""
        ints, vals = attachment_arrays(params)
        i, pi = _i(ints.reshape(-1))
        v, pv = _d(vals.reshape(-1))
        if _cloth_lib().oracle_cloth_scene_add_attachments(self._h, len(ints), pi, pv) != 0:
            raise ValueError("oracle: bad attachment")

    @property
    def time(self):
        return _cloth_lib().oracle_cloth_scene_time(self._h)

    def solve_failures(self):
        """Vertex solves whose 3x3 block had no solution so far."""
        return _cloth_lib().oracle_cloth_scene_solve_failures(self._h)

    @time.setter
    def time(self, t):
        _cloth_lib().oracle_cloth_scene_set_time(self._h, float(t))

    def get_state(self, mesh=0):
        nv = self.nv[mesh]
        x = np.zeros(3 * nv)
        v = np.zeros(3 * nv)
        _cloth_lib().oracle_cloth_scene_get_state(self._h, mesh, x.ctypes.data_as(_dp), v.ctypes.data_as(_dp))
        return x.reshape(nv, 3), v.reshape(nv, 3)

    def set_state(self, mesh, x=None, v=None):
        px = pv = None
        if x is not None:
            xa, px = _d(np.asarray(x).reshape(-1))
        if v is not None:
            va, pv = _d(np.asarray(v).reshape(-1))
        _cloth_lib().oracle_cloth_scene_set_state(self._h, mesh, px, pv)

    def __del__(self):
        try:
            if self._h:
                _cloth_lib().oracle_cloth_scene_destroy(self._h)
                self._h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------
# contact detector (detector_oracle.cpp)
# ---------------------------------------------------------------------------
_cd = None
_i64p = ctypes.POINTER(ctypes.c_int64)


def _cd_lib():
    global _cd
    if _cd is None:
        L = lib()
        L.oracle_cd_create.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.oracle_cd_create.restype = ctypes.c_void_p
        L.oracle_cd_destroy.argtypes = [ctypes.c_void_p]
        L.oracle_cd_destroy.restype = None
        L.oracle_cd_sizes.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        L.oracle_cd_sizes.restype = None
        L.oracle_cd_topology.argtypes = [ctypes.c_void_p, _ip, _ip]
        L.oracle_cd_topology.restype = None
        L.oracle_cd_detect.argtypes = [ctypes.c_void_p, _fp, ctypes.c_float, ctypes.c_int, ctypes.c_int]
        L.oracle_cd_detect.restype = None
        L.oracle_cd_counts.argtypes = [ctypes.c_void_p, _i64p]
        L.oracle_cd_counts.restype = None
        L.oracle_cd_vf_pairs.argtypes = [ctypes.c_void_p, _ip, _fp]
        L.oracle_cd_vf_pairs.restype = None
        L.oracle_cd_vf_contacts.argtypes = [ctypes.c_void_p, _ip, _fp]
        L.oracle_cd_vf_contacts.restype = None
        L.oracle_cd_ee_contacts.argtypes = [ctypes.c_void_p, _ip, _fp, _fp]
        L.oracle_cd_ee_contacts.restype = None
        L.oracle_cd_min_dis.argtypes = [ctypes.c_void_p, _fp, _fp, _fp]
        L.oracle_cd_min_dis.restype = None
        L.oracle_cd_bounds.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, _fp]
        L.oracle_cd_bounds.restype = None
        L.oracle_closest_point_triangle.argtypes = [_fp, _fp, _fp, _fp, _fp, _fp]
        L.oracle_segment_closest_points.argtypes = [_fp, _fp, _fp, _fp, _fp, _fp, _fp]
        L.oracle_segment_closest_points.restype = None
        _cd = L
    return _cd


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def closest_point_triangle(p, a, b, c):
    """(closest point, barycentrics, type) with type 0..6 = A, B, C, AB, BC, AC, interior."""
    args = [_f(x)[0] for x in (p, a, b, c)]
    out = np.zeros(3, np.float32)
    bary = np.zeros(3, np.float32)
    t = _cd_lib().oracle_closest_point_triangle(*[x.ctypes.data_as(_fp) for x in args], out.ctypes.data_as(_fp),
                                               bary.ctypes.data_as(_fp))
    return out, bary, t


def segment_closest_points(p0, p1, q0, q1):
    args = [_f(x)[0] for x in (p0, p1, q0, q1)]
    cp, cq, st = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(2, np.float32)
    _cd_lib().oracle_segment_closest_points(*[x.ctypes.data_as(_fp) for x in args], cp.ctypes.data_as(_fp),
                                           cq.ctypes.data_as(_fp), st.ctypes.data_as(_fp))
    return cp, cq, float(st[0]), float(st[1])


class Detector:
    """VF / EE proximity queries over triangle meshes (global vertex / face / edge ids,
    meshes concatenated in order). mode "grid" (uniform grid) or "brute" (all pairs)."""

    def __init__(self, meshes):
        L = _cd_lib()
        nv = np.array([len(p) for p, _ in meshes], np.int32)
        nf = np.array([len(f) for _, f in meshes], np.int32)
        faces = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32).reshape(-1) for _, f in meshes]))
        self._h = L.oracle_cd_create(len(meshes), nv.ctypes.data_as(_ip), nf.ctypes.data_as(_ip),
                                     faces.ctypes.data_as(_ip))
        if not self._h:
            raise ValueError("invalid mesh (index out of range, non-manifold or inconsistent orientation)")
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.oracle_cd_sizes(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        self.nv, self.nf, self.ne = a.value, b.value, c.value

    def topology(self):
        edges = np.zeros((self.ne, 4), np.int32)
        fe = np.zeros((self.nf, 3), np.int32)
        _cd_lib().oracle_cd_topology(self._h, edges.ctypes.data_as(_ip), fe.ctypes.data_as(_ip))
        return edges, fe

    def detect(self, positions, radius, mode="grid", threads=8):
        L = _cd_lib()
        p = np.ascontiguousarray(positions, np.float32).reshape(-1)
        self.radius = np.float32(radius)
        L.oracle_cd_detect(self._h, p.ctypes.data_as(_fp), ctypes.c_float(radius), 1 if mode == "grid" else 0,
                           threads)
        cnt = np.zeros(3, np.int64)
        L.oracle_cd_counts(self._h, cnt.ctypes.data_as(_i64p))
        n_pairs, n_vf, n_ee = (int(x) for x in cnt)
        out = {}
        vf = np.zeros((n_pairs, 2), np.int32)
        d = np.zeros(n_pairs, np.float32)
        L.oracle_cd_vf_pairs(self._h, vf.ctypes.data_as(_ip), d.ctypes.data_as(_fp))
        out["vf_pairs"], out["vf_pair_d"] = vf, d
        rec = np.zeros((n_vf, 4), np.int32)
        d = np.zeros(n_vf, np.float32)
        L.oracle_cd_vf_contacts(self._h, rec.ctypes.data_as(_ip), d.ctypes.data_as(_fp))
        out["vf_contacts"], out["vf_contact_d"] = rec, d
        ee = np.zeros((n_ee, 2), np.int32)
        mu = np.zeros((n_ee, 2), np.float32)
        d = np.zeros(n_ee, np.float32)
        L.oracle_cd_ee_contacts(self._h, ee.ctypes.data_as(_ip), mu.ctypes.data_as(_fp), d.ctypes.data_as(_fp))
        out["ee_contacts"], out["ee_mu"], out["ee_d"] = ee, mu, d
        a, b, c = np.zeros(self.nv, np.float32), np.zeros(self.nf, np.float32), np.zeros(self.ne, np.float32)
        L.oracle_cd_min_dis(self._h, a.ctypes.data_as(_fp), b.ctypes.data_as(_fp), c.ctypes.data_as(_fp))
        out["vf_min"], out["face_min"], out["edge_min"] = a, b, c
        return out

    def bounds(self, thickness, relaxation):
        b = np.zeros(self.nv, np.float32)
        _cd_lib().oracle_cd_bounds(self._h, ctypes.c_float(thickness), ctypes.c_float(relaxation),
                                   b.ctypes.data_as(_fp))
        return b

    def __del__(self):
        if getattr(self, "_h", None):
            _cd_lib().oracle_cd_destroy(self._h)
            self._h = None
