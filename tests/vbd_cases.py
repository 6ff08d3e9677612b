"""Shared scene specs for the parity tests: the same spec builds a HIP scene
(through the C ABI) and an oracle scene."""
import numpy as np

from gaia_amd import VBDPhysics, scenes
from gaia_amd.vbd import physics_dict, physics_params


def rel_err(x, ref):
    """Normwise relative position error: max |x - ref| / max |ref|."""
    return float(np.abs(np.asarray(x, np.float64) - ref).max() / max(np.abs(ref).max(), 1e-30))


def c1_spec():
    c = scenes.c1_soft_cube()
    mesh = dict(pos=c["positions"], tets=c["tets"], material=c["material"], fixed=(), vel=None,
                coloring=c["coloring"])
    return dict(physics=physics_dict(**c["physics"]), meshes=[mesh], steps=c["steps"])


def multi_spec():
    """Two cubes of different materials, one with fixed points, damping and a
    non-default initialisation type; ground at y = 0 with friction."""
    p1, t1 = scenes.cube_tet_mesh(4, 0.1, origin=(-0.6, 0.05, -0.2))
    p2, t2 = scenes.cube_tet_mesh(3, 0.12, origin=(0.3, 0.3, -0.2))
    rng = np.random.default_rng(4)
    v1 = rng.normal(scale=0.2, size=p1.shape)
    m1 = dict(density=800.0, mu=5e4, lmbd=4e5, damping_hydrostatic=0.01, damping_deviatoric=0.02,
              exponential_vel_damping=0.9, constant_vel_damping=0.01)
    m2 = dict(density=1200.0, mu=2e5, lmbd=1e6, initialization_type=2)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=2, iterations=6,
                           world_lower=(-1e7, 0.0, -1e7), world_upper=(1e7, 1e7, 1e7),
                           boundary_friction_dynamic=0.3, friction_start_iter=1)
    top = [i for i in range(len(p2)) if p2[i, 1] > p2[:, 1].max() - 1e-9][:3]
    meshes = [dict(pos=p1, tets=t1, material=m1, fixed=(), vel=v1, coloring="mcs"),
              dict(pos=p2, tets=t2, material=m2, fixed=top, vel=None, coloring="greedy")]
    return dict(physics=physics, meshes=meshes, steps=20)


def build_gpu(ctx, spec):
    import gaia_amd

    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**spec["physics"])
    for m in spec["meshes"]:
        if "faces" in m:
            colors = gaia_amd.color_tri_mesh(len(m["pos"]), m["faces"], m["coloring"])
            sim.addClothMesh(m["pos"], m["faces"], fixed_points=m["fixed"], vertex_colors=colors,
                             velocities=m["vel"], **m["material"])
        else:
            colors = gaia_amd.color_tet_mesh(len(m["pos"]), m["tets"], m["coloring"])
            sim.addTetMesh(m["pos"], m["tets"], fixed_points=m["fixed"], vertex_colors=colors, velocities=m["vel"],
                           **m["material"])
    for d in spec.get("deformers", []):
        sim.addDeformer(d)
    if "attachments" in spec:
        sim.addAttachments(spec["attachments"])
    sim.initialize()
    return sim


def build_oracle_cloth(oracle, spec, precision=32):
    import gaia_amd

    sc = oracle.ClothScene(spec["physics"], precision=precision)
    for m in spec["meshes"]:
        colors = gaia_amd.color_tri_mesh(len(m["pos"]), m["faces"], m["coloring"])
        pos = np.asarray(m["pos"], np.float32).astype(np.float64) if precision == 32 else m["pos"]
        vel = None if m["vel"] is None else (np.asarray(m["vel"], np.float32).astype(np.float64)
                                             if precision == 32 else m["vel"])
        sc.add_mesh(pos, m["faces"], m["material"], colors, fixed=m["fixed"], vel=vel)
    for d in spec.get("deformers", []):
        sc.add_deformer(d)
    if "attachments" in spec:
        sc.add_attachments(spec["attachments"])
    sc.finalize()
    return sc


def curtain_spec(n=24, steps=40, init=2, **mat):
    """Vertical sheet hanging from its whole top row: the membrane is in tension, so one VBD
    step is well-conditioned (FP32 and FP64 oracles agree to ~1e-6 over 30 steps). A sheet
    pinned at two corners sags into compression, where the exact StVK Hessian is indefinite:
    there FP32 and FP64 of the same algorithm already differ by ~1e-2 after one step, so the
    1e-5 per-step contract is checked on tensioned sheets and on the first step from rest."""
    pos, faces, top = scenes.cloth_curtain(n, 1.0, 2.0)
    mu, lm = scenes.stvk_lame(1e5, 0.3)
    material = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4, initialization_type=init)
    material.update(mat)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                           world_lower=(-1e7, 0.0, -1e7))
    return dict(physics=physics, steps=steps,
                meshes=[dict(pos=pos, faces=faces, material=material, fixed=top, vel=None, coloring="mcs")])


def two_corner_sheet_spec(n=24):
    """The batched workload's cloth scene: horizontal sheet pinned at two corners."""
    pos, faces = scenes.cloth_sheet(n, 1.0, 1.5)
    mu, lm = scenes.stvk_lame(1e5, 0.3)
    material = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4, initialization_type=0)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                           world_lower=(-1e7, 0.0, -1e7))
    return dict(physics=physics, steps=1,
                meshes=[dict(pos=pos, faces=faces, material=material, fixed=[0, n - 1], vel=None, coloring="mcs")])


def cloth_drop_spec():
    """Two meshes (merged parallel groups, colours of one mesh may share a group): a flat sheet
    sliding onto the ground (penalty + friction, init type 0, membrane damping) and a damped
    curtain (init type 2), with substeps and greedy colouring on the curtain."""
    p1, f1 = scenes.cloth_sheet(10, 0.6, 0.02)
    p2, f2, top = scenes.cloth_curtain(7, 0.4, 1.0)
    v1 = np.tile([0.05, -0.2, 0.0], (len(p1), 1))
    m1 = dict(density=0.3, mu=2e3, lmbd=3e3, bending_stiffness=1e-3, damping_stvk=1e-3, damping_bending=1e-3,
              initialization_type=0, exponential_vel_damping=0.98)
    m2 = dict(density=0.1, mu=5e3, lmbd=5e3, bending_stiffness=1e-4, damping_stvk=2e-3, initialization_type=2)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=2, iterations=8,
                           world_lower=(-1e7, 0.0, -1e7), boundary_friction_dynamic=0.4)
    return dict(physics=physics, steps=30,
                meshes=[dict(pos=p1, faces=f1, material=m1, fixed=[], vel=v1, coloring="mcs"),
                        dict(pos=p2, faces=f2, material=m2, fixed=top, vel=None, coloring="greedy")])


def build_oracle(oracle, spec, precision=32):
    import gaia_amd

    sc = oracle.TetScene(spec["physics"], precision=precision)
    for m in spec["meshes"]:
        colors = gaia_amd.color_tet_mesh(len(m["pos"]), m["tets"], m["coloring"])
        pos = np.asarray(m["pos"], np.float32).astype(np.float64) if precision == 32 else m["pos"]
        vel = None if m["vel"] is None else (np.asarray(m["vel"], np.float32).astype(np.float64)
                                             if precision == 32 else m["vel"])
        sc.add_mesh(pos, m["tets"], m["material"], colors, fixed=m["fixed"], vel=vel)
    for d in spec.get("deformers", []):
        sc.add_deformer(d)
    ph = spec["physics"]
    if ph.get("handle_collision") and ph.get("tet_contact_model") == 1:
        sc.set_collision(penetration_params(ph))
    sc.finalize()
    return sc


def penetration_params(ph):
    """The oracle's collision parameters from a physics dict with the penetration contact model
    (values rounded to float as the device takes them)."""
    f32 = lambda x: float(np.float32(x))  # noqa: E731
    return dict(allow_dcd=ph["allow_dcd"], allow_ccd=ph["allow_ccd"], self_collision=ph["handle_self_collision"],
                feasible_region=ph["check_feasible_region"], tet_traverse=ph["check_tet_traverse"],
                traverse_non_self=ph["tet_traverse_non_self"], shift_to_center=ph["shift_query_point_to_center"],
                stop_after_passing=ph["stop_traversing_after_passing"], stiffness=f32(ph["collision_stiffness"]),
                off_height=f32(ph["collision_off_height"]), feasible_eps=f32(ph["feasible_region_epsilon"]),
                ray_eps=f32(ph["ray_tri_intersection_epsilon"]),
                max_search_mult=f32(ph["max_search_distance_multiplier"]), center_shift=f32(ph["center_shift_level"]))


def penetration_spec(steps=30, iterations=10, stiffness=1e5, vy=-1.5, bodies=2, n=3, h=0.1, seed=0):
    """Tet cubes falling onto each other and the ground with the penetration contact model: cube 0
    rests on the ground, the others start above it with a downward velocity and a seeded lateral
    offset, so they hit it (and each other) within a few frames."""
    rng = np.random.default_rng(seed)
    meshes = []
    for b in range(bodies):
        off = (0.0, 0.0, 0.0) if b == 0 else (rng.uniform(-0.1, 0.1), 0.32 * b, rng.uniform(-0.1, 0.1))
        pos, tets = scenes.cube_tet_mesh(n, h, origin=off)
        vel = None if b == 0 else np.tile([0.0, vy, 0.0], (len(pos), 1))
        meshes.append(dict(pos=pos, tets=tets, material=dict(density=1000.0, mu=2e5, lmbd=1e6, friction_dynamic=0.2,
                                                             friction_eps_v=0.1),
                           fixed=(), vel=vel, coloring="greedy"))
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=iterations,
                           world_lower=(-1e7, 0.0, -1e7), handle_collision=1, tet_contact_model=1,
                           collision_stiffness=stiffness)
    return dict(physics=physics, meshes=meshes, steps=steps)


def contact_drape_spec(steps=20, energy=1, density=1.0, stiffness=30.0, friction=0.2, iterations=8):
    """Cloth contact (VF/FV/EE repulsion + IPC friction + conservative bounds): a free sheet
    falling with a sideways drift onto a fixed, slightly tilted sheet 1.2 contact radii below."""
    lo, flo = scenes.cloth_sheet(12, 0.6, 0.0, origin=(-0.3, -0.3))
    lo[:, 1] = 0.02 * lo[:, 0]  # tilt: the upper sheet slides under friction
    up, fup = scenes.cloth_sheet(10, 0.4, 0.0, origin=(-0.2, -0.2))
    up[:, 1] = 0.02 * up[:, 0] + 0.012
    mu, lm = scenes.stvk_lame(1e4, 0.3)
    m_lo = dict(density=density, mu=mu, lmbd=lm, bending_stiffness=1e-4, friction_dynamic=friction,
                friction_eps_v=0.01)
    m_up = dict(m_lo, initialization_type=0)
    vel = np.tile([0.05, -0.3, 0.0], (len(up), 1))
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=iterations,
                           world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=stiffness,
                           contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45,
                           contact_energy_type=energy, contact_query_radius=0.02, contact_bvh_rebuild_steps=4)
    return dict(physics=physics, steps=steps,
                meshes=[dict(pos=lo, faces=flo, material=m_lo, fixed=list(range(len(lo))), vel=None, coloring="mcs"),
                        dict(pos=up, faces=fup, material=m_up, fixed=[], vel=vel, coloring="mcs")])


def single_mesh(spec):
    """The same two sheets as one mesh of two components (self-contact within a mesh)."""
    lo, up = spec["meshes"]
    pos = np.concatenate([lo["pos"], up["pos"]])
    faces = np.concatenate([lo["faces"], up["faces"] + len(lo["pos"])])
    vel = np.concatenate([np.zeros_like(lo["pos"]), up["vel"]])
    spec["meshes"] = [dict(pos=pos, faces=faces, material=up["material"], fixed=list(lo["fixed"]), vel=vel,
                           coloring="mcs")]
    return spec


def folded_sheet(rows=24, cols=8, h=0.02, gap=0.016, fold_at=0.2):
    """Strip of rows x cols vertices (spacing h), folded over itself along its length: the
    lower layer lies at y = 0, a half circle of diameter gap turns it, and the upper layer runs
    back at y = gap. Row r sits at arc length r h."""
    u = np.arange(rows) * h
    arc = np.pi * gap / 2
    x = np.where(u <= fold_at, u, np.where(u <= fold_at + arc,
                                           fold_at + 0.5 * gap * np.cos((u - fold_at) / (0.5 * gap) - np.pi / 2),
                                           fold_at - (u - fold_at - arc)))
    y = np.where(u <= fold_at, 0.0, np.where(u <= fold_at + arc,
                                             0.5 * gap + 0.5 * gap * np.sin((u - fold_at) / (0.5 * gap) - np.pi / 2),
                                             gap))
    ii, jj = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    pos = np.stack([x[ii.ravel()], y[ii.ravel()], jj.ravel() * h], axis=1)
    ci, cj = np.meshgrid(np.arange(rows - 1), np.arange(cols - 1), indexing="ij")
    v = (cj + ci * cols).ravel()
    f1 = np.stack([v, v + cols + 1, v + 1], axis=1)
    f2 = np.stack([v, v + cols, v + cols + 1], axis=1)
    faces = np.stack([f1, f2], axis=1).reshape(-1, 3).astype(np.int32)
    return pos, faces


def self_contact_spec(steps=20, energy=1, iterations=8):
    """Self-contact of one mesh: a folded strip whose upper layer falls onto its lower layer,
    which rests on the ground (boundary penalty at y = -0.002)."""
    pos, faces = folded_sheet()
    mu, lm = scenes.stvk_lame(1e4, 0.3)
    material = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-5, friction_dynamic=0.1, friction_eps_v=0.01,
                    initialization_type=0)
    vel = np.zeros_like(pos)
    vel[pos[:, 1] > 0.015, 1] = -0.2
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=iterations,
                           world_lower=(-1e7, -0.002, -1e7), handle_collision=1, contact_stiffness=1e3,
                           contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45,
                           contact_energy_type=energy, contact_query_radius=0.02, contact_bvh_rebuild_steps=8)
    return dict(physics=physics, steps=steps,
                meshes=[dict(pos=pos, faces=faces, material=material, fixed=[], vel=vel, coloring="mcs")])


def tet_drop_spec(steps=20, energy=1, stiffness=300.0, density=30.0, iterations=16, second_cube=False, vy=-0.3):
    """Tet bodies with contact on their boundary surface: a soft cube (3^3 cells of 4 cm) falling
    onto a fixed, slightly tilted cloth sheet 1.2 contact radii below; optionally a second cube
    above it (tet-tet contact). With these defaults the cubes come to rest inside the contact
    radius without touching (bottom 7-8 mm above the sheet, 2-7 mm between the cubes); a softer
    contact or fewer iterations leave the per-step solve too far from equilibrium to hold the
    weight (8 iterations settle a cube at 7.3 mm where 128 reach the 9.9 mm equilibrium)."""
    lo, flo = scenes.cloth_sheet(12, 0.6, 0.0, origin=(-0.3, -0.3))
    lo[:, 1] = 0.02 * lo[:, 0]
    m_lo = dict(density=1.0, mu=1e4, lmbd=1e4, bending_stiffness=1e-4, friction_dynamic=0.2, friction_eps_v=0.01)
    cp, ct = scenes.cube_tet_mesh(3, 0.04, origin=(-0.06, 0.0, -0.06))
    cp[:, 1] += 0.02 * cp[:, 0] * (cp[:, 1] < 1e-9) + 0.012  # bottom layer follows the tilt
    mat = dict(density=density, mu=2e4, lmbd=1e5, friction_dynamic=0.2, friction_eps_v=0.01)
    meshes = [dict(pos=lo, faces=flo, material=m_lo, fixed=list(range(len(lo))), vel=None, coloring="mcs"),
              dict(pos=cp, tets=ct, material=mat, fixed=[], vel=np.tile([0.02, vy, 0.0], (len(cp), 1)),
                   coloring="mcs")]
    if second_cube:
        cp2 = cp + np.array([0.01, 0.12 + 0.012, 0.0])
        meshes.append(dict(pos=cp2, tets=ct, material=mat, fixed=[], vel=np.tile([0.0, vy, 0.0], (len(cp), 1)),
                           coloring="mcs"))
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=iterations,
                           world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=stiffness,
                           contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45,
                           contact_energy_type=energy, contact_query_radius=0.02, contact_bvh_rebuild_steps=4)
    return dict(physics=physics, steps=steps, meshes=meshes)


def build_oracle_mixed(oracle, spec, precision=32):
    """Oracle scene of cloth and tet meshes with collision handling."""
    import gaia_amd

    sc = oracle.ClothScene(spec["physics"], precision=precision)
    for m in spec["meshes"]:
        pos = np.asarray(m["pos"], np.float32).astype(np.float64) if precision == 32 else m["pos"]
        vel = None if m["vel"] is None else (np.asarray(m["vel"], np.float32).astype(np.float64)
                                             if precision == 32 else m["vel"])
        if "faces" in m:
            colors = gaia_amd.color_tri_mesh(len(m["pos"]), m["faces"], m["coloring"])
            sc.add_mesh(pos, m["faces"], m["material"], colors, fixed=m["fixed"], vel=vel)
        else:
            colors = gaia_amd.color_tet_mesh(len(m["pos"]), m["tets"], m["coloring"])
            sc.add_tet_mesh(pos, m["tets"], m["material"], colors, fixed=m["fixed"], vel=vel)
    sc.finalize()
    return sc


def twist_spec(n=16, steps=30, iterations=10):
    """The twisting experiment's set-up at small size (S01_Twisting, Parameters.json "Deformers"):
    a horizontal sheet whose two opposite edges are turned by Rotators about the x axis in
    opposite directions (angularVelocity pi/4, rotationEndTime 60)."""
    pos, faces = scenes.cloth_sheet(n, 0.5, 0.0, origin=(-0.25, -0.25))
    left = [i * n for i in range(n)]
    right = [i * n + n - 1 for i in range(n)]
    mu, lm = scenes.stvk_lame(1e5, 0.3)
    material = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4, initialization_type=2)
    physics = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=1.0 / 60.0, num_substeps=2, iterations=iterations,
                           world_lower=(-1e7, -1e7, -1e7))
    rot = lambda axis, sel: dict(DeformerName="Rotator", center=[0, 0, 0], axis=axis,
                                 angularVelocity=np.pi / 4, selectedMeshes=[0], selectedVertices=[sel],
                                 rotationEndTime=60)
    return dict(physics=physics, steps=steps,
                meshes=[dict(pos=pos, faces=faces, material=material, fixed=[], vel=None, coloring="mcs")],
                deformers=[rot([1, 0, 0], right), rot([-1, 0, 0], left)])
