"""Seeded synthetic scene generators for the benchmark configurations.

Every generator is deterministic given its arguments (and seed). Outputs are
numpy arrays: positions (N, 3) float32/float64, tets (T, 4) int32.
"""
import json
import os

import numpy as np


def cube_tet_mesh(n=10, h=0.1, origin=(0.0, 0.0, 0.0), dtype=np.float64):
    """Regular n^3 hex grid with spacing h, each hex split into 5 tets.

    Vertex (i, j, k) has index i + (n+1) * (j + (n+1) * k). In hex (i, j, k) the
    central tet joins the four corners of even global parity (i+j+k+a+b+c even),
    and each odd corner forms a corner tet with its three edge neighbours. Using
    global parity makes the split alternate from hex to hex, so shared faces get
    the same diagonal (a conforming mesh). Tets are oriented to positive volume.
    At n = 10: 1331 vertices, 5000 tets, incidence 15.03 on average, 32 at most.
    """
    n1 = n + 1
    kk, jj, ii = np.meshgrid(np.arange(n1), np.arange(n1), np.arange(n1), indexing="ij")
    pos = np.stack([origin[0] + ii.ravel() * h, origin[1] + jj.ravel() * h, origin[2] + kk.ravel() * h],
                   axis=1).astype(dtype)
    # hexes in (k, j, i) order, i fastest; per hex: central tet, then the four corner tets
    hk, hj, hi = [a.ravel() for a in np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")]
    par = (hi + hj + hk) % 2

    def corner(a, b, c):
        return (hi + a) + n1 * ((hj + b) + n1 * (hk + c))

    def templates(p):
        cs = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
        even = sorted(x for x in cs if (p + sum(x)) % 2 == 0)
        odd = sorted(x for x in cs if (p + sum(x)) % 2 == 1)
        out = [even]
        for o in odd:
            out.append([o, (o[0] ^ 1, o[1], o[2]), (o[0], o[1] ^ 1, o[2]), (o[0], o[1], o[2] ^ 1)])
        return out

    tets = np.zeros((n ** 3, 5, 4), dtype=np.int64)
    for p in (0, 1):
        sel = par == p
        for t, tmpl in enumerate(templates(p)):
            for v, (a, b, c) in enumerate(tmpl):
                tets[sel, t, v] = corner(a, b, c)[sel]
    tets = tets.reshape(-1, 4).astype(np.int32)
    # orient to positive volume
    p = pos.astype(np.float64)
    d = np.einsum("ij,ij->i", np.cross(p[tets[:, 1]] - p[tets[:, 0]], p[tets[:, 2]] - p[tets[:, 0]]),
                  p[tets[:, 3]] - p[tets[:, 0]])
    neg = d < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return pos, tets


def c1_soft_cube():
    """Plain VBD soft cube: 10x10x10 cells of 0.1 m (1331 vertices, 5000 tets), bottom at y = 0.5."""
    pos, tets = cube_tet_mesh(10, 0.1, origin=(-0.5, 0.5, -0.5))
    physics = dict(
        gravity=(0.0, -9.8, 0.0),
        time_step=1.0 / 60.0,
        num_substeps=1,
        iterations=10,
        world_lower=(-1e7, 0.0, -1e7),
        world_upper=(1e7, 1e7, 1e7),
    )
    material = dict(density=1000.0, mu=1e5, lmbd=1e6)
    return dict(positions=pos, tets=tets, physics=physics, material=material, coloring="greedy", steps=300)


def cloth_sheet(n, size=1.0, height=1.0, origin=(0.0, 0.0), dtype=np.float64):
    """n x n vertex grid in the horizontal plane y = height, tessellated as the reference's
    grid generator does (Simulator/VBDCloth/ParameterGen/DataGen/S01_GenCloth.py, writeOBj):
    vertex (i, j) = j + i n with x along j and z along i; each cell (i, j) gets faces
    (v, v+1, v+n+1) and (v, v+n+1, v+n)."""
    xs = np.linspace(origin[0], origin[0] + size, n)
    zs = np.linspace(origin[1], origin[1] + size, n)
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([xs[jj.ravel()], np.full(n * n, height), zs[ii.ravel()]], axis=1).astype(dtype)
    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v = (cj + ci * n).ravel()
    f1 = np.stack([v, v + 1, v + n + 1], axis=1)
    f2 = np.stack([v, v + n + 1, v + n], axis=1)
    faces = np.stack([f1, f2], axis=1).reshape(-1, 3).astype(np.int32)
    return pos, faces


def cloth_curtain(n, size=1.0, top=2.0, dtype=np.float64):
    """n x n vertical sheet in the plane z = 0 hanging from its top row (vertices 0..n-1 at
    y = top), same tessellation as cloth_sheet; returns (positions, faces, top_row)."""
    pos, faces = cloth_sheet(n, size, 0.0, dtype=np.float64)
    pos = np.stack([pos[:, 0], top - pos[:, 2], np.zeros(len(pos))], axis=1).astype(dtype)
    return pos, faces[:, [0, 2, 1]].copy(), list(range(n))


def stvk_lame(young, poisson):
    """Membrane (plane-stress) Lame parameters of a sheet from Young's modulus and Poisson ratio:
    mu = E / (2 (1 + nu)), lambda = E nu / (1 - nu^2)."""
    return young / (2.0 * (1.0 + poisson)), young * poisson / (1.0 - poisson * poisson)


C4_PHYSICS = dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                  world_lower=(-1e7, 0.0, -1e7), world_upper=(1e7, 1e7, 1e7))


def c4_scenes(seed=0, num_scenes=256, vmin=500, vmax=50000, cloth_fraction=0.5, poisson=0.3):
    """Batched-VBD workload: num_scenes independent scenes, vertex counts log-uniform in
    [vmin, vmax], each (seeded) either a tet cube or a cloth sheet.

    Tet cube: n^3 cells of 5 cm, (n+1)^3 close to the target, 10 cm above the ground,
    stable neo-Hookean mu = 1e5, lambda = 1e6, rho = 1000, greedy colouring.
    Cloth sheet: N x N vertices (N^2 close to the target), 1 m square at 1.5 m height, pinned
    at two corners, StVK from E = 1e5 with nu = 0.3 (plane-stress Lame parameters), bending
    stiffness 1e-4, rho = 0.2 kg/m^2, MCS colouring of the vertex graph with bending edges.
    dt = 1/60 s, 10 iterations."""
    rng = np.random.default_rng(seed)
    targets = np.exp(rng.uniform(np.log(vmin), np.log(vmax), size=num_scenes))
    kinds = rng.random(num_scenes) < cloth_fraction
    mu_c, lm_c = stvk_lame(1e5, poisson)
    out = []
    for target, is_cloth in zip(targets, kinds):
        if is_cloth:
            n = max(3, int(round(np.sqrt(target))))
            pos, faces = cloth_sheet(n, 1.0, 1.5, dtype=np.float32)
            out.append(dict(kind="cloth", positions=pos, faces=faces, physics=dict(C4_PHYSICS),
                            fixed=[0, n - 1],
                            material=dict(density=0.2, mu=mu_c, lmbd=lm_c, bending_stiffness=1e-4),
                            coloring="mcs"))
        else:
            n = max(2, int(round(target ** (1.0 / 3.0))) - 1)
            pos, tets = cube_tet_mesh(n, 0.05, origin=(0.0, 0.1, 0.0), dtype=np.float32)
            out.append(dict(kind="tet", positions=pos, tets=tets, physics=dict(C4_PHYSICS), fixed=[],
                            material=dict(density=1000.0, mu=1e5, lmbd=1e6), coloring="greedy"))
    return out


def zipf_size_classes(num_scenes, vmin=1000, vmax=200000, exponent=1.2, classes=32, seed=0):
    """Vertex-count targets of a served population: 32 size classes spaced log-uniformly from
    vmin to vmax; class rank k (1 = smallest) drawn with probability proportional to k^-exponent
    (seeded Zipf over the ranks). The mean is about 9.3k vertices for the defaults."""
    rng = np.random.default_rng(seed)
    k = np.arange(1, classes + 1)
    w = k ** -float(exponent)
    ranks = rng.choice(classes, size=num_scenes, p=w / w.sum())
    sizes = vmin * (vmax / vmin) ** (np.arange(classes) / (classes - 1))
    return sizes[ranks]


def c5_scene_kinds(num_scenes=2000, vmin=1000, vmax=200000, exponent=1.2, cloth_fraction=0.5, seed=0):
    """Served-VBD population: (kind, grid size) per scene, with the batched workload's generators
    (tet cube of n^3 cells, or an N x N cloth sheet pinned at two corners) at Zipf-drawn vertex
    targets. Scenes of one (kind, grid size) share their mesh; each has its own state."""
    sizes = zipf_size_classes(num_scenes, vmin, vmax, exponent, seed=seed)
    rng = np.random.default_rng(seed + 1)
    cloth = rng.random(num_scenes) < cloth_fraction
    out = []
    for target, is_cloth in zip(sizes, cloth):
        if is_cloth:
            out.append(("cloth", max(3, int(round(np.sqrt(target))))))
        else:
            out.append(("tet", max(2, int(round(target ** (1.0 / 3.0))) - 1)))
    return out


def c4_template(kind, n):
    """One scene of the batched workload's generators (see c4_scenes) for a given grid size."""
    mu_c, lm_c = stvk_lame(1e5, 0.3)
    if kind == "cloth":
        pos, faces = cloth_sheet(n, 1.0, 1.5, dtype=np.float32)
        return dict(kind="cloth", positions=pos, faces=faces, physics=dict(C4_PHYSICS), fixed=[0, n - 1],
                    material=dict(density=0.2, mu=mu_c, lmbd=lm_c, bending_stiffness=1e-4), coloring="mcs")
    pos, tets = cube_tet_mesh(n, 0.05, origin=(0.0, 0.1, 0.0), dtype=np.float32)
    return dict(kind="tet", positions=pos, tets=tets, physics=dict(C4_PHYSICS), fixed=[],
                material=dict(density=1000.0, mu=1e5, lmbd=1e6), coloring="greedy")


def c4_tet_scene_sizes(seed=0, num_scenes=256, vmin=500, vmax=50000):
    """Vertex-count targets drawn log-uniform in [vmin, vmax] (seeded)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(vmin), np.log(vmax), size=num_scenes))


def c4_tet_scenes(seed=0, num_scenes=256):
    """Batched-VBD workload, tet-cube part: per scene a cube of n^3 cells with
    (n+1)^3 close to a log-uniform vertex target in [500, 50k], 5 cm cells, resting
    10 cm above the ground, stable neo-Hookean mu=1e5, lambda=1e6, rho=1000,
    dt = 1/60 s, 10 iterations, greedy colouring."""
    out = []
    for i, target in enumerate(c4_tet_scene_sizes(seed, num_scenes)):
        n = max(2, int(round(target ** (1.0 / 3.0))) - 1)
        pos, tets = cube_tet_mesh(n, 0.05, origin=(0.0, 0.1, 0.0), dtype=np.float32)
        physics = dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                       world_lower=(-1e7, 0.0, -1e7), world_upper=(1e7, 1e7, 1e7))
        out.append(dict(kind="tet", positions=pos, tets=tets, physics=physics,
                        material=dict(density=1000.0, mu=1e5, lmbd=1e6), coloring="greedy"))
    return out


def write_t_file(path, pos, tets):
    """MeshFrame .t text format ("Vertex i x y z", "Tet i a b c d")."""
    with open(path, "w") as f:
        for i, p in enumerate(pos):
            f.write(f"Vertex {i} {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        for i, t in enumerate(tets):
            f.write(f"Tet {i} {t[0]} {t[1]} {t[2]} {t[3]}\n")


def write_scene_json(folder, name, pos, tets, physics, material, colors=None, translation=(0.0, 0.0, 0.0)):
    """Writes <name>.t, optional colouring, Models.json and Parameters.json in the reference schema
    (Materials/Materials.h:55-75, VBD/VBD_NeoHookean.cpp:1613-1625, Parameters/PhysicsParameters.h:107-178)."""
    os.makedirs(folder, exist_ok=True)
    tpath = os.path.join(folder, name + ".t")
    write_t_file(tpath, pos, tets)
    model = {
        "path": "${REPO_ROOT}/" + name + ".t",
        "materialName": "NeoHookean",
        "density": material.get("density", 1.0),
        "miu": material.get("mu", 1e4),
        "lmbd": material.get("lmbd", 1e4),
        "translation": list(translation),
        "rotation": [0.0, 0.0, 0.0],
        "scale": [1.0, 1.0, 1.0],
        "initialVelocity": [0.0, 0.0, 0.0],
    }
    if colors is not None:
        cats = [np.flatnonzero(colors == c).tolist() for c in range(int(colors.max()) + 1)]
        cpath = tpath + ".vertexColoring.json"
        with open(cpath, "w") as f:
            json.dump(cats, f)
        model["verticesColoringCategoriesPath"] = model["path"] + ".vertexColoring.json"
    p = physics
    params = {
        "PhysicsParams": {
            "gravity": list(p.get("gravity", (0.0, -10.0, 0.0))),
            "timeStep": p.get("time_step", 1.0 / 60.0),
            "numSubsteps": p.get("num_substeps", 20),
            "iterations": p.get("iterations", 1),
            "worldBounds": [list(p.get("world_lower", (-1e7,) * 3)), list(p.get("world_upper", (1e7,) * 3))],
            "usePlaneGround": True,
            "boundaryFrictionDynamic": p.get("boundary_friction_dynamic", 0.1),
        },
        "CollisionParams": {"allowDCD": False, "allowCCD": False},
    }
    mpath = os.path.join(folder, "Models.json")
    ppath = os.path.join(folder, "Parameters.json")
    with open(mpath, "w") as f:
        json.dump({"Models": [model]}, f, indent=1)
    with open(ppath, "w") as f:
        json.dump(params, f, indent=1)
    return mpath, ppath


def torus(nu=128, nv=64, major=0.08, minor=0.03, dtype=np.float64):
    """Closed torus triangle mesh: vertex (i, j) = i nv + j at angles u_i = 2 pi i / nu (around
    the axis z) and v_j = 2 pi j / nv (around the tube); every quad (i, j) (i+1, j) (i+1, j+1)
    (i, j+1) is split into two consistently oriented triangles. nu nv vertices, 2 nu nv faces,
    3 nu nv edges."""
    u = 2.0 * np.pi * np.arange(nu) / nu
    v = 2.0 * np.pi * np.arange(nv) / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    ring = major + minor * np.cos(V)
    pos = np.stack([ring * np.cos(U), ring * np.sin(U), minor * np.sin(V)], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.stack([np.stack([a, b, c], -1), np.stack([a, c, d], -1)], axis=2).reshape(-1, 3)
    return pos.astype(dtype), faces.astype(np.int32)


def random_rotation(rng):
    """Rotation from three seeded angles, each uniform in [0, 2 pi) (z-y-x Euler)."""
    a, b, c = rng.uniform(0.0, 2.0 * np.pi, 3)
    rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rx = np.array([[1, 0, 0], [0, np.cos(c), -np.sin(c)], [0, np.sin(c), np.cos(c)]])
    return rz @ ry @ rx


def tet_ball(n=12, radius=0.15, dtype=np.float64):
    """Tet ball: the n^3-cell tet cube on [-1, 1]^3 mapped onto the ball by
    (x, y, z) -> (x sqrt(1 - y^2/2 - z^2/2 + y^2 z^2/3), and cyclically), a smooth bijection of the
    cube onto the ball whose Jacobian stays positive, so every tet keeps a positive volume.
    (n+1)^3 vertices and 5 n^3 tets; the boundary is the mapped cube surface."""
    pos, tets = cube_tet_mesh(n, 2.0 / n, origin=(-1.0, -1.0, -1.0))
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    x2, y2, z2 = x * x, y * y, z * z
    ball = np.stack([x * np.sqrt(1 - y2 / 2 - z2 / 2 + y2 * z2 / 3),
                     y * np.sqrt(1 - z2 / 2 - x2 / 2 + z2 * x2 / 3),
                     z * np.sqrt(1 - x2 / 2 - y2 / 2 + x2 * y2 / 3)], axis=1)
    return (radius * ball).astype(dtype), tets


def c2_bodies(seed=0, n=12, cube_size=0.28, ball_radius=0.14, spacing=0.5, layers_y=(0.3, 0.8)):
    """Tet bodies in contact: 32 bodies on a 4 x 4 x 2 lattice (spacing 0.5 m, layers at 0.3 m and
    0.8 m) in a 2 m box, alternating 16 subdivided cubes (n^3 cells, edge cube_size) and 16 tet
    balls (tet_ball(n), radius ball_radius), each with 5 n^3 tets (8,640 at n = 12) and a seeded
    rotation; shear modulus log-uniform in [1e4, 1e6] Pa, lambda = 10 mu, density 1000 kg/m^3,
    contact friction 0.2. The sizes keep every rotated body clear of its neighbours.
    Returns a list of dicts (pos, tets, material)."""
    rng = np.random.default_rng(seed)
    cube, cube_tets = cube_tet_mesh(n, cube_size / n, origin=(-cube_size / 2,) * 3)
    ball, ball_tets = tet_ball(n, ball_radius)
    bodies = []
    coords = [(-0.75 + spacing * i, y, -0.75 + spacing * k) for y in layers_y for k in range(4) for i in range(4)]
    for b, centre in enumerate(coords):
        i, k, layer = b % 4, (b // 4) % 4, b // 16
        is_cube = (i + k + layer) % 2 == 0
        base, tets = (cube, cube_tets) if is_cube else (ball, ball_tets)
        rot = random_rotation(rng)
        mu = float(10.0 ** rng.uniform(4.0, 6.0))
        pos = base @ rot.T + np.asarray(centre)
        # a rotation is proper, so tet orientation (and positive volume) is kept
        bodies.append(dict(pos=pos, tets=tets, kind="cube" if is_cube else "ball",
                           material=dict(density=1000.0, mu=mu, lmbd=10.0 * mu, friction_dynamic=0.2,
                                         friction_eps_v=0.01)))
    return bodies


def c2_physics(iterations=20, query_radius=0.003):
    """Physics of the tet-contact workload: dt = 1/120, one substep, contact radius 2 mm,
    contact stiffness 1e4 (the cloth framework's default, VBDClothPhysicsParameters.h:61) with the
    quadratic-log energy, query radius 1.5 contact radii (the ratio of the reference's cloth
    scenes, e.g. contactRadius 0.2 / maxQueryDis 0.3 in
    VBDCloth/ParameterGen/Parameters/S01_Twisting/*/Parameters.json), BVH rebuilt every 32 steps
    (contactBVHReconstructionSteps' default, refit between), box boundary of 2 m."""
    return dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 120.0, num_substeps=1, iterations=iterations,
                world_lower=(-1.0, 0.0, -1.0), world_upper=(1.0, 2.0, 1.0), handle_collision=1,
                contact_stiffness=1e4, contact_radius=0.002, thickness=0.0, conservative_step_relaxation=0.45,
                contact_energy_type=1, contact_query_radius=query_radius, contact_bvh_rebuild_steps=32)


def c3_tori(seed=0, grid=4, nu=128, nv=64, spacing=0.3, major=0.12, minor=0.04, drift=0.005):
    """Collision-detection workload: grid^3 tori (default 64 x 16,384 triangles, edges of about
    4 mm) on a lattice with `spacing` between centres (neighbouring tori of outer radius 0.16 m
    interpenetrate),
    each with a seeded random rotation and a seeded random drift direction of `drift` metres
    per step. Returns (meshes [(positions, faces)], drift vectors (M, 3))."""
    rng = np.random.default_rng(seed)
    base, faces = torus(nu, nv, major, minor)
    meshes, drifts = [], []
    for ix in range(grid):
        for iy in range(grid):
            for iz in range(grid):
                R = random_rotation(rng)
                centre = spacing * np.array([ix, iy, iz], np.float64)
                meshes.append(((base @ R.T + centre).astype(np.float32), faces))
                d = rng.normal(size=3)
                drifts.append(drift * d / np.linalg.norm(d))
    return meshes, np.asarray(drifts, np.float32)
