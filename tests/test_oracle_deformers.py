"""Pins of the oracle's deformers (VBD/VBDDeformer.h:6-222, VBDCloth/VBDClothDeformer.h:6-246)
and vertex-face attachments (VBDCloth/VBDClothPhysics.cpp:148-224, 1465-1498).

Deformed vertices join the fixed points and follow prescribed paths, checked against closed forms
evaluated here in numpy: rotation by Rodrigues' formula, linear translation, and the step-wise
cloth translation. Attachments add a spring k b^2 to one vertex's Newton system; with material
constants that vanish one iteration lands exactly on the minimiser of
m/(2 dt^2) |x - y|^2 + k/2 |b (x_read - target)|^2."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_oracle_cloth, twist_spec

DT = float(np.float32(1.0 / 60.0))


def clock(n, dt=DT):
    """The scene time after n substeps: float sums of dt, as the reference's curTime."""
    t = np.float32(0)
    for _ in range(n):
        t = np.float32(t + np.float32(dt))
    return float(t)


def rodrigues(p, axis, theta):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = np.cos(theta), np.sin(theta)
    return p * c + np.cross(a, p) * s + np.outer(p @ a, a) * (1 - c)


def tet_cube(oracle, deformer, precision=64, iterations=4):
    pos, tets = scenes.cube_tet_mesh(3, 0.1, origin=(-0.15, 0.2, -0.15))
    phys = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=iterations,
                        world_lower=(-1e7, -1e7, -1e7))
    sc = oracle.TetScene(phys, precision=precision)
    sc.add_mesh(pos, tets, dict(density=1000.0, mu=2e4, lmbd=1e5), gaia_amd.color_tet_mesh(len(pos), tets, "greedy"))
    sc.add_deformer(deformer)
    sc.finalize()
    return sc, pos


def top_layer(pos):
    return [int(i) for i in np.nonzero(np.isclose(pos[:, 1], pos[:, 1].max()))[0]]


@pytest.mark.parametrize("precision", [32, 64])
def test_tet_rotator_turns_the_selected_vertices(oracle, precision):
    """A Rotator about the y axis through the origin: after n steps the selected vertices are
    the initial ones turned by omega t(n) (Rodrigues), their velocities zero; the rest of the
    cube follows."""
    sel = None
    pos, _ = scenes.cube_tet_mesh(3, 0.1, origin=(-0.15, 0.2, -0.15))
    sel = top_layer(pos)
    w = 2.0
    d = dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 1, 0], angularVelocity=w, selectedMeshes=[0],
             selectedVertices=[sel])
    sc, pos = tet_cube(oracle, d, precision)
    n = 12
    sc.step(n, threads=1)
    x, v = sc.get_state(0)
    np.testing.assert_allclose(x[sel], rodrigues(pos[sel], (0, 1, 0), w * clock(n)), rtol=0, atol=2e-6)
    assert np.all(v[sel] == 0)
    assert abs(sc.time - clock(n)) == 0
    free = [i for i in range(len(pos)) if i not in sel]
    assert np.abs(x[free] - pos[free]).max() > 1e-3  # the body moves under gravity and the twist


def test_rotator_end_time_and_centre_on_the_axis(oracle):
    """rotationEndTime stops the turn at omega * end; a centre on the axis line gives the same
    motion as the origin (root = ((P - o) . a) a, RootP = P - root)."""
    pos, _ = scenes.cube_tet_mesh(3, 0.1, origin=(-0.15, 0.2, -0.15))
    sel = top_layer(pos)
    base = dict(DeformerName="Rotator", axis=[0, 1, 0], angularVelocity=3.0, selectedMeshes=[0],
                selectedVertices=[sel], rotationEndTime=0.1)
    a, _ = tet_cube(oracle, dict(base, center=[0, 0, 0]))
    b, _ = tet_cube(oracle, dict(base, center=[0, 0.7, 0]))
    a.step(10, threads=1)
    b.step(10, threads=1)
    xa, xb = a.get_state(0)[0], b.get_state(0)[0]
    np.testing.assert_allclose(xa[sel], rodrigues(pos[sel], (0, 1, 0), 3.0 * 0.1), rtol=0, atol=2e-6)
    np.testing.assert_allclose(xa[sel], xb[sel], rtol=0, atol=1e-12)


def test_tet_translator_is_linear_in_time_until_the_end(oracle):
    pos, _ = scenes.cube_tet_mesh(3, 0.1, origin=(-0.15, 0.2, -0.15))
    sel = top_layer(pos)
    speed = np.array([0.3, 0.0, -0.2])
    d = dict(DeformerName="Translator", speed=list(speed), selectedMeshes=[0], selectedVertices=[sel],
             deformationEndTime=0.1)
    sc, pos = tet_cube(oracle, d)
    for n in (3, 12):
        sc.step(3 if n == 3 else 9, threads=1)
        x, _ = sc.get_state(0)
        t = min(clock(n), float(np.float32(0.1)))
        np.testing.assert_allclose(x[sel], pos[sel] + speed * t, rtol=0, atol=1e-6)


def cloth_strip(oracle, deformers, precision=64, num_substeps=1, attachments=None, **phys):
    pos, faces = scenes.cloth_sheet(6, 0.5, 0.0)
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT * num_substeps, num_substeps=num_substeps,
                     iterations=4, world_lower=(-1e7, -1e7, -1e7))
    p.update(phys)
    mu, lm = scenes.stvk_lame(1e4, 0.3)
    spec = dict(physics=p, meshes=[dict(pos=pos, faces=faces, material=dict(density=0.2, mu=mu, lmbd=lm),
                                        fixed=[], vel=None, coloring="mcs")], deformers=deformers)
    if attachments:
        spec["attachments"] = attachments
    return build_oracle_cloth(oracle, spec, precision), pos


def test_cloth_translator_moves_step_wise_inside_its_window(oracle):
    """The cloth Translator adds speed * dt in every substep whose time t satisfies
    start <= t < end (VBDClothDeformer.h:170-186)."""
    sel = [0, 1, 2]
    speed = np.array([0.0, 0.5, 0.25])
    start, end = 0.05, 0.12
    d = dict(DeformerName="Translator", speed=list(speed), selectedMeshes=[0], selectedVertices=[sel],
             deformationStartTime=start, deformationEndTime=end)
    sc, pos = cloth_strip(oracle, [d], num_substeps=2)
    sc.step(6, threads=1)  # 12 substeps
    x, v = sc.get_state(0)
    f32 = np.float32
    moves = sum(1 for n in range(1, 13) if f32(clock(n)) >= f32(start) and f32(clock(n)) < f32(end))
    assert 0 < moves < 12
    np.testing.assert_allclose(x[sel], pos[sel] + moves * speed * DT, rtol=0, atol=1e-9)
    assert np.all(v[sel] == 0)


def test_cloth_rotator_matches_rodrigues(oracle):
    sel = [0, 5, 30, 35]
    d = dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 0, 1], angularVelocity=1.5, selectedMeshes=[0],
             selectedVertices=[sel])
    sc, pos = cloth_strip(oracle, [d])
    sc.step(8, threads=1)
    x, _ = sc.get_state(0)
    np.testing.assert_allclose(x[sel], rodrigues(pos[sel], (0, 0, 1), 1.5 * clock(8)), rtol=0, atol=2e-6)


def test_twist_fp32_follows_fp64(oracle):
    """The twisting set-up: FP32 and FP64 oracles agree on the twisted sheet (a well-posed
    prescribed-boundary problem)."""
    spec = twist_spec(n=10, steps=10)
    a = build_oracle_cloth(oracle, spec, 32)
    b = build_oracle_cloth(oracle, spec, 64)
    a.step(10, threads=1)
    b.step(10, threads=1)
    xa, xb = a.get_state(0)[0], b.get_state(0)[0]
    assert np.abs(xa - xb).max() < 1e-4
    left = [i * 10 for i in range(10)]
    # the left edge turned by -pi/4 * t about +x (axis (-1, 0, 0))
    np.testing.assert_allclose(xb[left], rodrigues(spec["meshes"][0]["pos"][left], (-1, 0, 0),
                                                         np.pi / 4 * clock(20, np.float32(np.float32(1 / 60) / 2))),
                               rtol=0, atol=2e-6)


def test_deformer_arguments(oracle):
    sc = oracle.TetScene(physics_dict(), precision=64)
    pos, tets = scenes.cube_tet_mesh(2, 0.1)
    sc.add_mesh(pos, tets, dict(), gaia_amd.color_tet_mesh(len(pos), tets, "greedy"))
    with pytest.raises(ValueError):
        sc.add_deformer(dict(DeformerName="Spinner", selectedMeshes=[0], selectedVertices=[[0]]))
    with pytest.raises(ValueError):
        sc.add_deformer(dict(DeformerName="Translator", speed=[1, 0, 0], selectedMeshes=[0], selectedVertices=[[99]]))


# --------------------------------------------------------------------------- attachments

def lone_vertices(oracle, attachments, fixed=(), iterations=1, precision=64):
    """Two 2 x 2 cloth patches with vanishing material constants and no gravity: every vertex
    solve is its own inertia plus attachment springs, so one iteration solves it exactly."""
    pa, fa = scenes.cloth_sheet(2, 0.1, 0.0)
    pb, fb = scenes.cloth_sheet(2, 0.1, 0.3, origin=(0.5, 0.0))
    vel = np.tile([0.2, -0.1, 0.05], (4, 1))
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=iterations,
                     world_lower=(-1e7, -1e7, -1e7))
    mat = dict(density=0.2, mu=1e-12, lmbd=1e-12, initialization_type=2)
    spec = dict(physics=p, attachments=attachments,
                meshes=[dict(pos=pa, faces=fa, material=mat, fixed=list(fixed[0]) if fixed else [], vel=vel,
                             coloring="mcs"),
                        dict(pos=pb, faces=fb, material=mat, fixed=list(fixed[1]) if fixed else [], vel=vel,
                             coloring="mcs")])
    sc = build_oracle_cloth(oracle, spec, precision)
    return sc, pa, pb, vel


def lumped_mass(pos, faces, density=0.2):
    m = np.zeros(len(pos))
    for f in faces:
        a = 0.5 * np.linalg.norm(np.cross(pos[f[1]] - pos[f[0]], pos[f[2]] - pos[f[0]]))
        m[f] += density * a / 3
    return m


def test_attachment_to_a_fixed_point(oracle):
    """Vertex 3 of mesh 0 tied to the point p: x = (m/dt^2 y + k p) / (m/dt^2 + k)."""
    k, point = 50.0, np.array([0.3, 0.2, -0.1])
    att = dict(vertexConstraintVertexInfos=[0, 3], vertexConstraintTargetFacesInfos=[-1, 0, 0, 0],
               vertexConstraintBarys=[list(point)], vertexConstraintStiffness=[k])
    sc, pa, _, vel = lone_vertices(oracle, att)
    sc.step(1, threads=1)
    x, _ = sc.get_state(0)
    _, fa = scenes.cloth_sheet(2, 0.1, 0.0)
    m = lumped_mass(pa, fa)[3]
    y = pa[3] + DT * vel[3]
    c = m / DT ** 2
    p32 = point.astype(np.float32).astype(np.float64)  # the reference stores the point in float
    np.testing.assert_allclose(x[3], (c * y + k * p32) / (c + k), rtol=0, atol=1e-10)
    np.testing.assert_allclose(x[0], pa[0] + DT * vel[0], rtol=0, atol=1e-10)  # no record: its inertia


def test_attachment_default_stiffness_and_face_side(oracle):
    """Vertex 1 of mesh 0 tied to face (0, 1, 2) of mesh 1 with barycentrics (0.2, 0.3, 0.5), no
    stiffness given (1e4). With mesh 1 fixed the vertex side lands on
    (c y + k target) / (c + k). With mesh 0 fixed instead, each face vertex j with b_j != 0 gets
    f = -k b_j (x_read - target), H = k b_j^2, where x_read is mesh 0's vertex with the face
    vertex's index (VBDClothPhysics.cpp:1471)."""
    bary = np.array([0.2, 0.3, 0.5])
    att = dict(vertexConstraintVertexInfos=[0, 1], vertexConstraintTargetFacesInfos=[1, 0, 1, 2],
               vertexConstraintBarys=[list(bary)])
    k = 1e4
    _, fa = scenes.cloth_sheet(2, 0.1, 0.0)
    sc, pa, pb, vel = lone_vertices(oracle, att, fixed=([], [0, 1, 2, 3]))
    sc.step(1, threads=1)
    x, _ = sc.get_state(0)
    m = lumped_mass(pa, fa)
    c = m[1] / DT ** 2
    target = bary.astype(np.float32).astype(np.float64) @ pb[[0, 1, 2]]
    y = pa[1] + DT * vel[1]
    np.testing.assert_allclose(x[1], (c * y + k * target) / (c + k), rtol=0, atol=1e-10)

    # face side, one face vertex at a time (b = (1, 0, 0)): the target is that vertex itself
    att1 = dict(att, vertexConstraintBarys=[[1.0, 0.0, 0.0]])
    sc, pa, pb, vel = lone_vertices(oracle, att1, fixed=([0, 1, 2, 3], []))
    sc.step(1, threads=1)
    xb, _ = sc.get_state(1)
    mb = lumped_mass(pb, fa)
    y0 = pb[0] + DT * vel[0]
    c0 = mb[0] / DT ** 2
    # Newton on f(x) = c (y - x) - k (-1) (x_read - x) with x_read = mesh 0's vertex 0 (fixed)
    np.testing.assert_allclose(xb[0], (c0 * y0 + k * pa[0]) / (c0 + k), rtol=0, atol=1e-10)
    for j in (1, 2, 3):  # zero barycentrics: no record, pure inertia
        np.testing.assert_allclose(xb[j], pb[j] + DT * vel[j], rtol=0, atol=1e-10)


def test_attachment_arguments(oracle):
    for bad in (dict(vertexConstraintVertexInfos=[2, 0], vertexConstraintTargetFacesInfos=[-1, 0, 0, 0],
                     vertexConstraintBarys=[[0, 0, 0]]),
                dict(vertexConstraintVertexInfos=[0, 0], vertexConstraintTargetFacesInfos=[1, 0, 1, 9],
                     vertexConstraintBarys=[[0.3, 0.3, 0.4]])):
        with pytest.raises(ValueError):
            lone_vertices(oracle, bad)
    with pytest.raises(ValueError):
        lone_vertices(oracle, dict(vertexConstraintVertexInfos=[0, 0, 0, 1], vertexConstraintTargetFacesInfos=[-1, 0, 0, 0],
                                   vertexConstraintBarys=[[0, 0, 0]]))


def test_reference_experiment_deformers_parse():
    """The "Deformers" arrays of the reference's twisting (S01) and knot-pulling (S02) parameter
    files (fixture written by tests/golden/make_deformer_fixtures.py) read as the set-ups built
    here: S01 turns the two x-extreme columns of its 50 x 50 sheet about +x and -x at pi/4 rad/s
    until t = 60 (the layout of twist_spec), S02 pulls two vertex sets apart at 10 m/s."""
    import json
    import os

    from oracle.pyoracle import deformer_args

    path = os.path.join(os.path.dirname(__file__), "golden", "deformers", "reference_deformers.json")
    ref = json.load(open(path))
    s01 = [deformer_args(d) for d in ref["S01_Twisting"]]
    assert [k for k, *_ in s01] == [0, 0]
    n = 50
    right, left = [i * n + n - 1 for i in range(n)], [i * n for i in range(n)]
    for (kind, prm, meshes, counts, verts), axis, sel in zip(s01, ([1, 0, 0], [-1, 0, 0]), (right, left)):
        assert list(meshes) == [0] and list(counts) == [n] and sorted(verts.tolist()) == sel
        np.testing.assert_allclose(prm[3:6], axis)
        np.testing.assert_allclose(prm[6], np.pi / 4)
        assert prm[7] == 60
    s02 = [deformer_args(d) for d in ref["S02_KnotPulling"]]
    assert [k for k, *_ in s02] == [1, 1]
    np.testing.assert_allclose(s02[0][1][8:11], [-10, 0, 0])
    np.testing.assert_allclose(s02[1][1][8:11], [10, 0, 0])
    for kind, prm, meshes, counts, verts in s02:
        assert prm[7] == -1 and prm[11] == 0  # no end time: the pull runs on; starts at 0
        assert len(verts) == counts.sum() > 0
