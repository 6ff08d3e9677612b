"""GPU parity of deformers (VBD/VBDDeformer.h:6-222, VBDCloth/VBDClothDeformer.h:6-246) and
vertex-face attachments (VBDCloth/VBDClothPhysics.cpp:148-224, 1465-1498): the HIP path through
the C ABI against the FP32 oracle per step from identical state (1e-5 normwise-relative), the
deformed vertices against closed forms, and the JSON loader, batches and the scene time."""
import json

import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import (build_gpu, build_oracle, build_oracle_cloth, c1_spec, contact_drape_spec, rel_err,
                       twist_spec)

pytestmark = pytest.mark.gpu
TOL = 1e-5
DT = float(np.float32(1.0 / 60.0))


def clock(n, dt=DT):
    t = np.float32(0)
    for _ in range(n):
        t = np.float32(t + np.float32(dt))
    return float(t)


def rodrigues(p, axis, theta):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    c, s = np.cos(theta), np.sin(theta)
    return p * c + np.cross(a, p) * s + np.outer(p @ a, a) * (1 - c)


def lockstep(gpu_ctx, spec, osc, steps):
    """Per step from the oracle's state: worst normwise-relative position error."""
    sim = build_gpu(gpu_ctx, spec)
    nm = len(spec["meshes"])
    errs = []
    for _ in range(steps):
        for m in range(nm):
            sim.setState(m, *osc.get_state(m))
        sim.runStep()
        osc.step(1, threads=8)
        errs.append(max(rel_err(sim.getState(m)[0], osc.get_state(m)[0]) for m in range(nm)))
    t = sim.curTime
    sim.destroy()
    return np.array(errs), t


def test_twist_per_step_parity(gpu_ctx, oracle):
    """The twisting set-up (two Rotators on opposite edges of a sheet, 2 substeps)."""
    spec = twist_spec(n=16, steps=30)
    osc = build_oracle_cloth(oracle, spec)
    errs, t = lockstep(gpu_ctx, spec, osc, 30)
    print(f"\n[twist] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: {errs.max():.3e}"
    assert t == osc.time == clock(60, np.float32(np.float32(1 / 60) / 2))


def test_tet_rotator_per_step_parity_and_closed_form(gpu_ctx, oracle):
    spec = c1_spec()
    pos = spec["meshes"][0]["pos"]
    top = [int(i) for i in np.nonzero(np.isclose(pos[:, 1], pos[:, 1].max()))[0]]
    spec["deformers"] = [dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 1, 0], angularVelocity=2.0,
                              selectedMeshes=[0], selectedVertices=[top], rotationEndTime=0.3)]
    spec["meshes"][0]["fixed"] = []
    osc = build_oracle(oracle, spec)
    errs, _ = lockstep(gpu_ctx, spec, osc, 25)
    print(f"\n[tet rotator] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL
    # free run on the device: the top layer turned by min(t, 0.3) * 2
    sim = build_gpu(gpu_ctx, spec)
    sim.runSteps(25)
    x, v = sim.getState(0)
    p32 = np.asarray(pos, np.float32).astype(np.float64)
    np.testing.assert_allclose(x[top], rodrigues(p32[top], (0, 1, 0), 2.0 * min(clock(25), 0.3)), rtol=0,
                               atol=2e-6)
    assert np.all(v[top] == 0)
    sim.destroy()


def small_tet(ctx, i, deformer):
    pos, tets = scenes.cube_tet_mesh(3, 0.1, origin=(0.0, 0.3 + 0.1 * i, 0.0))
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=6,
                                        world_lower=(-1e7, 0.0, -1e7)))
    sim.addTetMesh(pos, tets, density=1000.0, mu=1e5, lmbd=1e6)
    if deformer:
        sim.addDeformer(deformer)
    sim.initialize()
    return sim, pos


def test_translator_in_a_batch_equals_alone(gpu_ctx):
    """Scenes with and without deformers in one batch (graph replay): each as when stepped alone;
    the tet Translator follows p0 + speed min(t, end)."""
    speed = [0.2, 0.0, -0.1]
    ds = [dict(DeformerName="Translator", speed=speed, selectedMeshes=[0], selectedVertices=[[0, 1, 2]],
               deformationEndTime=0.05), None,
          dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 0, 1], angularVelocity=1.0, selectedMeshes=[0],
               selectedVertices=[[5, 6]])]
    alone = [small_tet(gpu_ctx, i, d) for i, d in enumerate(ds)]
    mixed = [small_tet(gpu_ctx, i, d) for i, d in enumerate(ds)]
    for s, _ in alone:
        s.runSteps(6)
    batch = Batch(gpu_ctx, [s for s, _ in mixed])
    batch.record_graph(True)
    batch.run_steps(6)
    for (a, pos), (m, _) in zip(alone, mixed):
        xa, va = a.getState(0)
        xm, vm = m.getState(0)
        assert np.array_equal(xa, xm) and np.array_equal(va, vm)
        assert a.curTime == m.curTime == clock(6)
    x, _ = alone[0][0].getState(0)
    p32 = np.asarray(alone[0][1], np.float32).astype(np.float64)
    np.testing.assert_allclose(x[:3], p32[:3] + np.array(speed) * float(np.float32(0.05)), rtol=0, atol=1e-6)
    batch.destroy()
    for s, _ in alone + mixed:
        s.destroy()


def test_cloth_translator_window_and_time(gpu_ctx, tmp_path):
    pos, faces = scenes.cloth_sheet(6, 0.5, 0.0)
    sel = [0, 1, 2]
    speed = np.array([0.0, 0.5, 0.25])
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=2 * DT, num_substeps=2, iterations=4,
                                        world_lower=(-1e7, -1e7, -1e7)))
    mu, lm = scenes.stvk_lame(1e4, 0.3)
    sim.addClothMesh(pos, faces, density=0.2, mu=mu, lmbd=lm)
    sim.addDeformer(dict(DeformerName="Translator", speed=list(speed), selectedMeshes=[0], selectedVertices=[sel],
                         deformationStartTime=0.05, deformationEndTime=0.12))
    sim.initialize()
    sim.runSteps(6)
    x, v = sim.getState(0)
    f32 = np.float32
    dt = float(f32(f32(2 * DT) / 2))
    moves = sum(1 for n in range(1, 13) if f32(clock(n, dt)) >= f32(0.05) and f32(clock(n, dt)) < f32(0.12))
    p32 = np.asarray(pos, np.float32).astype(np.float64)
    np.testing.assert_allclose(x[sel], p32[sel] + moves * speed * dt, rtol=0, atol=1e-6)
    assert np.all(v[sel] == 0)
    assert sim.curTime == clock(12, dt)
    # the recovery state carries the time; recovering sets it
    sim.saveState(str(tmp_path / "s.json"), frame_id=6)
    sim.curTime = 0.0
    assert sim.curTime == 0.0
    sim.recoverFromState(str(tmp_path / "s.json"))
    assert sim.curTime == clock(12, dt)
    sim.destroy()


def lone(ctx, attachments, fixed):
    pa, fa = scenes.cloth_sheet(2, 0.1, 0.0)
    pb, fb = scenes.cloth_sheet(2, 0.1, 0.3, origin=(0.5, 0.0))
    vel = np.tile([0.2, -0.1, 0.05], (4, 1))
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=1,
                                        world_lower=(-1e7, -1e7, -1e7)))
    for p, f, fx in ((pa, fa, fixed[0]), (pb, fb, fixed[1])):
        sim.addClothMesh(p, f, fixed_points=fx, velocities=vel, density=0.2, mu=1e-12, lmbd=1e-12,
                         initialization_type=2)
    sim.addAttachments(attachments)
    sim.initialize()
    return sim, pa, pb, vel, fa


def lumped_mass(pos, faces, density=0.2):
    m = np.zeros(len(pos))
    for f in faces:
        m[f] += density * 0.5 * np.linalg.norm(np.cross(pos[f[1]] - pos[f[0]], pos[f[2]] - pos[f[0]])) / 3
    return m


def test_attachment_closed_forms_on_the_device(gpu_ctx):
    """A fixed-point attachment and the face side (b = -1, x_read from the vertex-side mesh at the
    face vertex's index): one iteration lands on (c y + k p) / (c + k), c = m / dt^2."""
    k, point = 50.0, np.array([0.3, 0.2, -0.1])
    sim, pa, _, vel, fa = lone(gpu_ctx, dict(vertexConstraintVertexInfos=[0, 3],
                                             vertexConstraintTargetFacesInfos=[-1, 0, 0, 0],
                                             vertexConstraintBarys=[list(point)], vertexConstraintStiffness=[k]),
                               ([], []))
    sim.runStep()
    x, _ = sim.getState(0)
    c = lumped_mass(pa, fa)[3] / DT ** 2
    y = pa[3] + DT * vel[3]
    np.testing.assert_allclose(x[3], (c * y + k * point) / (c + k), rtol=1e-5, atol=1e-6)
    sim.destroy()
    sim, pa, pb, vel, fa = lone(gpu_ctx, dict(vertexConstraintVertexInfos=[0, 1],
                                              vertexConstraintTargetFacesInfos=[1, 0, 1, 2],
                                              vertexConstraintBarys=[[1.0, 0.0, 0.0]]), ([0, 1, 2, 3], []))
    sim.runStep()
    xb, _ = sim.getState(1)
    c0 = lumped_mass(pb, fa)[0] / DT ** 2
    y0 = pb[0] + DT * vel[0]
    np.testing.assert_allclose(xb[0], (c0 * y0 + 1e4 * pa[0]) / (c0 + 1e4), rtol=1e-5, atol=1e-6)
    sim.destroy()


def hanging_pair_spec():
    """A curtain hanging from its top row with a second, free sheet tied below it: every vertex of
    the free sheet's top row is attached to a face of the curtain's bottom strip (face side
    active), and one corner to a fixed point."""
    n = 10
    pc, fc, top = scenes.cloth_curtain(n, 0.5, 1.0)
    ps, fs, _ = scenes.cloth_curtain(n, 0.5, 0.45)
    mu, lm = scenes.stvk_lame(1e5, 0.3)
    mat = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4, initialization_type=2)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                           world_lower=(-1e7, -1e7, -1e7))
    vi, tf, bary = [], [], []
    for j in range(n - 1):
        v = j  # top row of the free sheet
        a = (n - 2) * n + j  # bottom strip of the curtain: cell (n - 2, j)
        vi += [1, v]
        tf += [0, a, a + 1, a + n + 1]
        bary.append([0.5, 0.3, 0.2])
    vi += [1, n * n - 1]
    tf += [-1, 0, 0, 0]
    bary.append([0.5, 0.0, 0.0])
    att = dict(vertexConstraintVertexInfos=vi, vertexConstraintTargetFacesInfos=tf, vertexConstraintBarys=bary,
               vertexConstraintStiffness=[2e3] * (n - 1) + [50.0])
    return dict(physics=physics, steps=30, attachments=att,
                meshes=[dict(pos=pc, faces=fc, material=mat, fixed=top, vel=None, coloring="mcs"),
                        dict(pos=ps, faces=fs, material=mat, fixed=[], vel=None, coloring="mcs")])


def test_attachments_per_step_parity(gpu_ctx, oracle):
    spec = hanging_pair_spec()
    osc = build_oracle_cloth(oracle, spec)
    errs, _ = lockstep(gpu_ctx, spec, osc, 30)
    print(f"\n[attachments] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: {errs.max():.3e}"


def test_contact_scene_with_a_cloth_deformer(gpu_ctx, oracle):
    """The contact frame: the cloth deformer moves its vertices before detection at iteration 0.
    One step of this drape (step 7) lands within rounding of a contact switch, where the FP32 and
    FP64 oracles already part by ~5e-5 from the same state; per step the device stays within 1e-5
    of the FP32 oracle or within 5x the oracles' own FP32/FP64 distance on that step."""
    spec = contact_drape_spec(steps=15)
    spec["deformers"] = [dict(DeformerName="Translator", speed=[0.1, 0.0, 0.0], selectedMeshes=[1],
                              selectedVertices=[[0, 9]])]
    sim = build_gpu(gpu_ctx, spec)
    osc = build_oracle_cloth(oracle, spec)
    o64 = build_oracle_cloth(oracle, spec, 64)
    errs, errs64 = [], []
    for _ in range(15):
        for m in range(2):
            sim.setState(m, *osc.get_state(m))
            o64.set_state(m, *osc.get_state(m))
        o64.time = osc.time
        sim.runStep()
        osc.step(1, threads=8)
        o64.step(1, threads=8)
        errs.append(max(rel_err(sim.getState(m)[0], osc.get_state(m)[0]) for m in range(2)))
        errs64.append(max(rel_err(osc.get_state(m)[0], o64.get_state(m)[0]) for m in range(2)))
    sim.destroy()
    errs, errs64 = np.array(errs), np.array(errs64)
    print(f"\n[contact + deformer] per-step rel err {np.array2string(errs, precision=2)}\n"
          f"  fp32 vs fp64 {np.array2string(errs64, precision=2)}", flush=True)
    assert np.all(errs < np.maximum(TOL, 5 * errs64)), f"step {np.argmax(errs / np.maximum(TOL, 5 * errs64))}"
    assert np.sum(errs >= TOL) <= 1


def test_json_loader_reads_deformers_and_attachments(gpu_ctx, tmp_path):
    """Parameters.json "Deformers" and PhysicsParams.VertexFaceAttachmentConstraintParams load as
    the programmatic scene: identical states after 5 steps."""
    spec = hanging_pair_spec()
    spec["deformers"] = [dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 1, 0], angularVelocity=0.5,
                              selectedMeshes=[0], selectedVertices=[[0, 9]], rotationEndTime=-1)]
    models = {"Models": []}
    for i, m in enumerate(spec["meshes"]):
        path = tmp_path / f"m{i}.obj"
        with open(path, "w") as f:
            for p in m["pos"]:
                f.write(f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
            for t in m["faces"]:
                f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        models["Models"].append(dict(materialName="StVK_triMesh", path=str(path), density=0.2, miu=m["material"]["mu"], **{"lambda": m["material"]["lmbd"]},
                                     bendingStiffness=1e-4, initializationType=2, fixedPoints=list(m["fixed"])))
    params = {"PhysicsParams": dict(gravity=[0, -9.8, 0], timeStep=1.0 / 60.0, numSubsteps=1, iterations=10,
                                    worldBounds=[[-1e7, -1e7, -1e7], [1e7, 1e7, 1e7]], handleCollision=False,
                                    VertexFaceAttachmentConstraintParams=spec["attachments"]),
              "Deformers": spec["deformers"]}
    json.dump(models, open(tmp_path / "Models.json", "w"))
    json.dump(params, open(tmp_path / "Parameters.json", "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(str(tmp_path / "Models.json"), str(tmp_path / "Parameters.json"))
    a.initialize()
    b = build_gpu(gpu_ctx, spec)
    for m in range(2):  # the .obj round trip rounds positions: start both from the loaded state
        b.setState(m, *a.getState(m))
    a.runSteps(5)
    b.runSteps(5)
    for m in range(2):
        assert rel_err(a.getState(m)[0], b.getState(m)[0]) < 1e-6
    a.destroy()
    b.destroy()


def test_argument_errors(gpu_ctx):
    pos, faces = scenes.cloth_sheet(3, 0.1, 0.0)
    for bad_def in (dict(DeformerName="Translator", speed=[1, 0, 0], selectedMeshes=[1], selectedVertices=[[0]]),
                    dict(DeformerName="Translator", speed=[1, 0, 0], selectedMeshes=[0], selectedVertices=[[99]])):
        sim = VBDPhysics(gpu_ctx)
        sim.addClothMesh(pos, faces)
        sim.addDeformer(bad_def)
        with pytest.raises(gaia_amd.VBDError, match="deformer"):
            sim.initialize()
    for bad_att in (dict(vertexConstraintVertexInfos=[0, 0], vertexConstraintTargetFacesInfos=[0, 0, 1, 50],
                         vertexConstraintBarys=[[0.2, 0.3, 0.5]]),
                    dict(vertexConstraintVertexInfos=[3, 0], vertexConstraintTargetFacesInfos=[-1, 0, 0, 0],
                         vertexConstraintBarys=[[0, 0, 0]])):
        sim = VBDPhysics(gpu_ctx)
        sim.addClothMesh(pos, faces)
        sim.addAttachments(bad_att)
        with pytest.raises(gaia_amd.VBDError, match="attachment"):
            sim.initialize()


def test_vertex_selected_by_two_deformers_is_rejected(gpu_ctx):
    """Deformed vertices of one phase move in one launch, so a vertex may belong to one deformer
    only (the reference applies deformers one after another in list order)."""
    pos, faces = scenes.cloth_sheet(4, 0.1, 0.0)
    sim = VBDPhysics(gpu_ctx)
    sim.addClothMesh(pos, faces)
    sim.addDeformer(dict(DeformerName="Translator", speed=[1, 0, 0], selectedMeshes=[0], selectedVertices=[[0, 1]]))
    sim.addDeformer(dict(DeformerName="Translator", speed=[0, 1, 0], selectedMeshes=[0], selectedVertices=[[2, 1]]))
    with pytest.raises(gaia_amd.VBDError, match="already selected by a deformer"):
        sim.initialize()


def test_tet_deformer_with_collision_handling_is_rejected(gpu_ctx):
    """A tet deformer moves vertices after the frame's first detection, so contact bounds would not
    see the motion: such scenes are refused (cloth deformers, which move before detection, are not)."""
    from vbd_cases import tet_drop_spec

    spec = tet_drop_spec(steps=1)
    spec["deformers"] = [dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 1, 0], angularVelocity=1.0,
                              selectedMeshes=[1], selectedVertices=[[0, 1]], rotationEndTime=10)]
    with pytest.raises(gaia_amd.VBDError, match="deformers of tet meshes"):
        build_gpu(gpu_ctx, spec)


def test_served_deformer_scenes_equal_alone(gpu_ctx):
    """Scenes with deformers in the serving loop under a budget that forces copy-out between
    frames: the scene time and the deformed vertices travel with the arena, so each served scene
    equals the scene stepped alone."""
    from gaia_amd import Server

    rot = dict(DeformerName="Rotator", center=[0, 0, 0], axis=[0, 1, 0], angularVelocity=1.5, selectedMeshes=[0],
               selectedVertices=[[0, 1, 2, 3]])
    ds = [rot, None, dict(rot, axis=[1, 0, 0]), None]

    def make(i, upload):
        pos, tets = scenes.cube_tet_mesh(4 + i, 0.05, origin=(0.0, 0.3, 0.0))
        sim = VBDPhysics(gpu_ctx)
        sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=5,
                                            world_lower=(-1e7, 0.0, -1e7)))
        sim.addTetMesh(pos, tets, density=1000.0, mu=1e5, lmbd=1e6)
        if ds[i]:
            sim.addDeformer(ds[i])
        sim.initialize(defer_upload=not upload)
        return sim

    before = gpu_ctx.memory_bytes()
    probe = make(3, True)
    biggest = gpu_ctx.memory_bytes() - before
    probe.destroy()
    sims = [make(i, False) for i in range(4)]
    srv = Server(gpu_ctx, int(1.5 * biggest))
    for s in sims:
        srv.add_scene(s)
    rng = np.random.default_rng(2)
    counts = np.zeros(4, int)
    for _ in range(12):
        for i in rng.integers(0, 4, size=2):
            srv.request_step(sims[i])
            counts[i] += 1
        srv.tick()
    while srv.stats()["pending"]:
        srv.tick()
    assert srv.stats()["evictions"] > 0
    for i in range(4):
        alone = make(i, True)
        alone.runSteps(int(counts[i]))
        xa, va = alone.getState(0)
        xs, vs = sims[i].getState(0)
        assert np.array_equal(xa, xs) and np.array_equal(va, vs), f"scene {i}"
        assert alone.curTime == sims[i].curTime == clock(int(counts[i]))
        alone.destroy()
    srv.destroy()
    for s in sims:
        s.destroy()
