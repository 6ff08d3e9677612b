"""GPU parity of cloth contact (VBDClothPhysics.cpp:1223-1399 repulsion and friction in the
vertex solve, 1500-1573 conservative-bound truncation, ClothContactDetector queries): the HIP
path through the C ABI vs the FP32 oracle, per step from identical state, within 1e-5
normwise-relative."""
import json

import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import (build_gpu, build_oracle_cloth, contact_drape_spec, curtain_spec, rel_err,
                       self_contact_spec, single_mesh)

pytestmark = pytest.mark.gpu
TOL = 1e-5
THREADS = 8
DT = float(np.float32(1.0 / 60.0))


def contact_parity(gpu_ctx, oracle, spec, with_fp64=False):
    """Per step from identical state: the HIP path's error against the FP32 oracle, and (when
    asked) the FP32 oracle's own distance from the FP64 oracle on the same step."""
    sim = build_gpu(gpu_ctx, spec)
    osc = build_oracle_cloth(oracle, spec)
    o64 = build_oracle_cloth(oracle, spec, 64) if with_fp64 else None
    nm = len(spec["meshes"])
    errs, errs64 = [], []
    x0 = [osc.get_state(m)[0].copy() for m in range(nm)]
    for k in range(spec["steps"]):
        for m in range(nm):
            sim.setState(m, *osc.get_state(m))
            if o64 is not None:
                o64.set_state(m, *osc.get_state(m))
        sim.runStep()
        osc.step(1, threads=THREADS)
        if o64 is not None:
            o64.step(1, threads=THREADS)
            errs64.append(max(rel_err(osc.get_state(m)[0], o64.get_state(m)[0]) for m in range(nm)))
        errs.append(max(rel_err(sim.getState(m)[0], osc.get_state(m)[0]) for m in range(nm)))
        for m in range(nm):
            fixed = spec["meshes"][m]["fixed"]
            if len(fixed):
                assert np.all(sim.getState(m)[1][list(fixed)] == 0)
    moved = max(np.abs(osc.get_state(m)[0] - x0[m]).max() for m in range(nm))
    sim.destroy()
    return np.array(errs), np.array(errs64), moved


# Contact energies switch their Hessian on at d = contact radius (and, for the log barrier, at
# tau): the reference's vertex Newton step is discontinuous there, so two FP32 evaluations of one
# step part wherever a vertex lands within rounding of a switch. Stiff contact (k >> m / dt^2)
# makes that common; these configurations keep every step well-conditioned (the FP32 and FP64
# oracles agree to ~1e-6 on every step), so the 1e-5 per-step contract is checked step by step.
@pytest.mark.parametrize("energy,density", [(0, 3.0), (1, 1.0), (2, 1.0)])
def test_drape_on_fixed_sheet_per_step_parity(gpu_ctx, oracle, energy, density):
    errs, _, moved = contact_parity(gpu_ctx, oracle, contact_drape_spec(energy=energy, density=density))
    print(f"\n[contact drape, energy {energy}] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: rel err {errs.max():.3e}"
    assert moved > 0.001


def test_single_mesh_self_contact_per_step_parity(gpu_ctx, oracle):
    """Both sheets as components of one mesh: vertex-face and edge-edge pairs within a mesh,
    with the pairs that share a vertex excluded."""
    errs, _, moved = contact_parity(gpu_ctx, oracle, single_mesh(contact_drape_spec()))
    print(f"\n[self contact, one mesh] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: rel err {errs.max():.3e}"
    assert moved > 0.001


@pytest.mark.parametrize("case", ["stiff_drape", "fold"])
def test_ill_conditioned_contact_error_distribution(gpu_ctx, oracle, case):
    """Stiff contact (k / (m / dt^2) ~ 1e3) and a folded strip in self-contact: steps are often
    ill-conditioned (the FP32 and FP64 oracles differ by more than 1e-5 on many steps). The HIP
    path must then sit as close to the FP32 oracle as the FP32 oracle sits to the FP64 one: the
    median per-step error at most that of FP32 vs FP64 (a systematic error would show on every
    step)."""
    spec = (contact_drape_spec(steps=30, energy=0, density=0.2, stiffness=1e3) if case == "stiff_drape"
            else self_contact_spec(steps=30))
    errs, errs64, moved = contact_parity(gpu_ctx, oracle, spec, with_fp64=True)
    print(f"\n[contact {case}] per-step rel err gpu-vs-fp32 median {np.median(errs):.2e} max {errs.max():.2e}; "
          f"fp32-vs-fp64 median {np.median(errs64):.2e} max {errs64.max():.2e}", flush=True)
    assert np.median(errs) <= max(2.0 * np.median(errs64), 1e-6)
    assert moved > 0.001


def closed_form_scene(gpu_ctx, meshes, iterations=1, **phys):
    """GPU counterpart of the oracle's closed-form contact pins: every vertex its own colour."""
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=iterations,
                     world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=100.0,
                     contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45, contact_energy_type=0,
                     contact_query_radius=0.02)
    p.update(phys)
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(**p)
    for m in meshes:
        n = len(m["pos"])
        sim.addClothMesh(np.asarray(m["pos"], np.float64), m["faces"], fixed_points=m.get("fixed", ()),
                         vertex_colors=np.arange(n, dtype=np.int32), velocities=m.get("vel"), **m["material"])
    sim.initialize()
    return sim


BIG = dict(pos=[[-10.0, 0.0, -10.0], [10.0, 0.0, -10.0], [0.0, 0.0, 10.0]], faces=np.array([[0, 2, 1]], np.int32),
           material=dict(density=1.0, mu=0.0, lmbd=0.0), fixed=[0, 1, 2])
DENSITY = 16666.666666666668
M_INERTIA = DENSITY * 0.5 * 0.01 * 0.01 / 3.0 / DT ** 2


def small_triangle(d, vel=None, **mat):
    m = dict(density=DENSITY, mu=0.0, lmbd=0.0, bending_stiffness=0.0, initialization_type=2)
    m.update(mat)
    return dict(pos=[[0.0, d, 0.0], [0.01, d, 0.0], [0.0, d, 0.01]], faces=np.array([[0, 2, 1]], np.int32),
                material=m, vel=vel)


def test_closed_form_vf_contact_steps(gpu_ctx):
    """One free triangle over a large fixed one (only VF contacts): the quadratic energy's
    Newton step lands on (k r + M d) / (M + k), the log branch (d < r / 2) moves by
    (k2 / d) / (M + k2 / d^2), a stiff contact is cut to 0.45 d, and energies beyond the
    contact radius do nothing."""
    d, k, r = 0.004, 100.0, 0.01
    sim = closed_form_scene(gpu_ctx, [BIG, small_triangle(d)], iterations=3)
    sim.runStep()
    np.testing.assert_allclose(sim.getState(1)[0][:, 1], (k * r + M_INERTIA * d) / (M_INERTIA + k), rtol=2e-6)
    sim.destroy()
    d = 0.003
    sim = closed_form_scene(gpu_ctx, [BIG, small_triangle(d)], contact_energy_type=1)
    sim.runStep()
    k2 = 0.5 * (0.5 * r) ** 2 * k
    np.testing.assert_allclose(sim.getState(1)[0][:, 1], d + (k2 / d) / (M_INERTIA + k2 / d ** 2), rtol=2e-6)
    sim.destroy()
    d = 0.004
    sim = closed_form_scene(gpu_ctx, [BIG, small_triangle(d)], contact_stiffness=1e5)
    sim.runStep()
    f32 = np.float32
    np.testing.assert_allclose(sim.getState(1)[0][:, 1], d + float(f32(0.45) * f32(d)), rtol=2e-6)
    sim.destroy()
    sim = closed_form_scene(gpu_ctx, [BIG, small_triangle(0.015)], iterations=2)
    sim.runStep()
    np.testing.assert_array_equal(sim.getState(1)[0][:, 1], np.float32(0.015))
    sim.destroy()


def test_closed_form_friction_step(gpu_ctx):
    """Sliding at vx over the fixed triangle: the tangential step is -mu lambda / (M + mu lambda / u)."""
    d, k, r, vx, mu_f = 0.004, 100.0, 0.01, 0.05, 0.3
    big = dict(BIG, material=dict(BIG["material"], friction_dynamic=mu_f, friction_eps_v=0.01))
    sm = small_triangle(d, vel=np.tile([vx, 0.0, 0.0], (3, 1)), initialization_type=0, friction_dynamic=mu_f,
                        friction_eps_v=0.01)
    sim = closed_form_scene(gpu_ctx, [big, sm])
    sim.runStep()
    x, _ = sim.getState(1)
    lam, u = k * (r - d), DT * vx
    want_x = np.array(sm["pos"])[:, 0] + DT * vx - mu_f * lam / (M_INERTIA + mu_f * lam / u)
    # normwise 1e-5 of the positions: the FP32 oracle itself sits 4e-8 (3e-4 of the friction
    # displacement) from this closed form
    want = np.stack([want_x, np.full(3, (k * r + M_INERTIA * d) / (M_INERTIA + k))], axis=1)
    assert rel_err(x[:, :2], want) < TOL
    sim.destroy()


def test_closed_form_ee_contact_step(gpu_ctx):
    """Crossing edges at distance d (EE contacts only): the free endpoint (weight 1/2) moves down
    by 0.5 k (r - d) / (M + 0.25 k)."""
    d, k, r, density = 0.004, 100.0, 0.01, 3000.0
    A = dict(pos=[[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0]], faces=np.array([[0, 1, 2]], np.int32),
             material=dict(density=density, mu=0.0, lmbd=0.0), fixed=[0, 2])
    B = dict(pos=[[0.0, d, -1.0], [0.0, d, 1.0], [0.0, 1.0 + d, 0.0]], faces=np.array([[0, 2, 1]], np.int32),
             material=dict(density=1.0, mu=0.0, lmbd=0.0), fixed=[0, 1, 2])
    sim = closed_form_scene(gpu_ctx, [A, B])
    sim.runStep()
    x, _ = sim.getState(0)
    M = density * 1.0 / 3.0 / DT ** 2
    np.testing.assert_allclose(x[1], [1.0, -0.5 * k * (r - d) / (M + 0.25 * k), 0.0], rtol=2e-6, atol=1e-9)
    sim.destroy()


def test_drape_rests_on_the_lower_sheet(gpu_ctx):
    """Over 90 steps the upper sheet is caught by the lower one: every free vertex above the
    fixed sheet ends above its tilted plane (the log barrier of energy type 1 keeps d > 0), and
    the sheet has come down onto it (free fall would take it 0.4 m lower)."""
    spec = contact_drape_spec(steps=90)
    sim = build_gpu(gpu_ctx, spec)
    sim.runSteps(90)
    x, v = sim.getState(1)
    gap = x[:, 1] - 0.02 * x[:, 0]
    inside = np.all(np.abs(x[:, [0, 2]]) < 0.29, axis=1)
    print(f"\n[contact drape] gap after 90 steps min {gap[inside].min():.4e} median {np.median(gap[inside]):.4e}",
          flush=True)
    assert inside.sum() > 50
    assert gap[inside].min() > 0.0
    assert np.median(gap[inside]) < 0.02
    sim.destroy()


def test_contact_steps_are_deterministic(gpu_ctx):
    """Two copies of a contact scene step to bitwise the same state: contact sums are
    accumulated in a fixed order (no atomics) and detection sets are exact."""
    spec = contact_drape_spec(steps=3)
    a = build_gpu(gpu_ctx, spec)
    b = build_gpu(gpu_ctx, spec)
    for _ in range(3):
        a.runStep()
        b.runStep()
    xa, va = a.getState(1)
    xb, vb = b.getState(1)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    a.destroy()
    b.destroy()


def test_contact_json_loader(gpu_ctx, tmp_path):
    """A reference-schema scene with handleCollision and contact parameters (cloth defaults to
    collision handling) == the same scene built through the API."""
    spec = contact_drape_spec(steps=4)
    models = []
    for i, m in enumerate(spec["meshes"]):
        with open(tmp_path / f"m{i}.obj", "w") as f:
            for p in m["pos"]:
                f.write(f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
            for t in m["faces"]:
                f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
        mat = m["material"]
        d = {"path": "${REPO_ROOT}/" + f"m{i}.obj", "materialName": "StVK_triMesh", "density": mat["density"],
             "miu": mat["mu"], "lambda": mat["lmbd"], "bendingStiffness": mat["bending_stiffness"],
             "frictionDynamic": mat["friction_dynamic"], "frictionEpsV": mat["friction_eps_v"],
             "initializationType": mat.get("initialization_type", 2), "fixedPoints": list(m["fixed"])}
        if m["vel"] is not None:
            d["initialVelocity"] = [float(c) for c in m["vel"][0]]
        models.append(d)
    ph = spec["physics"]
    params = {"PhysicsParams": {"gravity": list(ph["gravity"]), "timeStep": ph["time_step"], "numSubsteps": 1,
                                "iterations": ph["iterations"], "contactStiffness": ph["contact_stiffness"],
                                "contactRadius": ph["contact_radius"], "thickness": 0.0,
                                "conservativeStepRelaxation": 0.45, "contactEnergyType": 0,
                                "contactBVHReconstructionSteps": ph["contact_bvh_rebuild_steps"],
                                "ContactDetectorParams": {"maxQueryDis": ph["contact_query_radius"]},
                                "worldBounds": [[-1e7, -1e7, -1e7], [1e7, 1e7, 1e7]]}}
    json.dump({"Models": models}, open(tmp_path / "Models.json", "w"))
    json.dump(params, open(tmp_path / "Parameters.json", "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(str(tmp_path / "Models.json"), str(tmp_path / "Parameters.json"),
                            repo_root=str(tmp_path))
    a.initialize()
    # contactEnergyType is not read from the parameter file (the reference's parser skips it):
    # the loaded scene keeps the default energy type, so the API scene uses that too
    spec["physics"]["contact_energy_type"] = gaia_amd.vbd.DEFAULT_PHYSICS["contact_energy_type"]
    b = build_gpu(gpu_ctx, spec)
    for _ in range(spec["steps"]):
        a.runStep()
        b.runStep()
    assert rel_err(a.getState(1)[0], b.getState(1)[0]) < 1e-6
    a.destroy()
    b.destroy()


def _frames(sim, n):
    out = []
    for _ in range(n):
        sim.runStep()
        out.append([sim.getState(m) for m in range(sim.numMeshes())])
    return out


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


def test_contact_graph_replay_equals_host_driven_bitwise(gpu_ctx):
    """A cloth-model contact scene captured in one HIP graph per frame, its re-detection after a
    truncating iteration decided on the device (a gate read by every kernel of the pass), equals
    the same scene stepped with the host reading the flag per iteration: positions and velocities
    bitwise on every frame, and the same number of detection passes; also across frames that
    rebuild the trees (every 4 frames here, launched directly)."""
    spec = contact_drape_spec(steps=1)
    spec["physics"]["contact_bvh_rebuild_steps"] = 4
    a = build_gpu(gpu_ctx, spec)
    b = build_gpu(gpu_ctx, spec)
    b.recordVBDSolveGraph(True)
    fa, fb = _frames(a, 14), _frames(b, 14)
    for k, (x, y) in enumerate(zip(fa, fb)):
        assert _same(x, y), f"frame {k}"
    sa, sb = a.contactStats(), b.contactStats()
    print(f"\n[contact graph] 14 frames bitwise; detection passes host {sa['detections']} device {sb['detections']}, "
          f"rebuilds {sa['rebuilds']} / {sb['rebuilds']}", flush=True)
    assert sa["detections"] == sb["detections"] > 14
    assert sa["rebuilds"] == sb["rebuilds"] == 3
    a.destroy()
    b.destroy()


def test_contact_scenes_in_a_batch_equal_alone_bitwise(gpu_ctx):
    """Two cloth-model contact scenes, a curtain without contact and a tet-only penetration-contact
    scene in one batch (launched directly and as a graph): every scene equals itself stepped
    alone, bitwise."""
    from vbd_cases import penetration_spec

    specs = [contact_drape_spec(steps=1, iterations=10),
             contact_drape_spec(steps=1, stiffness=60.0, friction=0.4, iterations=10), curtain_spec(n=6),
             penetration_spec(steps=0, bodies=2)]
    assert {sp["physics"]["iterations"] for sp in specs} == {10}
    frames = 9
    ref = []
    for sp in specs:
        sim = build_gpu(gpu_ctx, sp)
        ref.append(_frames(sim, frames)[-1])
        sim.destroy()
    for graph in (False, True):
        sims = [build_gpu(gpu_ctx, sp) for sp in specs]
        batch = Batch(gpu_ctx, sims)
        batch.record_graph(graph)
        batch.run_steps(frames)
        for i, s in enumerate(sims):
            got = [s.getState(m) for m in range(s.numMeshes())]
            assert _same(got, ref[i]), f"scene {i}, graph {graph}"
        batch.destroy()
        for s in sims:
            s.destroy()


def test_served_contact_scenes_equal_alone_under_eviction(gpu_ctx):
    """Cloth-model contact scenes served beside plain ones under a budget of their detectors plus
    about half of the arenas: idle scenes are copied out and back (detectors resident), the
    re-detection decided on the device inside the batched frames, and every served scene equals
    the same scene stepped alone for as many frames, bitwise."""
    from gaia_amd import Server

    specs = [contact_drape_spec(steps=1, stiffness=30.0 + 10.0 * k, iterations=8) for k in range(4)]
    specs += [curtain_spec(n=6)]
    for sp in specs:
        sp["physics"]["iterations"] = 8
    sims = [build_gpu(gpu_ctx, sp) for sp in specs]
    sizes = [s.deviceBytes() for s in sims]
    arenas = [b["arena"] for b in sizes]
    budget = int(sum(b["detector"] for b in sizes) + max(0.5 * sum(arenas), 1.1 * max(arenas)))
    srv = Server(gpu_ctx, budget)
    for s in sims:
        srv.add_scene(s)
    rng = np.random.default_rng(3)
    counts = np.zeros(len(sims), int)
    tag = 0
    for _ in range(30):
        for i in rng.integers(0, len(sims), size=rng.integers(1, 3)):
            srv.request_step(sims[i], tag)
            counts[i] += 1
            tag += 1
        srv.tick()
    for _ in range(1000):
        if not srv.stats()["pending"]:
            break
        srv.tick()
    st = srv.stats()
    done = srv.poll(capacity=tag + 16)
    assert st["pending"] == 0 and len(done) == tag and all(d["frame"] >= 0 for d in done)
    print(f"\n[served cloth contact] {tag} requests, evictions {st['evictions']}", flush=True)
    assert st["evictions"] > 0, st
    got = []
    for s in sims:
        srv.remove_scene(s)
        got.append([s.getState(m) for m in range(s.numMeshes())])
    srv.destroy()
    for i, sp in enumerate(specs):
        alone = build_gpu(gpu_ctx, sp)
        alone.runSteps(int(counts[i]))
        assert _same(got[i], [alone.getState(m) for m in range(alone.numMeshes())]), f"scene {i}"
        alone.destroy()
    for s in sims:
        s.destroy()


def test_contact_scene_with_tet_mesh_accepted(gpu_ctx):
    """Cloth and tet meshes in one scene with collision handling: the tet mesh takes part in
    contact through its boundary surface (parity in test_gpu_tet_contact.py)."""
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(handle_collision=1, num_substeps=1, iterations=2)
    pos, faces = scenes.cloth_sheet(4, 0.5, 1.0)
    sim.addClothMesh(pos, faces)
    p, t = scenes.cube_tet_mesh(2, 0.1, origin=(0.0, 0.5, 0.0))
    sim.addTetMesh(p, t)
    sim.initialize()
    sim.runSteps(2)
    x, _ = sim.getState(1)
    assert np.isfinite(x).all()
    sim.destroy()
