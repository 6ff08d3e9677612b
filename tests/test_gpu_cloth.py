"""GPU parity of the cloth VBD hot path (StVK + bending): HIP kernels through the C ABI vs
the FP32 oracle, per step from identical state, within 1e-5 normwise-relative."""
import os

import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from vbd_cases import (build_gpu, build_oracle_cloth, cloth_drop_spec, curtain_spec, rel_err,
                       two_corner_sheet_spec)

pytestmark = pytest.mark.gpu
TOL = 1e-5
THREADS = 8


def per_step_parity(gpu_ctx, oracle, spec):
    sim = build_gpu(gpu_ctx, spec)
    osc = build_oracle_cloth(oracle, spec)
    for a, b in zip(sim.parallelGroups(), osc.groups()):
        assert np.array_equal(a, b)
    nm = len(spec["meshes"])
    worst = 0.0
    for k in range(spec["steps"]):
        for m in range(nm):
            sim.setState(m, *osc.get_state(m))
        sim.runStep()
        osc.step(1, threads=THREADS)
        for m in range(nm):
            x, v = sim.getState(m)
            xr, _ = osc.get_state(m)
            e = rel_err(x, xr)
            worst = max(worst, e)
            assert e < TOL, f"step {k} mesh {m}: rel err {e:.3e}"
            fixed = spec["meshes"][m]["fixed"]
            if len(fixed):
                assert np.all(v[list(fixed)] == 0)
    sim.destroy()
    return worst


@pytest.mark.parametrize("init", [2, 0])
def test_curtain_per_step_parity(gpu_ctx, oracle, init):
    w = per_step_parity(gpu_ctx, oracle, curtain_spec(init=init))
    print(f"\n[cloth curtain init {init}] worst per-step rel err {w:.3e}", flush=True)


def test_two_corner_sheet_first_step_parity(gpu_ctx, oracle):
    """The batched workload's sheet (pinned at two corners), first step from rest."""
    w = per_step_parity(gpu_ctx, oracle, two_corner_sheet_spec())
    print(f"\n[cloth two-corner sheet] first-step rel err {w:.3e}", flush=True)


def test_two_sheet_drop_per_step_parity(gpu_ctx, oracle):
    w = per_step_parity(gpu_ctx, oracle, cloth_drop_spec())
    print(f"\n[cloth drop] worst per-step rel err {w:.3e}", flush=True)


def test_curtain_trajectory(gpu_ctx, oracle):
    spec = curtain_spec(steps=60, init=0)
    sim = build_gpu(gpu_ctx, spec)
    sim.recordVBDSolveGraph(True)
    o32 = build_oracle_cloth(oracle, spec, 32)
    o64 = build_oracle_cloth(oracle, spec, 64)
    for _ in range(spec["steps"]):
        sim.runStep()
        o32.step(1, threads=THREADS)
        o64.step(1, threads=THREADS)
    x, _ = sim.getState(0)
    x32, _ = o32.get_state(0)
    x64, _ = o64.get_state(0)
    e_gpu, e_64 = rel_err(x, x32), rel_err(x32, x64)
    print(f"\n[cloth curtain] 60-step drift gpu-vs-oracle32 {e_gpu:.3e}, oracle32-vs-oracle64 {e_64:.3e}", flush=True)
    assert e_gpu < 1e-4 and e_64 < 1e-4
    assert x[:, 1].min() < 0.99  # the curtain has stretched under its weight
    sim.destroy()


@pytest.mark.parametrize("lpv,block", [(1, 128), (4, 256), (16, 256)])
def test_cloth_schedules_agree(gpu_ctx, oracle, lpv, block):
    spec = curtain_spec(n=16, steps=5)
    sim = build_gpu(gpu_ctx, spec)
    sim.setSchedule(lpv, block)
    osc = build_oracle_cloth(oracle, spec)
    for _ in range(spec["steps"]):
        sim.runStep()
        osc.step(1, threads=THREADS)
    assert rel_err(sim.getState(0)[0], osc.get_state(0)[0]) < TOL
    sim.destroy()


def test_mixed_batch_equals_alone_bitwise(gpu_ctx):
    """Tet scenes and cloth scenes advanced by one launch per parallel group (mixed work
    items) give bit-identical results to each scene alone."""
    specs = scenes.c4_scenes(seed=3, num_scenes=8, vmin=300, vmax=3000)
    def build(spec):
        sim = VBDPhysics(gpu_ctx)
        sim.setPhysicsParams(**spec["physics"])
        if spec["kind"] == "cloth":
            sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"],
                             coloring=spec["coloring"], **spec["material"])
        else:
            sim.addTetMesh(spec["positions"], spec["tets"], coloring=spec["coloring"], **spec["material"])
        sim.initialize()
        return sim
    assert {s["kind"] for s in specs} == {"tet", "cloth"}
    alone = [build(s) for s in specs]
    mixed = [build(s) for s in specs]
    for s in alone:
        s.runSteps(5)
    b = Batch(gpu_ctx, mixed)
    b.record_graph(True)
    b.run_steps(5)
    for a, m in zip(alone, mixed):
        xa, va = a.getState(0)
        xm, vm = m.getState(0)
        assert np.array_equal(xa, xm) and np.array_equal(va, vm)
    b.destroy()
    for s in alone + mixed:
        s.destroy()


def test_cloth_json_loader(gpu_ctx, tmp_path, oracle):
    """Models.json with an "StVK_triMesh" .obj model (reference schema) == programmatic scene."""
    import json

    n = 8
    pos, faces, top = scenes.cloth_curtain(n, 0.5, 1.0)
    with open(tmp_path / "sheet.obj", "w") as f:
        for p in pos:
            f.write(f"v {p[0]:.9g} {p[1]:.9g} {p[2]:.9g}\n")
        for t in faces:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
    model = {"path": "${REPO_ROOT}/sheet.obj", "materialName": "StVK_triMesh", "density": 0.2, "miu": 1e4,
             "lambda": 2e4, "bendingStiffness": 1e-4, "fixedPoints": top,
             "translation": [0.0, 0.25, 0.0]}
    params = {"PhysicsParams": {"gravity": [0.0, -9.8, 0.0], "timeStep": 1 / 60, "numSubsteps": 1,
                                "iterations": 6, "handleCollision": False,
                                "worldBounds": [[-10, 0, -10], [10, 10, 10]]}}
    json.dump({"Models": [model]}, open(tmp_path / "Models.json", "w"))
    json.dump(params, open(tmp_path / "Parameters.json", "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(str(tmp_path / "Models.json"), str(tmp_path / "Parameters.json"),
                            repo_root=str(tmp_path))
    a.initialize()
    b = VBDPhysics(gpu_ctx)
    b.setPhysicsParams(gravity=(0.0, -9.8, 0.0), time_step=1 / 60, num_substeps=1, iterations=6,
                       world_lower=(-10.0, 0.0, -10.0), world_upper=(10.0, 10.0, 10.0))
    b.addClothMesh(pos + np.array([0.0, 0.25, 0.0]), faces, fixed_points=top, density=0.2, mu=1e4, lmbd=2e4,
                   bending_stiffness=1e-4)
    b.initialize()
    for _ in range(5):
        a.runStep()
        b.runStep()
    assert rel_err(a.getState(0)[0], b.getState(0)[0]) < 1e-6
    a.destroy()
    b.destroy()


def test_cloth_json_default_handle_collision(gpu_ctx, tmp_path):
    """The reference's cloth default handleCollision = true (VBDClothPhysicsParameters.h:60-62:
    contact stiffness 1e4, contact radius 1, query radius 0.4) turns collision handling on: the
    loaded scene runs the contact path (contact counters exist only for such scenes)."""
    import json

    pos, faces = scenes.cloth_sheet(4, 0.5, 1.0)
    with open(tmp_path / "s.obj", "w") as f:
        for p in pos:
            f.write(f"v {p[0]} {p[1]} {p[2]}\n")
        for t in faces:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")
    json.dump({"Models": [{"path": str(tmp_path / "s.obj"), "materialName": "StVK_triMesh"}]},
              open(tmp_path / "M.json", "w"))
    json.dump({"PhysicsParams": {"iterations": 2}}, open(tmp_path / "P.json", "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(str(tmp_path / "M.json"), str(tmp_path / "P.json"))
    a.initialize()
    a.runSteps(2)
    x, _ = a.getState(0)
    assert np.all(np.isfinite(x))
    st = a.contactStats()
    assert st["frames"] == 2 * 20 and st["detections"] >= st["frames"]  # 20 substeps per frame by default
    a.destroy()
