"""GPU parity of the Chebyshev-accelerated tet iterations (runStepGPU, VBD/VBDPhysics.cpp:872-906;
VBDSolveParallelGroup_applyAcceleration_kernel, VBDPhysicsCompute.cu:910-944): the HIP path
through the C ABI against the FP32 oracle, per step from identical state, 1e-5
normwise-relative; in a batch with plain scenes each equals the scene alone bit for bit."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_gpu, build_oracle, c1_spec, multi_spec, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
DT = float(np.float32(1.0 / 60.0))


def accelerated(spec, rho=0.93, iterations=None):
    spec["physics"] = dict(spec["physics"], use_accelerator=1, accelerator_rho=rho)
    if iterations:
        spec["physics"]["iterations"] = iterations
    return spec


@pytest.mark.parametrize("case", ["c1", "multi"])
def test_accelerated_per_step_parity(gpu_ctx, oracle, case):
    spec = accelerated(c1_spec(), iterations=10) if case == "c1" else accelerated(multi_spec())
    steps = 40
    sim = build_gpu(gpu_ctx, spec)
    o32 = build_oracle(oracle, spec)
    nm = len(spec["meshes"])
    errs = []
    for _ in range(steps):
        for m in range(nm):
            sim.setState(m, *o32.get_state(m))
        sim.runStep()
        o32.step(1, threads=8)
        errs.append(max(rel_err(sim.getState(m)[0], o32.get_state(m)[0]) for m in range(nm)))
    sim.destroy()
    errs = np.array(errs)
    print(f"\n[accelerator {case}] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: {errs.max():.3e}"


def test_free_tet_closed_form(gpu_ctx):
    """The oracle's closed form on the device: free vertices started at x0 = x + dt v by
    initialisation type 2 follow a_k = a_{k-2} + omega_k (y - a_{k-2}), a_{-1} = a_0 = x0."""
    rho, iters, g, v0 = 0.9, 5, np.array([0.0, -9.8, 0.0]), np.array([0.3, 0.1, -0.2])
    pos = np.array([[0.0, 1.0, 0.0], [0.1, 1.0, 0.0], [0.0, 1.1, 0.0], [0.0, 1.0, 0.1]])
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(**physics_dict(gravity=tuple(g), time_step=DT, num_substeps=1, iterations=iters,
                                        use_accelerator=1, accelerator_rho=rho, world_lower=(-1e7, -1e7, -1e7)))
    sim.addTetMesh(pos, np.int32([[0, 1, 2, 3]]), vertex_colors=np.arange(4, dtype=np.int32), density=1000.0,
                   mu=1e-12, lmbd=1e-12, initialization_type=2, velocities=np.tile(v0, (4, 1)))
    sim.initialize()
    sim.runStep()
    x, _ = sim.getState(0)
    x0 = pos + DT * v0
    y = x0 + DT * DT * g
    a, w = [x0, x0], 1.0
    for k in range(1, iters + 1):
        w = 1.0 if k == 1 else (2.0 / (2.0 - rho * rho) if k == 2 else 4.0 / (4.0 - rho * rho * w))
        a.append(a[-2] + w * (y - a[-2]))
    assert rel_err(x, a[-1]) < 1e-6
    sim.destroy()


def small(ctx, n, origin, seed, accel):
    pos, tets = scenes.cube_tet_mesh(n, 0.1, origin=origin)
    vel = np.random.default_rng(seed).normal(scale=0.3, size=pos.shape)
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=8,
                                        world_lower=(-1e7, 0.0, -1e7), use_accelerator=accel, accelerator_rho=0.85))
    sim.addTetMesh(pos, tets, velocities=vel, density=1000.0, mu=1e5, lmbd=1e6)
    sim.initialize()
    return sim


def test_batch_with_accelerated_and_plain_scenes_equals_alone(gpu_ctx):
    """Accelerated and plain scenes in one batch (graph replay): each as when stepped alone."""
    cfg = [(3, 1), (4, 0), (2, 1), (5, 0)]
    alone = [small(gpu_ctx, n, (0.0, 0.3 + 0.1 * i, 0.0), i, a) for i, (n, a) in enumerate(cfg)]
    mixed = [small(gpu_ctx, n, (0.0, 0.3 + 0.1 * i, 0.0), i, a) for i, (n, a) in enumerate(cfg)]
    for s in alone:
        s.runSteps(5)
    batch = Batch(gpu_ctx, mixed)
    batch.record_graph(True)
    batch.run_steps(5)
    for a, m in zip(alone, mixed):
        xa, va = a.getState(0)
        xm, vm = m.getState(0)
        assert np.array_equal(xa, xm) and np.array_equal(va, vm)
    batch.destroy()
    for s in alone + mixed:
        s.destroy()


def test_accelerator_arguments(gpu_ctx, tmp_path):
    pos, tets = scenes.cube_tet_mesh(2, 0.1, origin=(0.0, 0.5, 0.0))
    for bad in (dict(accelerator_rho=1.0), dict(handle_collision=1)):
        sim = VBDPhysics(gpu_ctx)
        sim.setPhysicsParams(use_accelerator=1, **bad)
        sim.addTetMesh(pos, tets)
        with pytest.raises(gaia_amd.VBDError, match="accelerator"):
            sim.initialize()
        sim.destroy()
    # the JSON loader reads useAccelerator / acceleratorRho: the loaded scene steps as the
    # programmatic one
    import json

    colors = gaia_amd.color_tet_mesh(len(pos), tets, "mcs")
    models, params = scenes.write_scene_json(str(tmp_path), "cube", pos, tets,
                                             physics_dict(iterations=6, num_substeps=1),
                                             dict(density=1000.0, mu=1e5, lmbd=1e6), colors=colors)
    p = json.load(open(params))
    p["PhysicsParams"].update(useAccelerator=True, acceleratorRho=0.7)
    json.dump(p, open(params, "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(models, params, repo_root=str(tmp_path))
    a.initialize()
    a.runSteps(3)
    b = VBDPhysics(gpu_ctx)
    b.setPhysicsParams(**physics_dict(iterations=6, num_substeps=1, use_accelerator=1, accelerator_rho=0.7))
    b.addTetMesh(pos, tets, vertex_colors=colors, density=1000.0, mu=1e5, lmbd=1e6)
    b.initialize()
    b.runSteps(3)
    assert np.array_equal(a.getState(0)[0], b.getState(0)[0])
    a.destroy()
    b.destroy()
