"""GPU parity of the tet VBD hot path: HIP kernels through the C ABI vs the FP32 oracle.

Contract: from identical (x, v), one step of the HIP path stays within 1e-5
normwise-relative position error of the oracle's FP32 step
(rel = max|x_gpu - x_ref| / max|x_ref|). Trajectory drift and FP32-vs-FP64
error are reported and bounded loosely.
"""
import json
import os

import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from vbd_cases import build_gpu, build_oracle, c1_spec, multi_spec, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
THREADS = 8


@pytest.fixture(scope="module")
def c1_oracle_traj(oracle):
    """FP32 and FP64 oracle trajectories of the soft-cube drop (300 steps)."""
    spec = c1_spec()
    out = {}
    for prec in (32, 64):
        sc = build_oracle(oracle, spec, precision=prec)
        xs, vs = [], []
        x, v = sc.get_state()
        xs.append(x)
        vs.append(v)
        for _ in range(spec["steps"]):
            sc.step(1, threads=THREADS)
            x, v = sc.get_state()
            xs.append(x)
            vs.append(v)
        out[prec] = (np.stack(xs), np.stack(vs))
    print(f"\n[c1] oracle trajectories done", flush=True)
    return spec, out


def test_c1_groups_match_oracle(gpu_ctx, oracle):
    spec = c1_spec()
    sim = build_gpu(gpu_ctx, spec)
    og = build_oracle(oracle, spec).groups()
    gg = sim.parallelGroups()
    assert len(gg) == len(og) == 8
    for a, b in zip(gg, og):
        assert np.array_equal(a, b)
    assert [len(g) for g in gg] == [665, 216, 138, 126, 110, 49, 19, 8]
    sim.destroy()


def test_c1_per_step_parity(gpu_ctx, c1_oracle_traj):
    """Every one of the 300 steps, started from the oracle's state, matches the
    oracle's next state within 1e-5."""
    spec, traj = c1_oracle_traj
    xs, vs = traj[32]
    sim = build_gpu(gpu_ctx, spec)
    worst = 0.0
    for k in range(spec["steps"]):
        sim.setState(0, xs[k], vs[k])
        sim.runStep()
        x, v = sim.getState(0)
        e = rel_err(x, xs[k + 1])
        worst = max(worst, e)
        assert e < TOL, f"step {k}: rel err {e:.3e}"
    print(f"\n[c1] worst per-step rel err over {spec['steps']} steps: {worst:.3e}", flush=True)
    sim.destroy()


def test_c1_trajectory(gpu_ctx, c1_oracle_traj):
    """Free-running 300-step trajectory vs the FP32 oracle, and FP32 vs FP64."""
    spec, traj = c1_oracle_traj
    x32, _ = traj[32]
    x64, _ = traj[64]
    sim = build_gpu(gpu_ctx, spec)
    sim.recordVBDSolveGraph(True)
    errs = []
    for k in range(spec["steps"]):
        sim.runStep()
        if k % 10 == 9:
            x, _ = sim.getState(0)
            errs.append(rel_err(x, x32[k + 1]))
    fp32_vs_fp64 = max(rel_err(x32[k], x64[k]) for k in range(len(x32)))
    print(f"\n[c1] trajectory drift gpu-vs-oracle32 max {max(errs):.3e}; oracle32-vs-oracle64 max {fp32_vs_fp64:.3e}",
          flush=True)
    # the cube lands and comes to rest: no chaotic amplification
    assert max(errs) < 1e-3
    assert fp32_vs_fp64 < 1e-3
    # it is resting on the ground at the end
    x, _ = sim.getState(0)
    assert x[:, 1].min() > -0.02 and x[:, 1].min() < 0.01
    sim.destroy()


def test_multi_mesh_per_step_parity(gpu_ctx, oracle):
    """Two meshes (merged parallel groups), damping, fixed points, init type 2,
    substeps, friction start iteration."""
    spec = multi_spec()
    sim = build_gpu(gpu_ctx, spec)
    osc = build_oracle(oracle, spec)
    for a, b in zip(sim.parallelGroups(), osc.groups()):
        assert np.array_equal(a, b)
    worst = 0.0
    for k in range(spec["steps"]):
        states = [osc.get_state(m) for m in range(2)]
        for m in range(2):
            sim.setState(m, *states[m])
        sim.runStep()
        osc.step(1, threads=THREADS)
        for m in range(2):
            x, v = sim.getState(m)
            xr, vr = osc.get_state(m)
            e = rel_err(x, xr)
            worst = max(worst, e)
            assert e < TOL, f"step {k} mesh {m}: {e:.3e}"
            if m == 1:
                fixed = spec["meshes"][1]["fixed"]
                assert np.all(v[list(fixed)] == 0)
    print(f"\n[multi] worst per-step rel err {worst:.3e}", flush=True)
    sim.destroy()


def test_graph_replay_equals_eager(gpu_ctx):
    spec = c1_spec()
    a = build_gpu(gpu_ctx, spec)
    b = build_gpu(gpu_ctx, spec)
    b.recordVBDSolveGraph(True)
    for _ in range(7):
        a.runStep()
        b.runStep()
    xa, va = a.getState(0)
    xb, vb = b.getState(0)
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    a.destroy()
    b.destroy()


def test_per_iteration_units_equal_run_step(gpu_ctx):
    """applyInitialStep + solveParallelGroup x iterations x groups + updateVelocities == runStep."""
    spec = c1_spec()
    a = build_gpu(gpu_ctx, spec)
    b = build_gpu(gpu_ctx, spec)
    ngroups = len(a.parallelGroups())
    for _ in range(3):
        a.runStep()
        b.applyInitialStep()
        for it in range(spec["physics"]["iterations"]):
            for g in range(ngroups):
                b.solveParallelGroup(g, it)
        b.updateVelocities()
    assert np.array_equal(a.getState(0)[0], b.getState(0)[0])
    a.destroy()
    b.destroy()


@pytest.mark.parametrize("lpv,block", [(1, 64), (2, 128), (4, 256), (8, 256), (16, 64)])
def test_schedules_agree(gpu_ctx, oracle, lpv, block):
    spec = c1_spec()
    sim = build_gpu(gpu_ctx, spec)
    sim.setSchedule(lpv, block)
    osc = build_oracle(oracle, spec)
    for _ in range(5):
        sim.runStep()
        osc.step(1, threads=THREADS)
    e = rel_err(sim.getState(0)[0], osc.get_state(0)[0])
    assert e < TOL
    sim.destroy()


def small_scene(ctx, n, origin, seed):
    p, t = scenes.cube_tet_mesh(n, 0.08, origin=origin)
    rng = np.random.default_rng(seed)
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(gravity=(0.0, -9.8, 0.0), time_step=1 / 60, num_substeps=1, iterations=10,
                         world_lower=(-1e7, 0.0, -1e7))
    sim.addTetMesh(p, t, velocities=rng.normal(scale=0.5, size=p.shape), density=1000.0,
                   mu=float(10 ** rng.uniform(4, 6)), lmbd=1e6, coloring="greedy")
    sim.initialize()
    return sim


def test_batch_equals_alone_bitwise(gpu_ctx):
    """Scenes advanced together in one launch per group give bit-identical results to
    the same scenes advanced alone."""
    sizes = [2, 5, 3, 7, 4]
    alone = [small_scene(gpu_ctx, n, (0.0, 0.2 + 0.1 * i, 0.0), i) for i, n in enumerate(sizes)]
    mixed = [small_scene(gpu_ctx, n, (0.0, 0.2 + 0.1 * i, 0.0), i) for i, n in enumerate(sizes)]
    for s in alone:
        s.runSteps(6)
    batch = Batch(gpu_ctx, mixed)
    batch.record_graph(True)
    batch.run_steps(6)
    for a, m in zip(alone, mixed):
        xa, va = a.getState(0)
        xm, vm = m.getState(0)
        assert np.array_equal(xa, xm) and np.array_equal(va, vm)
    batch.destroy()
    for s in alone + mixed:
        s.destroy()


def test_json_loader_equals_programmatic(gpu_ctx, tmp_path):
    pos, tets = scenes.cube_tet_mesh(4, 0.1, origin=(0.0, 0.0, 0.0))
    colors = gaia_amd.color_tet_mesh(len(pos), tets, "mcs")
    physics = dict(gravity=(0.0, -9.8, 0.0), time_step=1 / 60, num_substeps=2, iterations=8,
                   world_lower=(-10.0, 0.0, -10.0), world_upper=(10.0, 20.0, 10.0))
    material = dict(density=1000.0, mu=1e5, lmbd=1e6)
    models, params = scenes.write_scene_json(str(tmp_path), "cube", pos, tets, physics, material, colors=colors,
                                             translation=(0.0, 0.3, 0.0))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(models, params, repo_root=str(tmp_path))
    a.initialize()
    b = VBDPhysics(gpu_ctx)
    b.setPhysicsParams(**physics)
    b.addTetMesh(pos + np.array([0.0, 0.3, 0.0]), tets, vertex_colors=colors, **material)
    b.initialize()
    for _ in range(5):
        a.runStep()
        b.runStep()
    xa, _ = a.getState(0)
    xb, _ = b.getState(0)
    assert rel_err(xa, xb) < 1e-6
    a.destroy()
    b.destroy()


def test_free_fall_closed_form_gpu(gpu_ctx):
    pos, tets = scenes.cube_tet_mesh(3, 0.1)
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(gravity=(0.0, -9.8, 0.0), time_step=0.01, num_substeps=1, iterations=4, use_plane_ground=0)
    v0 = np.tile([0.3, 0.1, -0.2], (len(pos), 1))
    sim.addTetMesh(pos, tets, velocities=v0, density=1000.0, mu=1e5, lmbd=1e6)
    sim.initialize()
    sim.runStep()
    x, v = sim.getState(0)
    dt = float(np.float32(0.01))
    assert np.allclose(x, pos + dt * v0 + dt * dt * np.array([0, -9.8, 0]), atol=2e-7)
    sim.destroy()


def test_invalid_arguments_raise(gpu_ctx):
    sim = VBDPhysics(gpu_ctx)
    pos, tets = scenes.cube_tet_mesh(2, 0.1)
    sim.addTetMesh(pos, tets, vertex_colors=np.zeros(len(pos), np.int32))
    with pytest.raises(gaia_amd.VBDError, match="not valid"):
        sim.initialize()
    sim2 = VBDPhysics(gpu_ctx)
    sim2.addTetMesh(pos, tets)
    sim2.initialize()
    with pytest.raises(gaia_amd.VBDError, match="schedule"):
        sim2.setSchedule(3, 256)
    with pytest.raises(gaia_amd.VBDError, match="group index"):
        sim2.solveParallelGroup(99)
    sim2.destroy()
