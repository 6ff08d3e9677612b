"""Persistent work-queue solve (one launch per iteration over a flattened (scene, colour, chunk)
queue): bitwise equal to one launch per parallel group, and within 1e-5 of the FP32 oracle."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, VBDPhysics, scenes
from gaia_amd.autotune import autotune, default_candidates, schedule_key
from vbd_cases import build_gpu, build_oracle, build_oracle_cloth, cloth_drop_spec, multi_spec, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def build_spec_scene(ctx, spec):
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**spec["physics"])
    if spec["kind"] == "cloth":
        sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"], coloring=spec["coloring"],
                         **spec["material"])
    else:
        sim.addTetMesh(spec["positions"], spec["tets"], coloring=spec["coloring"], **spec["material"])
    sim.initialize()
    return sim


def states(sims):
    return [s.getState(0) for s in sims]


@pytest.mark.parametrize("lpv,block,xcd", [(8, 128, 1), (4, 256, 16), (16, 64, 64), (8, 256, 4)])
def test_xcd_block_mapping_equals_block_order_bitwise(gpu_ctx, lpv, block, xcd):
    """Group launches whose blocks take the work items in XCD-contiguous ranges give bit-identical
    positions and velocities to block order (each block's work is the same, only where it runs
    changes), captured or not."""
    specs = scenes.c4_scenes(seed=6, num_scenes=12, vmin=300, vmax=4000)
    ref = [build_spec_scene(gpu_ctx, s) for s in specs]
    mapped = [build_spec_scene(gpu_ctx, s) for s in specs]
    br = Batch(gpu_ctx, ref)
    br.set_schedule(lpv, block, mode="group")
    bx = Batch(gpu_ctx, mapped)
    bx.set_schedule(lpv, block, mode="group", xcd_blocks=xcd)
    assert bx.get_schedule()["xcd_blocks"] == xcd
    bx.record_graph(True)
    br.run_steps(4)
    bx.run_steps(4)
    for (xa, va), (xb, vb) in zip(states(ref), states(mapped)):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    br.destroy()
    bx.destroy()
    for s in ref + mapped:
        s.destroy()


QUEUE_SCHEDULES = [
    dict(lanes_per_vertex=8, block=128, vertices_per_entry=64),
    dict(lanes_per_vertex=4, block=256, vertices_per_entry=64),
    dict(lanes_per_vertex=16, block=256, vertices_per_entry=16, blocks_per_cu=1),
    dict(lanes_per_vertex=1, block=64, vertices_per_entry=512),
]


@pytest.mark.parametrize("sched", QUEUE_SCHEDULES, ids=lambda s: f"{s['lanes_per_vertex']}x{s['block']}"
                         f"e{s['vertices_per_entry']}g{s.get('blocks_per_cu', 0)}")
def test_queue_batch_equals_group_launches_bitwise(gpu_ctx, sched):
    """Mixed tet + cloth scenes: the persistent queue launch per iteration gives bit-identical
    positions and velocities to one launch per parallel group with the same lanes per vertex
    (same per-vertex arithmetic and summation order, same colour order inside each scene)."""
    specs = scenes.c4_scenes(seed=5, num_scenes=12, vmin=300, vmax=4000)
    assert {s["kind"] for s in specs} == {"tet", "cloth"}
    ref = [build_spec_scene(gpu_ctx, s) for s in specs]
    que = [build_spec_scene(gpu_ctx, s) for s in specs]
    br = Batch(gpu_ctx, ref)
    br.set_schedule(sched["lanes_per_vertex"], 64, mode="group")
    bq = Batch(gpu_ctx, que)
    bq.set_schedule(mode="queue", **sched)
    entries, grid = bq.queue_info()
    print(f"\n[queue {sched}] {entries} entries per iteration, grid {grid}", flush=True)
    assert entries > 0 and grid > 0
    bq.record_graph(True)
    br.run_steps(4)
    bq.run_steps(4)
    for (xa, va), (xb, vb) in zip(states(ref), states(que)):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    for b in (br, bq):
        b.destroy()
    for s in ref + que:
        s.destroy()


def test_queue_tet_multi_mesh_substeps_vs_oracle(gpu_ctx, oracle):
    """Two tet meshes (merged groups), 2 substeps per frame (counters reset per substep),
    friction start at iteration 1: per-step parity with the FP32 oracle."""
    spec = multi_spec()
    sim = build_gpu(gpu_ctx, spec)
    sim.setSchedule(8, 128, mode="queue", vertices_per_entry=32)
    osc = build_oracle(oracle, spec)
    nm = len(spec["meshes"])
    for k in range(10):
        for m in range(nm):
            sim.setState(m, *osc.get_state(m))
        sim.runStep()
        osc.step(1, threads=8)
        for m in range(nm):
            e = rel_err(sim.getState(m)[0], osc.get_state(m)[0])
            assert e < TOL, f"step {k} mesh {m}: {e:.3e}"
    sim.destroy()


def test_queue_cloth_commits_vs_oracle(gpu_ctx, oracle):
    """Two cloth meshes whose colours share parallel groups: those slots keep the positionsNext
    + commit scheme, which the queue runs as commit entries after the level's solves."""
    spec = cloth_drop_spec()
    sim = build_gpu(gpu_ctx, spec)
    sim.setSchedule(4, 256, mode="queue", vertices_per_entry=64)
    osc = build_oracle_cloth(oracle, spec)
    nm = len(spec["meshes"])
    for k in range(12):
        for m in range(nm):
            sim.setState(m, *osc.get_state(m))
        sim.runStep()
        osc.step(1, threads=8)
        for m in range(nm):
            e = rel_err(sim.getState(m)[0], osc.get_state(m)[0])
            assert e < TOL, f"step {k} mesh {m}: {e:.3e}"
    sim.destroy()


@pytest.mark.parametrize("mode", ["group", "queue"])
def test_iteration_hooks_equal_run_step(gpu_ctx, mode):
    """applyInitialStep + runIteration(i) for every iteration + updateVelocities == runStep."""
    spec = multi_spec()
    a = build_gpu(gpu_ctx, spec)
    b = build_gpu(gpu_ctx, spec)
    for s in (a, b):
        s.setSchedule(8, 128, mode=mode)
    phys = spec["physics"]
    for _ in range(3):
        a.runStep()
        for _sub in range(phys["num_substeps"]):
            b.applyInitialStep()
            for it in range(phys["iterations"]):
                b.runIteration(it)
            b.updateVelocities()
    for m in range(len(spec["meshes"])):
        xa, va = a.getState(m)
        xb, vb = b.getState(m)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    a.destroy()
    b.destroy()


def test_batch_per_group_calls_equal_run_iteration(gpu_ctx):
    """The per-group unit of a batch (gvbd_batch_solve_parallel_group) over every group of one
    iteration equals the iteration: cloth slots solved through positionsNext are committed after
    their group, and a contact scene's same-group partners are read from the iteration snapshot."""
    from vbd_cases import contact_drape_spec

    for spec in (cloth_drop_spec(), contact_drape_spec()):
        a, b = build_gpu(gpu_ctx, spec), build_gpu(gpu_ctx, spec)
        ba, bb = Batch(gpu_ctx, [a]), Batch(gpu_ctx, [b])
        contact = spec["physics"]["handle_collision"] != 0
        for _ in range(2):
            for sim, bt in ((a, ba), (b, bb)):
                if contact:
                    sim.applyCollisionDetection()
                bt.apply_initial_step()
            for it in range(3):
                ba.run_iteration(it)
                for g in range(bb.num_groups()):
                    bb.solve_group(g, it)
            ba.update_velocities()
            bb.update_velocities()
        for m in range(len(spec["meshes"])):
            (xa, va), (xb, vb) = a.getState(m), b.getState(m)
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"mesh {m}"
        for x in (ba, bb, a, b):
            x.destroy()


def test_schedule_validation(gpu_ctx):
    spec = multi_spec()
    sim = build_gpu(gpu_ctx, spec)
    with pytest.raises(gaia_amd.VBDError, match="vertices_per_entry"):
        sim.setSchedule(8, 128, mode="queue", vertices_per_entry=20)
    with pytest.raises(gaia_amd.VBDError, match="lanes_per_vertex"):
        sim.setSchedule(3, 128, mode="queue")
    sim.setSchedule(8, 128, mode="queue", vertices_per_entry=48)
    got = sim.getSchedule()
    assert got["mode"] == "queue" and got["vertices_per_entry"] == 48 and got["lanes_per_vertex"] == 8
    sim.destroy()


def test_autotuner_picks_a_measured_candidate(gpu_ctx):
    specs = scenes.c4_scenes(seed=9, num_scenes=6, vmin=300, vmax=2000)
    sims = [build_spec_scene(gpu_ctx, s) for s in specs]
    b = Batch(gpu_ctx, sims)
    cands = default_candidates()[:2] + default_candidates(("queue",))[:2]
    best, results = autotune(gpu_ctx, b, cands, reps=1)
    assert len(results) == len(cands)
    assert results[schedule_key(best)] == min(results.values())
    b.run_steps(1)
    gpu_ctx.synchronize()
    b.destroy()
    for s in sims:
        s.destroy()


@pytest.mark.parametrize("lpv,block,stage", [(8, 128, 64), (4, 128, 32), (16, 256, 64)])
def test_lds_staged_records_equal_direct_bitwise(gpu_ctx, lpv, block, stage):
    """Incidence records staged through LDS in block-wide chunks: each lane handles the same
    records in the same order as when it reads them from HBM, so tet and cloth scenes (with
    in-place and positionsNext groups, damping) step to identical bits."""
    specs = scenes.c4_scenes(seed=12, num_scenes=8, vmin=300, vmax=5000)
    a = [build_spec_scene(gpu_ctx, s) for s in specs]
    b = [build_spec_scene(gpu_ctx, s) for s in specs]
    ba, bb = Batch(gpu_ctx, a), Batch(gpu_ctx, b)
    ba.set_schedule(lpv, block, mode="group")
    bb.set_schedule(lpv, block, mode="group", stage_records=stage)
    assert bb.get_schedule()["stage_records"] == stage
    ba.run_steps(3)
    bb.run_steps(3)
    for (xa, va), (xb, vb) in zip(states(a), states(b)):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
    ba.destroy()
    bb.destroy()
    for s in a + b:
        s.destroy()


def test_lds_staged_schedule_vs_oracle(gpu_ctx, oracle):
    """A staged schedule on the two-sheet drop (merged groups, positionsNext) and the two-cube
    scene stays within 1e-5 of the FP32 oracle per step."""
    for spec, build_o in ((cloth_drop_spec(), build_oracle_cloth), (multi_spec(), build_oracle)):
        sim = build_gpu(gpu_ctx, spec)
        sim.setSchedule(8, 128, mode="group", stage_records=32)
        osc = build_o(oracle, spec)
        for _ in range(5):
            for m in range(len(spec["meshes"])):
                sim.setState(m, *osc.get_state(m))
            sim.runStep()
            osc.step(1, threads=8)
            for m in range(len(spec["meshes"])):
                assert rel_err(sim.getState(m)[0], osc.get_state(m)[0]) < TOL
        sim.destroy()


def test_stage_validation(gpu_ctx):
    sim = build_gpu(gpu_ctx, multi_spec())
    with pytest.raises(gaia_amd.VBDError, match="stage_records"):
        sim.setSchedule(8, 128, mode="group", stage_records=48)
    with pytest.raises(gaia_amd.VBDError, match="stage_records"):
        sim.setSchedule(1, 64, mode="group", stage_records=32)
    with pytest.raises(gaia_amd.VBDError, match="group launches only"):
        sim.setSchedule(8, 128, mode="queue", stage_records=32)
    sim.destroy()
