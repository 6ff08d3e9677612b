"""Scenes with the penetration contact model (tet_contact_model 1) batched, captured and served
with their detection trees resident: their frames need no host decision (DCD at the substep's
start, CCD after the initial step, no re-detection), so they join batches of plain scenes, replay
as HIP graphs (frames that rebuild the trees launched directly, the graph captured again after
them) and are served under a memory budget that copies arenas out and back. Every result equals
the scene stepped alone, bit for bit."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, Server, VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_gpu, penetration_spec

pytestmark = pytest.mark.gpu


def pen_specs(n=3, rebuild_steps=4):
    out = []
    for k in range(n):
        spec = penetration_spec(steps=0, bodies=2 + (k % 2), vy=-1.5 - 0.5 * k, seed=k)
        spec["physics"]["dcd_bvh_rebuild_steps"] = rebuild_steps
        out.append(spec)
    return out


def plain_tet_spec():
    pos, tets = scenes.cube_tet_mesh(4, 0.1, origin=(0.0, 0.3, 0.0))
    return dict(physics=physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                                     world_lower=(-1e7, 0.0, -1e7)),
                meshes=[dict(pos=pos, tets=tets, material=dict(density=1000.0, mu=1e5, lmbd=1e6), fixed=(),
                             vel=None, coloring="greedy")])


def plain_cloth_spec():
    pos, faces = scenes.cloth_sheet(16, 0.6, 0.8)
    mu, lm = scenes.stvk_lame(1e5, 0.3)
    return dict(physics=physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=10,
                                     world_lower=(-1e7, 0.0, -1e7)),
                meshes=[dict(pos=pos, faces=faces, material=dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4),
                             fixed=[0, 15], vel=None, coloring="mcs")])


def states(sim):
    return [sim.getState(m) for m in range(sim.numMeshes())]


def same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) for x, y in zip(a, b))


def alone(ctx, spec, frames, graph=False):
    sim = build_gpu(ctx, spec)
    if graph:
        sim.recordVBDSolveGraph(True)
    sim.runSteps(frames)
    out = states(sim)
    sim.destroy()
    return out


def test_batch_of_penetration_and_plain_scenes_equals_alone(gpu_ctx):
    specs = pen_specs() + [plain_tet_spec(), plain_cloth_spec()]
    frames = 10
    ref = [alone(gpu_ctx, sp, frames) for sp in specs]
    sims = [build_gpu(gpu_ctx, sp) for sp in specs]
    batch = Batch(gpu_ctx, sims)
    batch.run_steps(frames)
    for i, s in enumerate(sims):
        assert same(states(s), ref[i]), f"scene {i}"
    hits = sum(int((s.penetrationResults()[0][:, 3] >= 0).sum()) for s in sims[:3])
    print(f"\n[penetration batch] 3 contact + 2 plain scenes, {frames} frames bitwise alone; {hits} hits in the "
          f"last detection", flush=True)
    batch.destroy()
    for s in sims:
        s.destroy()


def test_graph_replay_equals_launches_across_tree_rebuilds(gpu_ctx):
    """A captured contact scene and batch over 13 frames with the trees rebuilt every 4: frames 4, 8
    and 12 run directly, the others replay the graph; results equal direct launches bitwise."""
    frames = 13
    for spec in pen_specs(n=1):
        assert same(alone(gpu_ctx, spec, frames, graph=True), alone(gpu_ctx, spec, frames))
    specs = pen_specs(n=2) + [plain_tet_spec()]
    ref = [alone(gpu_ctx, sp, frames) for sp in specs]
    sims = [build_gpu(gpu_ctx, sp) for sp in specs]
    batch = Batch(gpu_ctx, sims)
    batch.record_graph(True)
    batch.run_steps(frames)
    for i, s in enumerate(sims):
        assert same(states(s), ref[i]), f"scene {i}"
    stats = sims[0].contactStats()
    assert stats["frames"] == frames and stats["rebuilds"] == 3
    batch.destroy()
    for s in sims:
        s.destroy()


def test_served_penetration_scenes_equal_alone_under_eviction(gpu_ctx):
    """Six contact scenes and two plain ones served under a budget that holds about half of their
    arenas: idle scenes are copied out and back (the detector and its trees stay resident), and
    every served scene equals the same scene stepped alone for as many frames, bitwise."""
    specs = pen_specs(n=6) + [plain_tet_spec(), plain_cloth_spec()]
    sims = []
    for sp in specs:
        sim = VBDPhysics(gpu_ctx)
        sim.setPhysicsParams(**sp["physics"])
        for m in sp["meshes"]:
            if "faces" in m:
                sim.addClothMesh(m["pos"], m["faces"], fixed_points=m["fixed"],
                                 vertex_colors=gaia_amd.color_tri_mesh(len(m["pos"]), m["faces"], m["coloring"]),
                                 **m["material"])
            else:
                sim.addTetMesh(m["pos"], m["tets"], fixed_points=m["fixed"], velocities=m["vel"],
                               vertex_colors=gaia_amd.color_tet_mesh(len(m["pos"]), m["tets"], m["coloring"]),
                               **m["material"])
        sim.initialize(defer_upload=True)
        sims.append(sim)
    # the budget: every scene's detector (resident from its admission) plus half of the arenas,
    # as the server counts them (a probe of each scene)
    def device_bytes(spec):
        probe = build_gpu(gpu_ctx, spec)
        b = probe.deviceBytes()
        probe.destroy()
        return b["arena"], b["detector"]

    arenas, detectors = zip(*[device_bytes(sp) for sp in specs])
    budget = int(sum(detectors) + max(0.5 * sum(arenas), 1.1 * max(arenas)))
    print(f"\n[served penetration] arenas {arenas}, detectors {detectors}, budget {budget}", flush=True)
    srv = Server(gpu_ctx, budget)
    for s in sims:
        srv.add_scene(s)
    rng = np.random.default_rng(5)
    counts = np.zeros(len(sims), int)
    tag = 0
    for _ in range(40):
        for i in rng.integers(0, len(sims), size=rng.integers(1, 4)):
            srv.request_step(sims[i], tag)
            counts[i] += 1
            tag += 1
        srv.tick()
    for _ in range(1000):
        if not srv.stats()["pending"]:
            break
        srv.tick()
    assert srv.stats()["pending"] == 0
    done = srv.poll(capacity=tag + 16)
    st = srv.stats()
    assert len(done) == tag and all(d["frame"] >= 0 for d in done)
    print(f"\n[served penetration] {tag} requests, evictions {st['evictions']}, admissions {st['admissions']}",
          flush=True)
    assert st["evictions"] > 0, st
    got = []
    for s in sims:
        srv.remove_scene(s)
        got.append(states(s))
    srv.destroy()
    for i, sp in enumerate(specs):
        assert same(got[i], alone(gpu_ctx, sp, int(counts[i]))), f"scene {i} ({counts[i]} frames)"
    for s in sims:
        s.destroy()
