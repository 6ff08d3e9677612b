"""Parity at the benchmark configurations' own sizes, through the C ABI on the device.

- Batched VBD (256 scenes of log-uniform 500-50k vertices, tet cubes and two-corner sheets,
  3,274,854 free vertices): every tet scene's first frame within 1e-5 of the FP32 oracle, every
  cloth scene's within the spread of the FP32 and FP64 oracles (an unstressed sheet's first frame
  is ill-conditioned), and the batched
  frames bitwise equal to each scene stepped alone, under the benchmark's schedule and under a
  persistent work-queue schedule.
- The two-corner sheet over 60 frames: per-step error distributions of the device against the
  FP32 oracle and of the FP32 oracle against the FP64 one, with 3x3 solve failures counted.
- Tet bodies in contact (32 bodies of 8,640 tets, 70,304 vertices): per-step parity against the
  FP32 oracle from identical state on frames of the fall including impacts, stated against the
  FP32-vs-FP64 spread, and no vertex inside another body over a 60-frame fall.
- The same bodies with the tet solver's penetration contact, piled up on the floor: detections
  equal to the FP32 oracle's per vertex on 6 frames from identical state.
- Served VBD (2,000 scenes of 1k-200k vertices) under a memory budget of a quarter of the scenes'
  bytes: scenes are copied out and back between frames, and a seeded sample of served scenes
  equals the same scenes stepped alone, bitwise.
"""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, Server, VBDPhysics, scenes
from gaia_amd.autotune import apply_schedule, parse_schedule_key
from gaia_amd.vbd import physics_dict
from config_checks import build_spec_oracle, build_spec_scene, host_threads, vertices_inside_other_bodies
from vbd_cases import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
BENCH_SCHEDULE = "group:8x128"  # the batched benchmark's autotuned schedule
QUEUE_SCHEDULE = "queue:8x256:e512"


@pytest.fixture(scope="module")
def c4_specs():
    specs = scenes.c4_scenes(seed=0, num_scenes=256)
    assert sum(len(s["positions"]) - len(s.get("fixed", [])) for s in specs) == 3274854
    return specs


def _states(sims):
    return [s.getState(0) for s in sims]


def test_c4_batched_first_frame_vs_oracle_and_alone_bitwise(gpu_ctx, oracle, c4_specs):
    """All 256 scenes batched in one graph per frame: the first frame of every tet scene within 1e-5
    of the FP32 oracle (normwise relative) and the cloth scenes within the FP32-vs-FP64 spread, then
    after 3 frames bitwise equal to the scene alone, for the benchmark's group-launch schedule and a
    work-queue schedule (same lanes per vertex)."""
    specs = c4_specs
    alone = [build_spec_scene(gpu_ctx, s) for s in specs]
    for s in alone:
        s.setSchedule(8, 128, mode="group")
        s.runSteps(3)
    ref = _states(alone)
    for s in alone:
        s.destroy()

    sims = [build_spec_scene(gpu_ctx, s) for s in specs]
    init = _states(sims)
    batch = Batch(gpu_ctx, sims)
    apply_schedule(batch, parse_schedule_key(BENCH_SCHEDULE))
    batch.run_steps(1)
    first = _states(sims)
    threads = host_threads()
    errs, spread = [], []
    for i, spec in enumerate(specs):
        o = build_spec_oracle(oracle, spec)
        o.step(1, threads=threads)
        errs.append(rel_err(first[i][0], o.get_state(0)[0]))
        if spec["kind"] == "cloth":
            o64 = build_spec_oracle(oracle, spec, 64)
            o64.step(1, threads=threads)
            spread.append(rel_err(o.get_state(0)[0], o64.get_state(0)[0]))
    errs = np.array(errs)
    kinds = np.array([s["kind"] for s in specs])
    et, ec, spread = errs[kinds == "tet"], errs[kinds == "cloth"], np.array(spread)
    print(f"\n[C4 first frame vs FP32 oracle] tet max {et.max():.2e}; cloth median {np.median(ec):.2e} max "
          f"{ec.max():.2e}, FP32 vs FP64 of the cloth scenes median {np.median(spread):.2e} max {spread.max():.2e}",
          flush=True)
    assert et.max() < TOL, f"tet scene: {et.max():.3e}"
    # An unstressed two-corner sheet's first frame is ill-conditioned (its out-of-plane modes have
    # no stiffness and the reference's StVK Hessian has no PSD projection): the FP32 and FP64 oracles
    # of one scene part by up to ~7e-2. Cloth scenes are held to that spread: the device's median
    # within 2x of the FP32-vs-FP64 median (or 1e-6) and its worst scene within 4x of the worst
    # FP32-vs-FP64 scene (or 1e-5).
    assert np.median(ec) <= max(2.0 * np.median(spread), 1e-6)
    assert ec.max() <= max(4.0 * spread.max(), TOL)

    batch.run_steps(2)
    for i, ((xa, va), (xb, vb)) in enumerate(zip(ref, _states(sims))):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"{BENCH_SCHEDULE}: scene {i}"

    for i, (x, v) in enumerate(init):
        sims[i].setState(0, x, v)
    apply_schedule(batch, parse_schedule_key(QUEUE_SCHEDULE))
    entries, grid = batch.queue_info()
    batch.run_steps(3)
    for i, ((xa, va), (xb, vb)) in enumerate(zip(ref, _states(sims))):
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), f"{QUEUE_SCHEDULE}: scene {i}"
    print(f"[C4 batched == alone] 256 scenes bitwise after 3 frames, {BENCH_SCHEDULE} and {QUEUE_SCHEDULE} "
          f"({entries} queue entries per iteration, grid {grid})", flush=True)
    batch.destroy()
    for s in sims:
        s.destroy()


def test_c4_sheet_error_distribution(gpu_ctx, oracle):
    """The batched workload's two-corner sheet (N = 60, 3,600 vertices; StVK from E = 1e5,
    nu = 0.3, bending 1e-4) over 60 frames from identical state per step. The sheet sags into
    compression, where the exact StVK Hessian (no PSD projection in the reference,
    VBDTriMeshStVK.cpp:772-837) is indefinite and one FP32 step amplifies rounding: the device's
    per-step error against the FP32 oracle is bounded by the FP32 oracle's own error against the
    FP64 one (median within 2x of the FP32-vs-FP64 median, worst step within 4x of the
    FP32-vs-FP64 worst step); 3x3 solve failures are counted on both sides."""
    spec = scenes.c4_template("cloth", 60)
    sim = build_spec_scene(gpu_ctx, spec)
    sim.setSchedule(8, 128, mode="group")
    o32 = build_spec_oracle(oracle, spec, 32)
    o64 = build_spec_oracle(oracle, spec, 64)
    threads = host_threads()
    e_gpu, e_64 = [], []
    for _ in range(60):
        x, v = o32.get_state(0)
        sim.setState(0, x, v)
        o64.set_state(0, x, v)
        sim.runStep()
        o32.step(1, threads=threads)
        o64.step(1, threads=threads)
        xr = o32.get_state(0)[0]
        e_gpu.append(rel_err(sim.getState(0)[0], xr))
        e_64.append(rel_err(xr, o64.get_state(0)[0]))
    e_gpu, e_64 = np.array(e_gpu), np.array(e_64)
    fails = (sim.solveFailures(), o32.solve_failures(), o64.solve_failures())
    print(f"\n[C4 sheet, 60 frames] device vs FP32: median {np.median(e_gpu):.2e} max {e_gpu.max():.2e} "
          f"(step 1 {e_gpu[0]:.2e}, steps over 1e-5: {(e_gpu > TOL).sum()}); FP32 vs FP64: step 1 {e_64[0]:.2e}, median "
          f"{np.median(e_64):.2e} max {e_64.max():.2e}; 3x3 solve failures device/FP32/FP64 {fails}", flush=True)
    # its first step already parts the FP32 and FP64 oracles by ~1e-5: held to 4x that (or 1e-5)
    assert e_gpu[0] <= max(4.0 * e_64[0], TOL)
    assert np.median(e_gpu) <= max(2.0 * np.median(e_64), 1e-6)
    assert e_gpu.max() <= max(4.0 * e_64.max(), TOL)
    sim.destroy()


def _c2_scene(ctx, bodies, phys):
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**phys)
    colors = {}
    for b in bodies:
        key = (b["kind"], len(b["tets"]))
        if key not in colors:
            colors[key] = gaia_amd.color_tet_mesh(len(b["pos"]), b["tets"], "greedy")
        sim.addTetMesh(b["pos"], b["tets"], vertex_colors=colors[key], **b["material"])
    sim.initialize()
    return sim, colors


def _surface_vertices(tets):
    faces = np.concatenate([tets[:, [1, 2, 3]], tets[:, [0, 2, 3]], tets[:, [0, 1, 3]], tets[:, [0, 1, 2]]])
    u, c = np.unique(np.sort(faces, axis=1), axis=0, return_counts=True)
    return np.unique(u[c == 1])


def test_c2_full_size_parity_and_no_interpenetration(gpu_ctx, oracle):
    """32 tet bodies of 8,640 tets (16 cubes, 16 balls) falling in the 2 m box with boundary-surface
    contact (r = 2 mm, stiffness 1e4, friction 0.2, dt = 1/120, 20 iterations). On 7 frames of a
    60-frame fall, among them frames with impacts (more than one detection pass), the device's
    frame from the current state is compared with the FP32 and FP64 oracles' frames from the same
    state: the median over the checked frames within 2x of the FP32-vs-FP64 median (or 1e-6) and
    every frame within 4x of the worst FP32-vs-FP64 frame (or 1e-5). Over all 60 frames no
    boundary vertex of one body lies inside a tet of another."""
    bodies = scenes.c2_bodies(seed=0)
    phys = physics_dict(**scenes.c2_physics())
    sim, colors = _c2_scene(gpu_ctx, bodies, phys)
    nb = len(bodies)
    threads = host_threads()

    def oracle_scene(precision):
        sc = oracle.ClothScene(phys, precision=precision)
        for b in bodies:
            sc.add_tet_mesh(np.asarray(b["pos"], np.float32).astype(np.float64), b["tets"], b["material"],
                            colors[(b["kind"], len(b["tets"]))])
        sc.finalize()
        return sc

    o32, o64 = oracle_scene(32), oracle_scene(64)
    surfaces = [_surface_vertices(b["tets"]) for b in bodies]
    check = {0, 10, 16, 20, 24, 30, 45}
    rows = []
    worst_overlap = []
    for k in range(60):
        if k in check:
            state = [sim.getState(m) for m in range(nb)]
            for m, (x, v) in enumerate(state):
                o32.set_state(m, x, v)
                o64.set_state(m, x, v)
            d0 = sim.contactStats()["detections"]
            sim.runStep()
            detections = sim.contactStats()["detections"] - d0
            o32.step(1, threads=threads)
            o64.step(1, threads=threads)
            e = max(rel_err(sim.getState(m)[0], o32.get_state(m)[0]) for m in range(nb))
            e64 = max(rel_err(o32.get_state(m)[0], o64.get_state(m)[0]) for m in range(nb))
            rows.append((k, detections, e, e64))
        else:
            sim.runStep()
        found = vertices_inside_other_bodies([(sim.getState(m)[0], b["tets"]) for m, b in enumerate(bodies)],
                                             surfaces)
        worst_overlap.extend((k,) + f for f in found)
    for k, d, e, e64 in rows:
        print(f"[C2 frame {k}] detections {d}, device vs FP32 {e:.2e}, FP32 vs FP64 {e64:.2e}", flush=True)
    e = np.array([r[2] for r in rows])
    e64 = np.array([r[3] for r in rows])
    lowest = min(float(sim.getState(m)[0][:, 1].min()) for m in range(nb))
    print(f"[C2] lowest vertex after 60 frames {lowest:.4f} m; vertices inside other bodies: {worst_overlap[:5]}",
          flush=True)
    assert any(r[1] > 1 for r in rows), "no checked frame re-detected (no impact covered)"
    assert not worst_overlap
    assert np.median(e) <= max(2.0 * np.median(e64), 1e-6)
    assert e.max() <= max(4.0 * e64.max(), TOL)
    sim.destroy()


def test_c2_penetration_pileup_detections_equal_oracle(gpu_ctx, oracle):
    """The same 32 bodies with the tet solver's penetration contact (collision stiffness 1e5),
    stepped on the device until they pile up on the floor (frame 33), then 6 frames from identical
    state on the device and the FP32 oracle: every vertex's detection (hit mesh and face,
    contained, from CCD) equal on every frame, thousands of hits among them, and vertices deep
    enough in another body that the closest-face search goes past its nearest-face rounds. The
    positions are stated against the FP32-vs-FP64 spread of the same frames (ill-conditioned)."""
    bodies = scenes.c2_bodies(seed=0)
    phys = physics_dict(**dict(scenes.c2_physics(), tet_contact_model=1))
    sim, colors = _c2_scene(gpu_ctx, bodies, phys)
    nb = len(bodies)
    sim.runSteps(33)
    threads = host_threads()

    def oracle_scene(precision):
        sc = oracle.TetScene(phys, precision=precision)
        for b in bodies:
            sc.add_mesh(np.asarray(b["pos"], np.float32).astype(np.float64), b["tets"], b["material"],
                        colors[(b["kind"], len(b["tets"]))])
        sc.set_collision(dict(stiffness=float(np.float32(phys["collision_stiffness"]))))
        sc.finalize()
        return sc

    o32, o64 = oracle_scene(32), oracle_scene(64)
    sizes = [len(b["pos"]) for b in bodies]
    off = np.cumsum([0] + sizes)
    hits = full = 0
    rows_e = []
    for k in range(6):
        state = [sim.getState(m) for m in range(nb)]
        for m, (x, v) in enumerate(state):
            o32.set_state(m, x, v)
            o64.set_state(m, x, v)
        sim.runStep()
        o32.step(1, threads=threads)
        o64.step(1, threads=threads)
        got, counts = sim.penetrationResults()
        want = got.copy()
        want[:, 2:] = (-1, -1, 0, 0)
        for mesh, v, hm, hf, pen, ccd in o32.collisions():
            want[off[mesh] + v, 2:] = (hm, hf, pen, ccd)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert len(bad) == 0, (k, got[bad[:4]].tolist(), want[bad[:4]].tolist())
        hits += int((want[:, 3] >= 0).sum())
        full += int(counts[4])
        assert counts[3] == 0
        e = max(rel_err(sim.getState(m)[0], o32.get_state(m)[0]) for m in range(nb))
        e64 = max(rel_err(o32.get_state(m)[0], o64.get_state(m)[0]) for m in range(nb))
        rows_e.append((e, e64))
        # the device continues from the oracle's FP32 state
        for m in range(nb):
            sim.setState(m, *o32.get_state(m))
    e, e64 = np.array(rows_e).T
    print(f"\n[C2 penetration pile-up] frames 34-39: {hits} hits, {full} searches past the nearest rounds; "
          f"device vs FP32 max {e.max():.2e}, FP32 vs FP64 max {e64.max():.2e}", flush=True)
    assert hits > 1000 and full > 0
    assert np.all(e <= np.maximum(TOL, 4.0 * e64))


def test_c5_served_equals_alone_under_binding_budget(gpu_ctx):
    """The served population (2,000 scenes of 1k-200k vertices, Zipf(1.2) size classes, cloned from
    62 meshes) under a budget of 25 % of the scenes' device bytes: 40 ticks of seeded random step
    requests (about 50 per tick). Every tick stays within the budget, scenes are copied out and back
    (evictions > 0), and 50 seeded served scenes equal the same scenes stepped alone, bitwise."""
    kinds = scenes.c5_scene_kinds(2000, seed=0)
    protos, sizes = {}, {}
    for key in sorted(set(kinds)):
        spec = scenes.c4_template(*key)
        before = gpu_ctx.memory_bytes()
        probe = build_spec_scene(gpu_ctx, spec)
        sizes[key] = gpu_ctx.memory_bytes() - before
        probe.destroy()
        protos[key] = build_spec_scene(gpu_ctx, spec, defer_upload=True)
    total = sum(sizes[k] for k in kinds)
    budget = max(total // 4, max(sizes.values()))
    sims = [protos[k].clone() for k in kinds]
    srv = Server(gpu_ctx, budget)
    srv.set_schedule(8, 128, "group")
    for s in sims:
        srv.add_scene(s)
    rng = np.random.default_rng(3)
    counts = np.zeros(len(sims), np.int64)
    tag = 0
    seen = []
    for _ in range(40):
        for i in rng.integers(0, len(sims), size=rng.integers(30, 70)):
            srv.request_step(sims[i], tag)
            counts[i] += 1
            tag += 1
        seen.append(srv.tick())
    while srv.stats()["pending"]:
        seen.append(srv.tick())
    done = srv.poll(capacity=tag + 16)
    st = srv.stats()
    print(f"\n[C5 served] {len(sims)} scenes, {total / 1e9:.2f} GB, budget {budget / 1e9:.2f} GB: ticks {st['ticks']} "
          f"served {st['served']} admissions {st['admissions']} evictions {st['evictions']}", flush=True)
    assert len(done) == tag
    assert st["evictions"] > 0
    assert all(t["resident_bytes"] <= budget for t in seen)
    stepped = np.flatnonzero(counts > 0)
    sample = np.random.default_rng(4).choice(stepped, size=50, replace=False)
    for i in sample:
        alone = protos[kinds[i]].clone(defer_upload=False)
        alone.setSchedule(8, 128, mode="group")
        alone.runSteps(int(counts[i]))
        xa, va = alone.getState(0)
        xs, vs = sims[i].getState(0)
        assert np.array_equal(xa, xs) and np.array_equal(va, vs), f"scene {i} ({kinds[i]}), {counts[i]} frames"
        alone.destroy()
    srv.destroy()
    for s in sims + list(protos.values()):
        s.destroy()
