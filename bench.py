"""VBD throughput benchmark: vertex-solves/s at 10 iterations, batched vs plain.

One step = one frame of VBD (initial step, 10 iterations x parallel groups,
velocity update) over every scene of the workload, with all scene data
resident in HBM. `value` is the batched rate: all scenes advanced together, either
by one persistent work-queue launch per iteration or by one launch per parallel
group (the autotuner times both and keeps the faster), the frame captured in a HIP
graph. The plain rate (one scene per launch sequence, each scene its own graph)
is reported beside it, and a served leg ("served": 2000 scenes of 1k-200k vertices,
Poisson step requests continuously batched by the serving loop) reports served
throughput with p50/p99 step latency.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--scenes 256] [--seed 0] [--schedule KEY]
       python bench.py --workload detector | serve | contact   (the detector alone; the served leg
                                                               alone; tet bodies in contact)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
C4_ITERATIONS = 10
SERVE_SCHEDULE = "group:8x128"  # the batched workload's autotuned schedule


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scene(ctx, spec):
    import gaia_amd
    from gaia_amd import VBDPhysics

    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**spec["physics"])
    if spec["kind"] == "cloth":
        colors = gaia_amd.color_tri_mesh(len(spec["positions"]), spec["faces"], spec["coloring"])
        sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"], vertex_colors=colors,
                         **spec["material"])
    else:
        colors = gaia_amd.color_tet_mesh(len(spec["positions"]), spec["tets"], spec["coloring"])
        sim.addTetMesh(spec["positions"], spec["tets"], vertex_colors=colors, **spec["material"])
    sim.initialize()
    return sim


def free_vertices(spec):
    return len(spec["positions"]) - len(spec.get("fixed", []))


def solve_kernel_roofline(ctx, batch, schedule, reps=5):
    """Average duration of the dominant vertex-solve launches (HIP events on the library's
    stream) and their algorithmic bytes (gather model).

    group mode: `reps` iterations, each every parallel group's launch in order (as in a frame, so
    no launch reads data a repeat of itself left in the caches), plus each launch alone repeated
    `reps` times (warm caches) for the balance of the launches;
    queue mode: `reps` persistent launches, each one whole iteration over every group."""
    groups = batch.num_groups()
    work = [batch.group_work(g)[0] for g in range(groups)]
    if schedule["mode"] == "queue":
        batch.apply_initial_step()
        batch.run_iteration(0)
        ctx.synchronize()
        ctx.event_record(2)
        for it in range(1, reps + 1):
            batch.run_iteration(it)
        ctx.event_record(3)
        avg_ms = ctx.event_elapsed_ms(2, 3) / reps
        avg_bytes = float(sum(work))
        kernel = "queue_solve_kernel"
    else:
        for g in range(groups):
            batch.solve_group(g, 0)
        ctx.synchronize()
        ctx.event_record(2)
        for _ in range(reps):
            for g in range(groups):
                batch.solve_group(g, 0)
        ctx.event_record(3)
        avg_ms = ctx.event_elapsed_ms(2, 3) / (groups * reps)
        per = []
        for g in range(groups):
            batch.solve_group(g, 0)
            ctx.synchronize()
            ctx.event_record(2)
            for _ in range(reps):
                batch.solve_group(g, 0)
            ctx.event_record(3)
            per.append(round(ctx.event_elapsed_ms(2, 3) / reps * 1e3, 1))
        avg_bytes = sum(work) / groups
        kernel = "group_solve_kernel"
    gbs = avg_bytes / (avg_ms * 1e-3) / 1e9
    out = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None,
               kernel=kernel, avg_launch_us=round(avg_ms * 1e3, 2),
               algorithmic_bytes_per_launch=int(avg_bytes))
    if schedule["mode"] != "queue":
        # per launch of one iteration: µs when repeated alone (warm caches) and algorithmic MB
        out["launch_us_repeated"] = per
        out["launch_mb"] = [round(w / 1e6, 1) for w in work]
    return out


def host_cpus():
    """The host's CPUs as this process sees them: nproc, the CPUs it may run on (affinity), the
    cgroup CPU quota when one is set, the CPU model (lscpu), and the thread count the CPU baseline
    uses: every CPU available to the process, at most 16 (one GPU's share of the GPU host)."""
    import subprocess

    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((l.split(":", 1)[1].strip() for l in out.splitlines() if l.startswith("Model name")), None)
    except (OSError, subprocess.SubprocessError):
        pass
    if model is None:
        try:
            model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
    usable = affinity if quota is None else max(1, min(affinity, int(quota)))
    return dict(nproc=os.cpu_count(), affinity=affinity, cgroup_quota_cpus=quota, model=model,
                threads=min(usable, 16))


def c1_leg(ctx, threads, steps=300):
    """Config 1 (the plain soft cube: 1,331 vertices, 5,000 tets, greedy colouring, ground only):
    `steps` frames from the initial state on the GPU plain path (one scene, its own HIP graph) and
    on the FP32 oracle at 1 thread and at `threads`; ms per step and vertex-solves/s each."""
    from oracle import pyoracle

    import gaia_amd
    from gaia_amd import VBDPhysics, scenes
    from gaia_amd.vbd import physics_dict

    c = scenes.c1_soft_cube()
    colors = gaia_amd.color_tet_mesh(len(c["positions"]), c["tets"], "greedy")
    solves = len(c["positions"]) * c["physics"]["iterations"] * steps
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**c["physics"])
    sim.addTetMesh(c["positions"], c["tets"], vertex_colors=colors, **c["material"])
    sim.initialize()
    sim.recordVBDSolveGraph(True)
    sim.runStep()  # graph capture, untimed
    x0, v0 = c["positions"], np.zeros_like(c["positions"])
    sim.setState(0, x0, v0)
    ctx.synchronize()
    t = time.perf_counter()
    sim.runSteps(steps)
    ctx.synchronize()
    gpu_s = time.perf_counter() - t
    sim.destroy()
    out = dict(workload=f"soft cube 1,331 vertices / 5,000 tets, greedy colouring (8 colours), 10 iterations, "
                        f"{steps} frames from rest", steps=steps,
               gpu_plain=dict(ms_per_step=round(gpu_s / steps * 1e3, 4), vertex_solves_per_s=round(solves / gpu_s, 1)))
    for n in sorted({1, threads}):
        sc = pyoracle.TetScene(physics_dict(**c["physics"]), precision=32)
        sc.add_mesh(c["positions"], c["tets"], c["material"], colors)
        sc.finalize()
        t = time.perf_counter()
        sc.step(steps, threads=n)
        s = time.perf_counter() - t
        out[f"cpu_{n}_threads"] = dict(ms_per_step=round(s / steps * 1e3, 3), vertex_solves_per_s=round(solves / s, 1))
    return out


def cpu_baseline(specs, threads, budget_s=12.0):
    """The oracle (plain C++ restatement of the reference CPU path, full 9x9
    assembly per tet) on `threads` host threads, over the first scenes of the workload."""
    from oracle import pyoracle

    from gaia_amd.vbd import physics_dict

    import gaia_amd

    solves = 0
    elapsed = 0.0
    used = 0
    for spec in specs:
        if spec["kind"] == "cloth":
            sc = pyoracle.ClothScene(physics_dict(**spec["physics"]), precision=32)
            colors = gaia_amd.color_tri_mesh(len(spec["positions"]), spec["faces"], spec["coloring"])
            sc.add_mesh(spec["positions"].astype(np.float64), spec["faces"], spec["material"], colors,
                        fixed=spec["fixed"])
        else:
            sc = pyoracle.TetScene(physics_dict(**spec["physics"]), precision=32)
            colors = gaia_amd.color_tet_mesh(len(spec["positions"]), spec["tets"], spec["coloring"])
            sc.add_mesh(spec["positions"].astype(np.float64), spec["tets"], spec["material"], colors)
        sc.finalize()
        t0 = time.perf_counter()
        sc.step(1, threads=threads)
        elapsed += time.perf_counter() - t0
        solves += free_vertices(spec) * spec["physics"]["iterations"]
        used += 1
        if elapsed > budget_s:
            break
    return dict(value=round(solves / elapsed, 1), unit="vertex-solves/s", cores=threads, kind="port",
                sample=f"first {used} scenes of the workload, 1 step each, FP32 oracle on {threads} threads",
                seconds=round(elapsed, 2))


def detector_bench(args, rebuild_every=32, sigma=0.001, radius=0.002):
    """Contact detection alone: 64 tori x 16,384 triangles (524,288 vertices, 1,048,576 faces,
    1,572,864 edges); per step the vertices get fresh seeded Gaussian noise (sigma 1 mm) on top
    of a per-torus rigid drift of 5 mm per step (generator, untimed), then the timed detector
    work: BVH refit (a full rebuild every 32 steps, the reference's contactBVHReconstructionSteps)
    and VF + EE queries at r = 2 mm over every vertex and edge."""
    import gaia_amd
    from gaia_amd import scenes
    from gaia_amd.detector import ContactDetector

    meshes, drift = scenes.c3_tori(seed=args.seed)
    ctx = gaia_amd.Context(0)
    cd = ContactDetector(ctx, meshes, max_query_dis=radius)
    log(f"detector: {cd.nv} vertices, {cd.nf} faces, {cd.ne} edges; device memory "
        f"{ctx.memory_bytes() / 2**30:.2f} GiB")
    refit_ms = query_ms = 0.0
    pairs = contacts = 0
    total = args.warmup + args.steps
    for step in range(total):
        cd.perturb(seed=args.seed, step=step, sigma=sigma, drift=drift)
        ctx.event_record(0)
        cd.updateBVH(rebuild=(step % rebuild_every == 0))
        ctx.event_record(1)
        cd.query(record=False)
        ctx.event_record(2)
        ctx.synchronize()
        if step >= args.warmup:
            refit_ms += ctx.event_elapsed_ms(0, 1)
            query_ms += ctx.event_elapsed_ms(1, 2)
            c = cd.counts()
            pairs += c["vf_pairs"]
            contacts += c["ee_contacts"]
    n = args.steps
    step_ms = (refit_ms + query_ms) / n
    # algorithmic bytes of the queries: per vertex its position (12 B), per VF candidate within
    # r the face's 3 positions (36 B) and an 8 B result; per edge its 2 ids and 2 positions
    # (32 B), per EE contact 24 B of the other edge and an 8 B result
    qbytes = cd.nv * 12 + cd.ne * 32 + (pairs / n) * 44 + (contacts / n) * 32
    gbs = qbytes / (query_ms / n * 1e-3) / 1e9
    out = {
        "metric": "contact detection steps/sec (refit + VF + EE queries)",
        "value": round(1e3 / step_ms, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"64 tori x 16,384 triangles (seed {args.seed}), noise sigma 1 mm + 5 mm drift "
                               f"per step, VF + EE proximity queries at r = 2 mm, rebuild every {rebuild_every} "
                               f"steps, refit otherwise",
                   "vertices": cd.nv, "faces": cd.nf, "edges": cd.ne,
                   "refit_ms": round(refit_ms / n, 3), "query_ms": round(query_ms / n, 3),
                   "vf_pairs_per_step": int(pairs / n), "ee_contacts_per_step": int(contacts / n)},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 5), "traffic": None, "kernel": "vf + ee query kernels",
                     "algorithmic_bytes_per_launch": int(qbytes)},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the oracle's uniform-grid detector (same predicates) on the host cores, one step
        from oracle import pyoracle

        threads = host_cpus()["threads"]
        D = pyoracle.Detector(meshes)
        pos = cd.positions()
        t = time.perf_counter()
        D.detect(pos, radius, "grid", threads=threads)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(1.0 / cpu_s, 4), "unit": "steps/s", "cores": threads, "kind": "port",
                               "sample": "one step at the last positions, uniform-grid oracle detector"}
    log(f"detector: {out}")
    print(json.dumps(out), flush=True)
    cd.destroy()
    ctx.close()


def contact_bench(args):
    """Tet bodies in contact: 32 bodies (16 subdivided cubes and 16 tet balls of 8,640 tets each;
    70,304 vertices, 276,480 tets) falling in a 2 m box, dt = 1/120, 20 iterations. Contact model
    "cloth": contact radius 2 mm on every body's boundary surface (VF/FV/EE energies, friction,
    conservative-bound truncation, re-detection after truncating iterations); "penetration": the
    reference tet solver's DCD + CCD with the collision force (stiffness 1e5) and friction 0.2 in
    the solve. One step = one frame of the scene."""
    import gaia_amd
    from gaia_amd import Batch, VBDPhysics, scenes
    from gaia_amd.autotune import parse_schedule_key, schedule_key

    t0 = time.perf_counter()
    bodies = scenes.c2_bodies(seed=args.seed)
    phys = scenes.c2_physics()
    penetration = args.contact_model == "penetration"
    if penetration:
        phys.update(tet_contact_model=1)
    if args.query_radius is not None:
        phys["contact_query_radius"] = args.query_radius
    ctx = gaia_amd.Context(0)
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**phys)
    colors = {}
    for b in bodies:
        key = (b["kind"], len(b["tets"]))
        if key not in colors:
            colors[key] = gaia_amd.color_tet_mesh(len(b["pos"]), b["tets"], "greedy")
        sim.addTetMesh(b["pos"], b["tets"], vertex_colors=colors[key], **b["material"])
    sim.initialize()
    nverts = sum(len(b["pos"]) for b in bodies)
    ntets = sum(len(b["tets"]) for b in bodies)
    iters = phys["iterations"]
    log(f"contact: {len(bodies)} bodies, {nverts} vertices, {ntets} tets, built in {time.perf_counter() - t0:.1f}s; "
        f"device memory {ctx.memory_bytes() / 2**30:.2f} GiB")
    for _ in range(args.warmup):
        sim.runStep()
    ctx.synchronize()
    # schedule: lanes per vertex x block of the group launches, each candidate timed over 3 frames
    # from the same state (the state is restored after every candidate and before the timed run)
    saved = [sim.getState(m) for m in range(len(bodies))]

    def restore():
        for m, (x, v) in enumerate(saved):
            sim.setState(m, x, v)

    if args.schedule:
        cands = [args.schedule]
    else:
        cands = ["group:4x64", "group:4x128", "group:8x64", "group:8x128", "group:8x256", "group:16x128",
                 "group:16x256", "group:8x128:x", "group:16x256:x"]
    tune = {}
    for key in cands:
        c = parse_schedule_key(key)
        sim.setSchedule(c["lanes_per_vertex"], c["block"], mode="group", xcd_blocks=c.get("xcd_blocks", 0))
        restore()
        sim.runStep()
        restore()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            sim.runStep()
        ctx.synchronize()
        tune[key] = round((time.perf_counter() - t) / 3 * 1e3, 3)
    best = min(tune, key=tune.get)
    sched = {"vertices_per_entry": 0, "blocks_per_cu": 0, **parse_schedule_key(best)}
    sim.setSchedule(sched["lanes_per_vertex"], sched["block"], mode="group", xcd_blocks=sched.get("xcd_blocks", 0))
    graph = False
    if not args.no_graph:
        # one HIP graph per frame is a candidate too (the cloth model's re-detection then decided
        # on the device; detection is not timed inside the graph, its time comes from direct frames)
        sim.recordVBDSolveGraph(True)
        restore()
        sim.runStep()
        restore()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            sim.runStep()
        ctx.synchronize()
        tune[best + "+graph"] = round((time.perf_counter() - t) / 3 * 1e3, 3)
        graph = tune[best + "+graph"] < tune[best]
        sim.recordVBDSolveGraph(graph)
    restore()
    log(f"contact autotune (ms/frame): {tune} -> {best}{' + graph' if graph else ''}")
    st0 = sim.contactStats()
    t = time.perf_counter()
    for _ in range(args.steps):
        sim.runStep()
    ctx.synchronize()
    wall = time.perf_counter() - t
    st1 = sim.contactStats()
    n = args.steps
    rate = nverts * iters * n / wall
    detections = (st1["detections"] - st0["detections"]) / n  # counted on the device in graph frames
    if graph:
        # detection time per pass from direct frames after the timed window, times the timed
        # window's passes per frame
        sim.recordVBDSolveGraph(False)
        sd = sim.contactStats()
        for _ in range(3):
            sim.runStep()
        se = sim.contactStats()
        per_pass = (se["detect_ms"] - sd["detect_ms"]) / max(se["detections"] - sd["detections"], 1)
        detect_ms = per_pass * detections
    else:
        detect_ms = (st1["detect_ms"] - st0["detect_ms"]) / n
    lowest = min(float(sim.getState(m)[0][:, 1].min()) for m in range(len(bodies)))
    if penetration:
        rows, counts = sim.penetrationResults()
        found = f"last detection {int(counts[0])} DCD / {int(counts[2])} CCD hits"
    else:
        found = f"last detection {st1['vf_pairs']} VF / {st1['ee_pairs']} EE candidates"
    log(f"contact ({args.contact_model}): {wall / n * 1e3:.2f} ms/step, {rate:.3e} vertex-solves/s, "
        f"{detections:.2f} detections/step ({detect_ms:.2f} ms/step), {found}")

    # the solve kernel's launches (5 iterations of every parallel group in order) at the final state,
    # timed with HIP events on the library stream; algorithmic bytes from the tet gather model (contact
    # gathers of surface vertices are not counted, so the figure is a lower bound)
    batch = Batch(ctx, [sim])
    groups = batch.num_groups()
    work = [batch.group_work(g)[0] for g in range(groups)]
    reps = 5
    for g in range(groups):
        sim.solveParallelGroup(g, 0)
    ctx.synchronize()
    ctx.event_record(2)
    for _ in range(reps):
        for g in range(groups):
            sim.solveParallelGroup(g, 0)
    ctx.event_record(3)
    tot_ms = ctx.event_elapsed_ms(2, 3)
    batch.destroy()
    avg_ms = tot_ms / (groups * reps)
    avg_bytes = sum(work) / groups
    gbs = avg_bytes / (avg_ms * 1e-3) / 1e9
    solve_ms_per_step = avg_ms * groups * iters
    roof = dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                traffic=None, kernel="group_solve_kernel", avg_launch_us=round(avg_ms * 1e3, 2),
                algorithmic_bytes_per_launch=int(avg_bytes))
    if penetration:
        model = ("penetration contact (DCD of surface vertices at each substep's start, CCD of their paths, "
                 "collision stiffness 1e5, friction 0.2), tet and surface LBVHs rebuilt every 16 frames")
    else:
        model = (f"contact radius 2 mm, stiffness 1e4 (quadratic-log), friction 0.2, query radius "
                 f"{phys['contact_query_radius'] * 1000:g} mm, BVH rebuilt every 32 frames")
    out = {
        "metric": "tet-contact VBD vertex-solves/sec at 20 iters",
        "value": round(rate, 1),
        "unit": "vertex-solves/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": round(wall / n * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"tet bodies in contact: 16 cubes + 16 tet balls of 8,640 tets (seed {args.seed}, seeded "
                        f"rotations, mu log-uniform 1e4-1e6, lambda 10 mu) on a 4x4x2 lattice in a 2 m box, dt=1/120, "
                        f"20 iterations, {model}; frames {args.warmup + 1}-{args.warmup + n} of the fall",
            "contact_model": args.contact_model,
            "bodies": len(bodies), "vertices": nverts, "tets": ntets, "iterations": iters,
            "parallel_groups": groups,
            "schedule": dict(sched, key=schedule_key(sched)),
            "autotune_ms_per_frame": tune,
            "detections_per_step": round(detections, 3),
            "detect_ms_per_step": round(detect_ms, 3),
            "solve_ms_per_step": round(solve_ms_per_step, 3),
            "hip_graph": graph,
            **({"dcd_hits_last": int(counts[0]), "ccd_hits_last": int(counts[2])} if penetration else
               {"vf_candidates_last": st1["vf_pairs"], "ee_candidates_last": st1["ee_pairs"]}),
            "lowest_vertex_y": round(lowest, 4),
        },
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        # the oracle (FP32) on the host cores: one frame of the same scene from its initial state
        from oracle import pyoracle

        from gaia_amd.vbd import physics_dict

        threads = host_cpus()["threads"]
        if penetration:
            sc = pyoracle.TetScene(physics_dict(**phys), precision=32)
            for b in bodies:
                sc.add_mesh(np.asarray(b["pos"], np.float32).astype(np.float64), b["tets"], b["material"],
                            colors[(b["kind"], len(b["tets"]))])
            sc.set_collision(dict(stiffness=float(np.float32(phys.get("collision_stiffness", 1e5)))))
        else:
            sc = pyoracle.ClothScene(physics_dict(**phys), precision=32)
            for b in bodies:
                sc.add_tet_mesh(np.asarray(b["pos"], np.float32).astype(np.float64), b["tets"], b["material"],
                                colors[(b["kind"], len(b["tets"]))])
        sc.finalize()
        t = time.perf_counter()
        sc.step(1, threads=threads)
        cpu_s = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(nverts * iters / cpu_s, 1), "unit": "vertex-solves/s",
                               "cores": threads, "kind": "port",
                               "sample": "the first frame of the same scene, FP32 oracle (detection included)"}
    log(f"contact: {out}")
    sim.destroy()
    ctx.close()
    return out


def serve_bench(args, schedule=None):
    """Served VBD: Poisson-arriving step requests over a Zipf-sized population of scenes,
    continuously batched by the native serving loop (one frame of every ready scene per tick).
    Returns the result line (a dict)."""
    import gaia_amd
    from gaia_amd import Server, VBDPhysics, scenes
    from gaia_amd.autotune import parse_schedule_key, schedule_key

    t0 = time.perf_counter()
    kinds = scenes.c5_scene_kinds(args.serve_scenes, seed=args.seed)
    templates = {}
    for key in sorted(set(kinds)):
        spec = scenes.c4_template(*key)
        if spec["kind"] == "cloth":
            colors = gaia_amd.color_tri_mesh(len(spec["positions"]), spec["faces"], spec["coloring"])
        else:
            colors = gaia_amd.color_tet_mesh(len(spec["positions"]), spec["tets"], spec["coloring"])
        templates[key] = (spec, colors)
    ctx = gaia_amd.Context(0)
    # one prepared (not uploaded) scene per mesh; every served scene is a clone of one of them
    protos = {}
    for key, (spec, colors) in templates.items():
        sim = VBDPhysics(ctx)
        sim.setPhysicsParams(**spec["physics"])
        if spec["kind"] == "cloth":
            sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"], vertex_colors=colors,
                             **spec["material"])
        else:
            sim.addTetMesh(spec["positions"], spec["tets"], vertex_colors=colors, **spec["material"])
        sim.initialize(defer_upload=True)
        protos[key] = sim
    sims = [protos[key].clone() for key in kinds]
    free = np.array([free_vertices(templates[k][0]) for k in kinds], dtype=np.int64)
    iters = C4_ITERATIONS
    log(f"registered {len(sims)} scenes ({len(templates)} meshes), {free.sum()} free vertices, "
        f"in {time.perf_counter() - t0:.1f}s")
    budget = int(args.budget_gb * 1e9)
    srv = Server(ctx, budget)
    sched = {"vertices_per_entry": 0, "blocks_per_cu": 0,
             **(schedule or parse_schedule_key(args.schedule or SERVE_SCHEDULE))}
    if sched["mode"] != "group":  # the served plan changes every tick: group launches
        sched = {"vertices_per_entry": 0, "blocks_per_cu": 0, **parse_schedule_key(SERVE_SCHEDULE)}
    srv.set_schedule(sched["lanes_per_vertex"], sched["block"], sched["mode"], sched["vertices_per_entry"],
                     sched["blocks_per_cu"], stage_records=sched.get("stage_records", 0),
                     xcd_blocks=sched.get("xcd_blocks", 0))
    for sim in sims:
        srv.add_scene(sim)
    # warm-up (untimed): one frame of every scene, which makes each resident on its first request
    t = time.perf_counter()
    for i, sim in enumerate(sims):
        srv.request_step(sim, -1 - i)
    while srv.stats()["pending"]:
        srv.tick()
    srv.poll(capacity=len(sims) + 16)
    warm = srv.stats()
    log(f"warm-up: {len(sims)} scenes admitted and stepped in {time.perf_counter() - t:.1f}s; resident "
        f"{warm['resident_bytes'] / 1e9:.2f} GB of a {warm['budget_bytes'] / 1e9:.1f} GB budget, "
        f"{warm['evictions']} evictions")

    def timed(req_rate, R, seed):
        """R Poisson step requests at req_rate per second, served as they arrive: wall time,
        latencies from arrival and from submission, per-tick scene counts, vertex solves."""
        rng = np.random.default_rng(seed)
        arrivals = np.cumsum(rng.exponential(1.0 / req_rate, size=R))
        targets = rng.integers(0, len(sims), size=R)
        submit = np.full(R, np.nan)
        native = np.full(R, np.nan)
        stepped = []
        submitted = completed = 0
        ctx.synchronize()
        t0 = time.perf_counter()
        while completed < R:
            now = time.perf_counter() - t0
            while submitted < R and arrivals[submitted] <= now:
                srv.request_step(sims[targets[submitted]], submitted)
                submit[submitted] = time.perf_counter() - t0
                submitted += 1
            if submitted > completed:
                st = srv.tick()
                if st["last_stepped"]:
                    stepped.append(st["last_stepped"])
                recs = srv.poll(capacity=R)
                native[recs["tag"]] = recs["latency_ms"]
                completed += len(recs)
            else:
                time.sleep(max(0.0, arrivals[submitted] - now))
        wall = time.perf_counter() - t0
        # from arrival: the wait for the host loop to post the request, then the server's own record
        # (request to the completion of its frame, timed by the frame's completion event)
        lat = (submit - arrivals) * 1e3 + native
        return wall, lat, native, stepped, int(free[targets].sum()) * iters, float(free[targets].mean()) * iters

    pct = lambda a: {"p50": round(float(np.percentile(a, 50)), 3), "p99": round(float(np.percentile(a, 99)), 3),
                     "mean": round(float(a.mean()), 3), "max": round(float(a.max()), 3)}
    R = args.requests
    wall, lat, native, stepped, solves, mean_solves = timed(args.rate, R, args.seed + 7)
    rate = solves / wall
    stats = srv.stats()
    offered = mean_solves * args.rate
    # a saturating point: requests offered faster than the device can step them, so the served
    # rate is the loop's capacity (latency then measures the queue, not the service)
    sat = None
    if args.saturation_factor > 0:
        sat_rate = args.rate * args.saturation_factor
        sw, slat, _, sstep, ssolves, smean = timed(sat_rate, R, args.seed + 11)
        sat = {"request_rate_per_s": sat_rate, "offered_vertex_solves_per_s": round(smean * sat_rate, 1),
               "served_vertex_solves_per_s": round(ssolves / sw, 1), "latency_ms": pct(slat),
               "scenes_per_tick_mean": round(float(np.mean(sstep)), 2), "ticks": len(sstep)}
        log(f"serve, saturating: {sat}")
    out = {
        "metric": "served VBD vertex-solves/sec at 10 iters; step latency p50/p99 ms",
        "value": round(rate, 1),
        "unit": "vertex-solves/s",
        "n_gpus": 1,
        "steps": R,
        "warmup": 1,
        "ms_per_step": round(wall / max(len(stepped), 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"served VBD: {len(sims)} scenes of 1k-200k vertices (32 log-spaced size classes, Zipf(1.2) "
                        f"over class rank, seed {args.seed}), half neo-Hookean tet cubes and half StVK+bending "
                        f"sheets, dt=1/60, 10 iterations; {R} step requests, Poisson at {args.rate:g}/s, uniform "
                        f"over scenes; one batched frame of every ready scene per tick",
            "scenes": len(sims),
            "vertices": int(free.sum()),
            "request_rate_per_s": args.rate,
            "requests": R,
            "budget_gb": round(stats["budget_bytes"] / 1e9, 1),
            "schedule": dict(sched, key=schedule_key(sched)),
        },
        "latency_ms": pct(lat),
        "latency_from_submission_ms": pct(native),
        "offered_vertex_solves_per_s": round(offered, 1),
        "ticks": len(stepped),
        "scenes_per_tick_mean": round(float(np.mean(stepped)), 2),
        "resident_gb": round(stats["resident_bytes"] / 1e9, 3),
        "admissions": stats["admissions"],
        "evictions": stats["evictions"],
        "saturation": sat,
    }
    log(f"serve: {out}")
    srv.destroy()
    for sim in sims + list(protos.values()):
        sim.destroy()
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scenes", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the served leg of the default run")
    ap.add_argument("--schedule", default=None,
                    help="fixed schedule instead of autotuning: group:LxB or queue:LxB:eN[:gM] "
                         "(lanes per vertex x block, N vertices per queue entry, M blocks per CU)")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30, help="size of the copy-bandwidth calibration")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying HIP graphs (profiler runs)")
    ap.add_argument("--workload", choices=["batched", "detector", "serve", "contact"], default="batched",
                    help="batched VBD (default), the contact detector alone on 64 tori, served VBD, or tet "
                         "bodies in contact (32 bodies of 8,640 tets)")
    ap.add_argument("--contact-model", choices=["cloth", "penetration"], default="cloth",
                    help="contact workload: the cloth contact model on the bodies' surfaces, or the reference tet "
                         "solver's penetration contact (DCD + CCD)")
    ap.add_argument("--query-radius", type=float, default=None,
                    help="tet-contact workload: contact query radius in m (default 0.003, 1.5 contact radii)")
    ap.add_argument("--serve-scenes", type=int, default=2000)
    ap.add_argument("--saturation-factor", type=float, default=3.0,
                    help="served VBD: a second timed window at this multiple of --rate (0: none)")
    ap.add_argument("--requests", type=int, default=20000, help="served VBD: step requests in the timed window")
    ap.add_argument("--rate", type=float, default=20000.0, help="served VBD: Poisson request rate per second")
    ap.add_argument("--budget-gb", type=float, default=288.0,
                    help="served VBD: device-memory budget in GB (clamped to the device's memory)")
    args = ap.parse_args()
    use_graph = not args.no_graph
    if args.gpus != 1:
        print("bench.py measures one GPU: every scene of the workload runs on one device; "
              "multi-GPU runs are not part of this build (scenes would be sharded whole per GPU).",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "detector":
        detector_bench(args)
        return
    if args.workload == "serve":
        print(json.dumps(serve_bench(args)), flush=True)
        return
    if args.workload == "contact":
        print(json.dumps(contact_bench(args)), flush=True)
        return

    import gaia_amd
    from gaia_amd import Batch, scenes
    from gaia_amd.autotune import apply_schedule, autotune, default_candidates, parse_schedule_key, schedule_key

    t0 = time.perf_counter()
    specs = scenes.c4_scenes(args.seed, args.scenes)
    ctx = gaia_amd.Context(0)
    sims = [build_scene(ctx, s) for s in specs]
    # vertex-solves count free (solved) vertices only
    nverts = sum(free_vertices(s) for s in specs)
    ncloth = sum(1 for s in specs if s["kind"] == "cloth")
    iters = specs[0]["physics"]["iterations"]
    log(f"built {len(sims)} scenes ({len(sims) - ncloth} tet, {ncloth} cloth), {nverts} free vertices, in "
        f"{time.perf_counter() - t0:.1f}s; device memory {ctx.memory_bytes() / 2**30:.2f} GiB")

    # ---- batched schedule: autotuned (or given) for the batch of all scenes
    batch = Batch(ctx, sims)
    if args.schedule:
        apply_schedule(batch, parse_schedule_key(args.schedule), use_graph)
        sched, tune = batch.get_schedule(), {}
    else:
        sched, tune = autotune(ctx, batch, default_candidates(), use_graph=use_graph)
    key = schedule_key(sched)
    entries, grid = batch.queue_info()
    log(f"autotune (ms/frame): {tune} -> {key}" + (f" ({entries} queue entries per iteration, grid {grid})"
                                                     if sched["mode"] == "queue" else ""))

    # ---- plain: one scene per launch sequence (each scene replays its own graph), with the
    # batched schedule's lanes per vertex and block size
    plain_rate = None
    if not args.no_plain:
        for s in sims:
            s.setSchedule(sched["lanes_per_vertex"], sched["block"], mode="group",
                          stage_records=sched.get("stage_records", 0), xcd_blocks=sched.get("xcd_blocks", 0))
            s.recordVBDSolveGraph(use_graph)
        for _ in range(args.warmup):
            for s in sims:
                s.runStep()
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            for s in sims:
                s.runStep()
        ctx.synchronize()
        plain_s = time.perf_counter() - t
        plain_rate = nverts * iters * args.steps / plain_s
        log(f"plain: {plain_s / args.steps * 1e3:.2f} ms/step, {plain_rate:.3e} vertex-solves/s")

    # ---- batched: all scenes together, one graph per frame
    for _ in range(args.warmup):
        batch.run_steps(1)
    ctx.synchronize()
    t = time.perf_counter()
    batch.run_steps(args.steps)
    ctx.synchronize()
    batched_s = time.perf_counter() - t
    rate = nverts * iters * args.steps / batched_s
    log(f"batched: {batched_s / args.steps * 1e3:.2f} ms/step, {rate:.3e} vertex-solves/s")

    roof = solve_kernel_roofline(ctx, batch, sched)
    # measured HBM bytes per launch (scaled PMC FETCH_SIZE + WRITE_SIZE) for this kernel,
    # workload and schedule, from the committed profile summary
    side = os.path.join(REPO, "profiles", f"{roof['kernel']}_batched_256_scenes.traffic.json")
    if os.path.exists(side) and args.scenes == 256 and args.seed == 0:
        t = json.load(open(side)).get("schedules", {}).get(key)
        if t:
            roof["traffic"] = t["hbm_bytes_per_launch"]
    # achievable-bandwidth reference points: the library's streaming copy kernel and the runtime's
    # device-to-device copy (each moves 2 x copy_bytes per copy)
    copy_ms = ctx.stream_copy_ms(args.copy_bytes, reps=5)
    memcpy_ms = ctx.stream_copy_ms(args.copy_bytes, reps=5, method=1)
    roof["measured_copy_gbs"] = round(2 * args.copy_bytes / (copy_ms * 1e-3) / 1e9, 1)
    roof["measured_memcpy_gbs"] = round(2 * args.copy_bytes / (memcpy_ms * 1e-3) / 1e9, 1)
    log(f"roofline: {roof}")
    cpu = None
    if not args.no_cpu_baseline:
        # the FP32 oracle on the host's cores (every CPU available to the process, at most 16) and on
        # one thread, over a bounded sample of the workload, and config 1 over its 300 frames
        host = host_cpus()
        cpu = cpu_baseline(specs, host["threads"])
        one = cpu_baseline(specs, 1, budget_s=10.0)
        cpu["one_thread"] = dict(value=one["value"], sample=one["sample"])
        cpu["host"] = host
        cpu["c1"] = c1_leg(ctx, host["threads"])
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "VBD vertex-solves/sec at 10 iters (batched vs plain); served step p50/p99 ms",
        "value": round(rate, 1),
        "unit": "vertex-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched_s / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched VBD: {len(specs)} scenes of log-uniform 500-50k vertices (seed {args.seed}), "
                        f"{len(specs) - ncloth} neo-Hookean tet cubes + {ncloth} StVK+bending cloth sheets pinned "
                        f"at two corners, dt=1/60, 10 iterations, "
                        + ("one persistent work-queue launch per iteration" if sched["mode"] == "queue"
                           else "one launch per parallel group") + ", HIP graph per frame",
            "scenes": len(specs),
            "cloth_scenes": ncloth,
            "vertices": nverts,
            "iterations": iters,
            "schedule": dict(sched, key=key, queue_entries=entries if sched["mode"] == "queue" else None,
                             grid=grid if sched["mode"] == "queue" else None),
            "autotune_ms_per_frame": tune,
        },
        "plain_value": None if plain_rate is None else round(plain_rate, 1),
        "batched_over_plain": None if plain_rate is None else round(rate / plain_rate, 2),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    batch.destroy()
    for s in sims:
        s.destroy()
    ctx.close()
    if not args.no_serve:
        # served leg: the served population under Poisson step requests, batched schedule
        sv = serve_bench(args, schedule=sched)
        out["served"] = {k: sv[k] for k in ("value", "unit", "latency_ms", "offered_vertex_solves_per_s",
                                            "scenes_per_tick_mean", "ticks", "resident_gb", "saturation")}
        out["served"]["workload"] = sv["config"]["workload"]
        out["served"]["schedule"] = sv["config"]["schedule"]["key"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
