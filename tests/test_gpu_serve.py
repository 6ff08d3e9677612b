"""Serving loop on the GPU: step requests for many scenes continuously batched into one frame
per tick, scenes admitted to HBM under a byte budget and copied out to host memory when idle.
Every served result must equal the same scene stepped alone (bitwise: the batched kernels keep
each scene's own colour order and arithmetic)."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import Batch, Server, VBDPhysics, scenes
from vbd_cases import contact_drape_spec, build_gpu

pytestmark = pytest.mark.gpu


def build(ctx, spec, defer_upload=False):
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**spec["physics"])
    if spec["kind"] == "cloth":
        sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"], coloring=spec["coloring"],
                         **spec["material"])
    else:
        sim.addTetMesh(spec["positions"], spec["tets"], coloring=spec["coloring"], **spec["material"])
    sim.initialize(defer_upload=defer_upload)
    return sim


def scene_bytes(ctx, spec):
    before = ctx.memory_bytes()
    s = build(ctx, spec)
    b = ctx.memory_bytes() - before
    s.destroy()
    return b


def run_served(ctx, specs, requests, budget, schedule=None):
    """Serves `requests` (a list of per-tick lists of scene indices) and returns the served
    scenes, the completed records and the stats seen after every tick."""
    sims = [build(ctx, sp, defer_upload=True) for sp in specs]
    srv = Server(ctx, budget)
    if schedule:
        srv.set_schedule(*schedule)
    for s in sims:
        srv.add_scene(s)
    tag = 0
    seen = []
    for batch in requests:
        for i in batch:
            srv.request_step(sims[i], tag)
            tag += 1
        seen.append(srv.tick())
    while srv.stats()["pending"]:
        seen.append(srv.tick())
    done = srv.poll(capacity=tag + 16)
    return sims, srv, done, seen


def random_requests(n_scenes, ticks, per_tick, seed):
    rng = np.random.default_rng(seed)
    return [list(rng.integers(0, n_scenes, size=rng.integers(0, per_tick + 1))) for _ in range(ticks)]


@pytest.mark.parametrize("budget_fraction", [2.0, 0.35])
def test_served_equals_alone_bitwise(gpu_ctx, budget_fraction):
    """Twelve tet and cloth scenes of uneven sizes, requests arriving at random (several for
    one scene in one tick, none for others); with a budget of 35% of the scenes' bytes, scenes
    are copied out and back between their frames. Each scene's state equals the scene stepped
    alone for the same number of frames."""
    specs = scenes.c4_scenes(seed=11, num_scenes=12, vmin=300, vmax=4000)
    total = sum(scene_bytes(gpu_ctx, sp) for sp in specs)
    biggest = max(scene_bytes(gpu_ctx, sp) for sp in specs)
    budget = max(int(budget_fraction * total), biggest)
    reqs = random_requests(len(specs), ticks=30, per_tick=6, seed=5)
    sims, srv, done, seen = run_served(gpu_ctx, specs, reqs, budget)
    counts = np.bincount(np.concatenate([np.asarray(r, int) for r in reqs if len(r)]), minlength=len(specs))
    assert len(done) == counts.sum() and np.all(done["frame"] > 0)
    assert all(st["resident_bytes"] <= budget for st in seen)
    st = srv.stats()
    print(f"\n[serve budget {budget_fraction}] ticks {st['ticks']} served {st['served']} admissions "
          f"{st['admissions']} evictions {st['evictions']}", flush=True)
    if budget_fraction < 1:
        assert st["evictions"] > 0
    for i, sp in enumerate(specs):
        alone = build(gpu_ctx, sp)
        alone.runSteps(int(counts[i]))
        xa, va = alone.getState(0)
        xs, vs = sims[i].getState(0)  # read from host memory when the scene is copied out
        assert np.array_equal(xa, xs) and np.array_equal(va, vs), f"scene {i}"
        alone.destroy()
    srv.destroy()
    for s in sims:
        s.destroy()


def test_frames_overlap_two_deep(gpu_ctx):
    """A tick queues its frame behind the one still running and completes only the older one:
    a scene with two requests gets both frames launched in consecutive ticks."""
    specs = scenes.c4_scenes(seed=3, num_scenes=2, vmin=500, vmax=800)
    sims = [build(gpu_ctx, sp) for sp in specs]  # uploaded: resident from registration
    srv = Server(gpu_ctx, 1 << 34)
    for s in sims:
        srv.add_scene(s)
    srv.request_step(sims[0], 0)
    srv.request_step(sims[0], 1)
    srv.request_step(sims[1], 2)
    st = srv.tick()
    assert st["last_stepped"] == 2 and st["pending"] == 3 and len(srv.poll()) == 0
    st = srv.tick()  # scene 0's second frame queued, the first tick's frame completed
    assert st["last_stepped"] == 1 and st["pending"] == 1
    assert sorted(srv.poll()["tag"].tolist()) == [0, 2]
    st = srv.tick()  # nothing new: drains
    assert st["last_stepped"] == 0 and st["pending"] == 0
    assert srv.poll()["tag"].tolist() == [1]
    alone = build(gpu_ctx, specs[0])
    alone.runSteps(2)
    assert np.array_equal(alone.getState(0)[0], sims[0].getState(0)[0])
    srv.destroy()
    for s in sims + [alone]:
        s.destroy()


def test_requests_of_a_scene_complete_in_order_one_per_tick(gpu_ctx):
    specs = scenes.c4_scenes(seed=2, num_scenes=4, vmin=300, vmax=2000)
    reqs = [[0, 0, 0, 1], [2], [], [0, 3, 3]]
    sims, srv, done, seen = run_served(gpu_ctx, specs, reqs, 1 << 34)
    for i in range(4):
        mine = done[done["scene"] == sims[i].scene]
        assert list(mine["frame"]) == list(range(1, len(mine) + 1))
        assert np.all(np.diff(mine["tag"]) > 0)
    assert np.all(done["latency_ms"] >= done["queue_ms"]) and np.all(done["queue_ms"] >= 0)
    # scene 0 has three requests queued in tick 0 and one in tick 3: four ticks at least
    assert seen[0]["last_stepped"] == 0  # nothing resident yet: the first tick only admits
    assert srv.stats()["served"] == 8
    srv.destroy()
    for s in sims:
        s.destroy()


def test_least_recently_stepped_idle_scene_leaves_first(gpu_ctx):
    specs = scenes.c4_scenes(seed=4, num_scenes=3, vmin=2000, vmax=2000)
    sizes = [scene_bytes(gpu_ctx, sp) for sp in specs]
    budget = sizes[0] + sizes[1] + sizes[2] // 2  # two of the three fit
    sims, srv, done, _ = run_served(gpu_ctx, specs, [[0], [1], [1]], budget)
    assert srv.resident(sims[0]) and srv.resident(sims[1]) and not srv.resident(sims[2])
    srv.request_step(sims[2], 99)
    srv.tick()  # admits scene 2: scene 0 (stepped least recently) is copied out
    assert not srv.resident(sims[0]) and srv.resident(sims[1]) and srv.resident(sims[2])
    srv.tick()  # launches scene 2's frame
    assert srv.stats()["pending"] == 1  # in flight until a later tick completes it
    srv.tick()
    assert srv.stats()["evictions"] == 1 and srv.stats()["pending"] == 0
    # a scene held in host memory cannot be stepped outside the server, but can be read
    with pytest.raises(gaia_amd.VBDError, match="not resident"):
        sims[0].runStep()
    assert np.all(np.isfinite(sims[0].getState(0)[0]))
    srv.remove_scene(sims[0])  # handed back resident
    sims[0].runStep()
    srv.destroy()
    for s in sims:
        s.destroy()


def test_server_rejections(gpu_ctx):
    specs = scenes.c4_scenes(seed=1, num_scenes=2, vmin=500, vmax=600)
    a = build(gpu_ctx, specs[0], defer_upload=True)
    srv = Server(gpu_ctx, 1 << 34)
    srv.add_scene(a)
    with pytest.raises(gaia_amd.VBDError, match="already registered"):
        srv.add_scene(a)
    with pytest.raises(gaia_amd.VBDError, match="registered with a server"):
        a.destroy()
    other = dict(specs[1], physics=dict(specs[1]["physics"], iterations=3))
    b = build(gpu_ctx, other, defer_upload=True)
    with pytest.raises(gaia_amd.VBDError, match="share iterations"):
        srv.add_scene(b)
    c = build_gpu(gpu_ctx, contact_drape_spec(steps=1, iterations=10))
    srv.add_scene(c)  # contact scenes are served with their detector resident
    srv.remove_scene(c)
    d = build(gpu_ctx, specs[1])
    srv.add_scene(d)
    with pytest.raises(gaia_amd.VBDError, match="registered with a server"):
        Batch(gpu_ctx, [d, c])
    srv.destroy()
    for s in (a, b, c, d):
        s.destroy()


def test_scene_larger_than_the_budget_is_rejected(gpu_ctx):
    specs = scenes.c4_scenes(seed=7, num_scenes=2, vmin=3000, vmax=3000)
    small = scenes.c4_scenes(seed=8, num_scenes=1, vmin=300, vmax=300)[0]
    budget = scene_bytes(gpu_ctx, small) * 2
    big = build(gpu_ctx, specs[0], defer_upload=True)
    ok = build(gpu_ctx, small, defer_upload=True)
    srv = Server(gpu_ctx, budget)
    srv.add_scene(big)
    srv.add_scene(ok)
    srv.request_step(big, 1)
    srv.request_step(ok, 2)
    with pytest.raises(gaia_amd.VBDError, match="budget"):
        srv.tick()
    while srv.stats()["pending"]:
        srv.tick()
    done = srv.poll()
    rec = {int(t): int(f) for t, f in zip(done["tag"], done["frame"])}
    assert rec == {1: -1, 2: 1}
    with pytest.raises(gaia_amd.VBDError, match="budget"):
        srv.request_step(big, 3)
    srv.destroy()
    big.destroy()
    ok.destroy()


def test_clone_equals_original(gpu_ctx):
    """A clone (host preparation copied) steps exactly like a scene built from the same input."""
    sp = scenes.c4_scenes(seed=9, num_scenes=2, vmin=800, vmax=1500)
    for spec in sp:
        proto = build(gpu_ctx, spec, defer_upload=True)
        a = proto.clone(defer_upload=False)
        b = build(gpu_ctx, spec)
        a.runSteps(3)
        b.runSteps(3)
        xa, va = a.getState(0)
        xb, vb = b.getState(0)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb)
        for s in (proto, a, b):
            s.destroy()
