"""Schedule autotuner for batched VBD: times whole graph-captured frames on the device
(HIP events on the library's stream) for each candidate schedule and keeps the fastest.

Knobs: launch mode (one launch per parallel group, or one persistent work-queue launch per
iteration), lanes per vertex (vertices per wavefront = 64 / lanes), block size, for group
launches the LDS staging depth (incidence records staged per chunk, 0 = read from HBM by the
lanes) and the XCD block mapping (each XCD's blocks take one contiguous range of the work
items), and for the work queue the entry size (vertices per entry) and the persistent grid per CU.
"""
import itertools

DEFAULT_GROUP = [(4, 256), (8, 256), (16, 256), (4, 128), (8, 128), (16, 128)]


def default_candidates(modes=("group", "queue")):
    out = []
    if "group" in modes:
        out += [dict(lanes_per_vertex=l, block=b, mode="group") for l, b in DEFAULT_GROUP]
        out += [dict(lanes_per_vertex=l, block=b, mode="group", xcd_blocks=x)
                for (l, b), x in itertools.product([(8, 256), (8, 128)], [1, 4, 16, 64])]
        for (l, b), st in itertools.product([(4, 128), (8, 128), (8, 256), (16, 256)], [32, 64]):
            out.append(dict(lanes_per_vertex=l, block=b, mode="group", stage_records=st))
    if "queue" in modes:
        for (l, b), passes in itertools.product([(4, 256), (8, 256), (8, 128)], [4, 16, 64]):
            out.append(dict(lanes_per_vertex=l, block=b, mode="queue", vertices_per_entry=passes * (b // l)))
    return out


def schedule_key(s):
    key = f"{s['mode']}:{s['lanes_per_vertex']}x{s['block']}"
    if s["mode"] == "group" and s.get("stage_records", 0):
        key += f":s{s['stage_records']}"
    if s["mode"] == "group" and s.get("xcd_blocks", 0):
        key += ":x" if s["xcd_blocks"] == 1 else f":x{s['xcd_blocks']}"
    if s["mode"] == "queue":
        key += f":e{s.get('vertices_per_entry', 0)}"
        if s.get("blocks_per_cu", 0):
            key += f":g{s['blocks_per_cu']}"
    return key


def parse_schedule_key(key):
    """Inverse of schedule_key: "group:8x128", "group:8x128:s64" (LDS staging, 64 records per
    chunk), "group:8x256:x" (XCD-contiguous blocks), "group:8x256:x16" (chunks of 16 items dealt to
    the XCDs in turn), "queue:8x128:e64", "queue:8x128:e64:g2"; a bare
    "8x128" means group launches."""
    parts = key.split(":")
    if len(parts) == 1:
        parts = ["group"] + parts
    mode = parts[0]
    if mode not in ("group", "queue"):
        raise ValueError(f"bad schedule {key!r}")
    l, b = (int(x) for x in parts[1].split("x"))
    s = dict(lanes_per_vertex=l, block=b, mode=mode)
    for p in parts[2:]:
        if p[0] == "e":
            s["vertices_per_entry"] = int(p[1:])
        elif p[0] == "g":
            s["blocks_per_cu"] = int(p[1:])
        elif p[0] == "s":
            s["stage_records"] = int(p[1:])
        elif p[0] == "x":
            s["xcd_blocks"] = int(p[1:]) if len(p) > 1 else 1
        else:
            raise ValueError(f"bad schedule {key!r}")
    return s


def apply_schedule(batch, s, use_graph=True):
    batch.set_schedule(s["lanes_per_vertex"], s["block"], mode=s["mode"],
                       vertices_per_entry=s.get("vertices_per_entry", 0), blocks_per_cu=s.get("blocks_per_cu", 0),
                       stage_records=s.get("stage_records", 0), xcd_blocks=s.get("xcd_blocks", 0))
    batch.record_graph(use_graph)


def autotune(ctx, batch, candidates=None, reps=2, use_graph=True, event_slots=(0, 1)):
    """Runs one warm-up frame and `reps` timed frames per candidate; leaves the batch on the
    fastest schedule (graph re-recorded) and returns (best schedule dict, {key: ms per frame}).
    Every candidate advances the batch's scenes by reps + 1 frames."""
    candidates = candidates or default_candidates()
    best, results = None, {}
    a, b = event_slots
    for s in candidates:
        apply_schedule(batch, s, use_graph)
        batch.run_steps(1)
        ctx.synchronize()
        ctx.event_record(a)
        batch.run_steps(reps)
        ctx.event_record(b)
        ms = ctx.event_elapsed_ms(a, b) / reps
        results[schedule_key(s)] = round(ms, 3)
        if best is None or ms < best[0]:
            best = (ms, s)
    apply_schedule(batch, best[1], use_graph)
    ctx.synchronize()
    return batch.get_schedule(), results
