"""Helpers of the parity tests at the benchmark configurations' own sizes: scene builders shared
with bench.py's generators, the host thread count for the oracle, and a vertex-in-tet check
between tet bodies."""
import os

import numpy as np

import gaia_amd
from gaia_amd import VBDPhysics


def host_threads():
    """Threads for the oracle: the CPUs this process may run on, at most 16 (one GPU's share of a
    GPU host)."""
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def build_spec_scene(ctx, spec, defer_upload=False):
    """A scene of the batched workload's generators (gaia_amd.scenes.c4_scenes / c4_template)."""
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**spec["physics"])
    if spec["kind"] == "cloth":
        colors = gaia_amd.color_tri_mesh(len(spec["positions"]), spec["faces"], spec["coloring"])
        sim.addClothMesh(spec["positions"], spec["faces"], fixed_points=spec["fixed"], vertex_colors=colors,
                         **spec["material"])
    else:
        colors = gaia_amd.color_tet_mesh(len(spec["positions"]), spec["tets"], spec["coloring"])
        sim.addTetMesh(spec["positions"], spec["tets"], vertex_colors=colors, **spec["material"])
    sim.initialize(defer_upload=defer_upload)
    return sim


def build_spec_oracle(oracle, spec, precision=32):
    """The same scene in the oracle (FP32 or FP64 restatement of the reference CPU path)."""
    from gaia_amd.vbd import physics_dict

    pos = np.asarray(spec["positions"], np.float32).astype(np.float64)
    if spec["kind"] == "cloth":
        sc = oracle.ClothScene(physics_dict(**spec["physics"]), precision=precision)
        colors = gaia_amd.color_tri_mesh(len(pos), spec["faces"], spec["coloring"])
        sc.add_mesh(pos, spec["faces"], spec["material"], colors, fixed=spec["fixed"])
    else:
        sc = oracle.TetScene(physics_dict(**spec["physics"]), precision=precision)
        colors = gaia_amd.color_tet_mesh(len(pos), spec["tets"], spec["coloring"])
        sc.add_mesh(pos, spec["tets"], spec["material"], colors, fixed=spec.get("fixed", ()))
    sc.finalize()
    return sc


def _tet_boxes(pos, tets):
    corners = pos[tets]  # (T, 4, 3)
    return corners.min(axis=1), corners.max(axis=1), corners


def vertices_inside_other_bodies(bodies, surfaces):
    """Vertex-in-tet check between bodies: for every ordered pair (a, b) of bodies whose bounding
    boxes overlap, the boundary-surface vertices of a that lie inside some tet of b (all four
    barycentric coordinates > 1e-7). `bodies` holds (positions (N, 3), tets (T, 4)) per body,
    `surfaces` the surface vertex ids per body. Returns a list of (a, b, count)."""
    from scipy.spatial import cKDTree

    boxes = [(p.min(axis=0), p.max(axis=0)) for p, _ in bodies]
    cache = {}
    found = []
    for b, (pb, tb) in enumerate(bodies):
        lo_b, hi_b = boxes[b]
        for a, (pa, _) in enumerate(bodies):
            if a == b:
                continue
            lo_a, hi_a = boxes[a]
            if np.any(hi_a < lo_b) or np.any(hi_b < lo_a):
                continue
            pts = pa[surfaces[a]]
            pts = pts[np.all((pts >= lo_b) & (pts <= hi_b), axis=1)]
            if len(pts) == 0:
                continue
            if b not in cache:
                lo, hi, corners = _tet_boxes(pb, tb)
                centroid = corners.mean(axis=1)
                radius = np.linalg.norm(corners - centroid[:, None, :], axis=2).max()
                dm = np.stack([corners[:, 1] - corners[:, 0], corners[:, 2] - corners[:, 0],
                               corners[:, 3] - corners[:, 0]], axis=2)
                cache[b] = (cKDTree(centroid), radius, np.linalg.inv(dm), corners[:, 0], lo, hi)
            tree, radius, inv, origin, lo, hi = cache[b]
            near = tree.query_ball_point(pts, radius)
            pi = np.repeat(np.arange(len(pts)), [len(n) for n in near])
            ti = np.concatenate([np.asarray(n, dtype=np.int64) for n in near]) if len(pi) else np.zeros(0, np.int64)
            if len(ti) == 0:
                continue
            inbox = np.all((pts[pi] >= lo[ti]) & (pts[pi] <= hi[ti]), axis=1)
            pi, ti = pi[inbox], ti[inbox]
            lam = np.einsum("pij,pj->pi", inv[ti], pts[pi] - origin[ti])
            inside = np.minimum(lam.min(axis=1), 1.0 - lam.sum(axis=1)) > 1e-7
            n = len(np.unique(pi[inside]))
            if n:
                found.append((a, b, n))
    return found
