"""Pins of the contact-detector oracle (CPU, no GPU): closest-point primitives against dense
sampling and closed forms, the uniform-grid detector against brute force, the candidate
definition (topology filter, radius) against float64 geometry, feasible regions and the
conservative-bound formula on configurations with known answers."""
import numpy as np
import pytest

from gaia_amd import scenes

R = 0.002


def tri_dense_min(p, a, b, c, n=400):
    """min distance from p to the triangle by dense barycentric sampling (float64)."""
    s = np.linspace(0.0, 1.0, n)
    u, v = np.meshgrid(s, s, indexing="ij")
    m = u + v <= 1.0
    u, v = u[m], v[m]
    q = a[None] + u[:, None] * (b - a)[None] + v[:, None] * (c - a)[None]
    return np.sqrt(((q - p[None]) ** 2).sum(1)).min()


def seg_dense_min(p0, p1, q0, q1, n=600):
    s = np.linspace(0.0, 1.0, n)
    P = p0[None] + s[:, None] * (p1 - p0)[None]
    Q = q0[None] + s[:, None] * (q1 - q0)[None]
    return np.sqrt(((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)).min()


def test_closest_point_regions_known_answers(oracle):
    a, b, c = np.array([0, 0, 0.]), np.array([1, 0, 0.]), np.array([0, 1, 0.])
    cases = [  # point, expected closest, expected type (A B C AB BC AC interior)
        ([-1, -1, 0.5], [0, 0, 0], 0),
        ([2, -0.5, 0], [1, 0, 0], 1),
        ([-0.2, 3, 0], [0, 1, 0], 2),
        ([0.5, -1, 0], [0.5, 0, 0], 3),
        ([1, 1, 0], [0.5, 0.5, 0], 4),
        ([-1, 0.25, 0.1], [0, 0.25, 0], 5),
        ([0.25, 0.25, 0.3], [0.25, 0.25, 0], 6),
    ]
    for p, q, t in cases:
        cp, bary, typ = oracle.closest_point_triangle(np.array(p, np.float32), a, b, c)
        assert typ == t, (p, typ)
        np.testing.assert_allclose(cp, q, atol=1e-7)
        np.testing.assert_allclose(bary[0] * a + bary[1] * b + bary[2] * c, cp, atol=1e-6)


def test_closest_point_matches_dense_sampling(oracle):
    rng = np.random.default_rng(0)
    for _ in range(60):
        a, b, c, p = rng.normal(size=(4, 3)).astype(np.float32)
        cp, bary, _ = oracle.closest_point_triangle(p, a, b, c)
        d = np.linalg.norm(cp.astype(np.float64) - p)
        dd = tri_dense_min(p.astype(np.float64), a.astype(np.float64), b.astype(np.float64), c.astype(np.float64))
        assert d <= dd + 1e-5 and dd - d < 1e-2 * max(1.0, dd)
        assert bary.min() >= -1e-6 and abs(bary.sum() - 1) < 1e-5


def test_segment_closest_points_match_dense_sampling(oracle):
    rng = np.random.default_rng(1)
    for _ in range(60):
        p0, p1, q0, q1 = rng.normal(size=(4, 3)).astype(np.float32)
        cp, cq, s, t = oracle.segment_closest_points(p0, p1, q0, q1)
        assert 0.0 <= s <= 1.0 and 0.0 <= t <= 1.0
        np.testing.assert_allclose(cp, p0 + s * (p1 - p0), atol=1e-5)
        np.testing.assert_allclose(cq, q0 + t * (q1 - q0), atol=1e-5)
        d = np.linalg.norm(cp.astype(np.float64) - cq)
        dd = seg_dense_min(*(x.astype(np.float64) for x in (p0, p1, q0, q1)))
        assert d <= dd + 1e-5 and dd - d < 2e-2 * max(1.0, dd)
    # parallel overlapping segments: distance is the offset, at metre and millimetre scale
    for sc in (1.0, 1e-3):
        cp, cq, s, t = oracle.segment_closest_points(np.array([0, 0, 0.]) * sc, np.array([1, 0, 0.]) * sc,
                                                     np.array([0.5, 0.3, 0]) * sc, np.array([2, 0.3, 0.]) * sc)
        assert abs(np.linalg.norm(cp - cq) - 0.3 * sc) < 1e-6 * sc
    # millimetre crossing segments fall in the reference's parallel branch (det < 1e-5): the
    # points stay on the segments and their distance bounds the true minimum from above
    cp, cq, s, t = oracle.segment_closest_points(np.array([0, 0, 0.]), np.array([2e-3, 0, 0.]),
                                                 np.array([1e-3, -1e-3, 5e-4]), np.array([1e-3, 1e-3, 5e-4]))
    assert np.linalg.norm(cp - cq) >= 5e-4 - 1e-9 and 0 <= s <= 1 and 0 <= t <= 1


def two_sheets(n=6, gap=0.0015, shift=(0.003, 0.0, 0.002), h=0.01):
    """Two horizontal sheets of spacing h, the second `gap` above the first and shifted."""
    p1, f1 = scenes.cloth_sheet(n, h * (n - 1), 0.0)
    p2, f2 = scenes.cloth_sheet(n, h * (n - 1), gap, origin=(shift[0], shift[2]))
    return [(p1.astype(np.float32), f1), (p2.astype(np.float32), f2)]


def crumpled(seed, n=7, h=0.004, amp=0.003):
    """A sheet folded onto itself plus noise: self-proximity within one mesh."""
    rng = np.random.default_rng(seed)
    p, f = scenes.cloth_sheet(n, h * (n - 1), 0.0)
    x = p[:, 0] / (h * (n - 1))
    p = np.stack([np.cos(np.pi * x) * h * (n - 1) / 3, np.sin(np.pi * x) * h * (n - 1) / 3 * 0.2, p[:, 2]], -1)
    p = p + rng.normal(scale=amp, size=p.shape)
    return [(p.astype(np.float32), f)]


@pytest.mark.parametrize("case", ["two_sheets", "crumpled0", "crumpled1", "tori"])
def test_grid_detector_equals_brute_force(oracle, case):
    if case == "two_sheets":
        meshes = two_sheets()
    elif case.startswith("crumpled"):
        meshes = crumpled(int(case[-1]))
    else:
        ms, _ = scenes.c3_tori(seed=2, grid=2, nu=16, nv=8, spacing=0.05, major=0.02, minor=0.008)
        meshes = ms
    D = oracle.Detector(meshes)
    pos = np.concatenate([m[0] for m in meshes])
    r = R if case != "tori" else 0.004
    g = D.detect(pos, r, "grid")
    b = D.detect(pos, r, "brute")
    for k in g:
        assert np.array_equal(g[k], b[k]), k
    assert len(g["vf_pairs"]) > 0 and len(g["ee_contacts"]) > 0


def test_candidate_definition_against_float64(oracle):
    """Every VF pair is a non-incident (vertex, face) closer than r, and every non-incident
    pair clearly closer than r (float64 sampling) is present; no adjacent edges in EE."""
    meshes = crumpled(3)
    D = oracle.Detector(meshes)
    pos = np.concatenate([m[0] for m in meshes]).astype(np.float64)
    faces = meshes[0][1]
    out = D.detect(pos.astype(np.float32), R, "brute")
    got = {tuple(x) for x in out["vf_pairs"]}
    for v in range(len(pos)):
        for f in range(len(faces)):
            if v in faces[f]:
                assert (v, f) not in got
                continue
            dd = tri_dense_min(pos[v], *pos[faces[f]], n=120)
            if dd < R * 0.98:
                assert (v, f) in got, (v, f, dd)
            if dd > R * 1.05:
                assert (v, f) not in got, (v, f, dd)
    edges, _ = D.topology()
    for e1, e2 in out["ee_contacts"]:
        assert e1 != e2 and not set(edges[e1, :2]) & set(edges[e2, :2])
    # EE distances are distances between two points on the segments (never below the true
    # minimum), below the radius. For millimetre edges det = |p|^2 |q|^2 - (p.q)^2 is below the
    # reference's absolute CMP_EPSILON (1e-5), so its parallel-segment branch picks the points
    # and the pair set is not symmetric in (e1, e2); no mirror property is asserted.
    assert len(out["ee_contacts"]) > 20
    for (e1, e2), d in zip(out["ee_contacts"], out["ee_d"]):
        true = seg_dense_min(*pos[edges[e1, :2]], *pos[edges[e2, :2]], n=200)
        assert true <= d + 1e-6 and d < R


def test_two_sheets_min_distances_and_bounds(oracle):
    """Parallel sheets `gap` apart: interior vertices of the upper sheet see the lower sheet at
    exactly `gap`; bounds = relax * (gap - thickness)."""
    gap = 0.0015
    meshes = two_sheets(gap=gap)
    D = oracle.Detector(meshes)
    pos = np.concatenate([m[0] for m in meshes])
    out = D.detect(pos, R, "grid")
    nv1 = len(meshes[0][0])
    upper = pos[nv1:]
    # upper-sheet vertices whose vertical projection is well inside the lower sheet
    inside = np.where((upper[:, 0] > 0.002) & (upper[:, 0] < 0.048) & (upper[:, 2] > 0.002) &
                      (upper[:, 2] < 0.048))[0] + nv1
    assert len(inside) > 0
    np.testing.assert_allclose(out["vf_min"][inside], gap, rtol=1e-5)
    b = D.bounds(0.0005, 0.45)
    # every bound is relax * (min over the vertex's own and its faces'/edges' distances - thickness)
    assert np.all(b <= 0.45 * (R - 0.0005) + 1e-9)
    np.testing.assert_allclose(b[inside], 0.45 * (gap - 0.0005), rtol=1e-4)
    # far apart: everything at the query radius
    far = [(meshes[0][0], meshes[0][1]), (meshes[1][0] + np.float32(1.0), meshes[1][1])]
    D2 = oracle.Detector(far)
    out2 = D2.detect(np.concatenate([m[0] for m in far]), R, "grid")
    assert len(out2["vf_pairs"]) == 0 and np.all(out2["vf_min"] == np.float32(R))
    np.testing.assert_allclose(D2.bounds(0.0, 0.45), 0.45 * R, rtol=1e-6)


def test_feasible_regions(oracle):
    """Metre-scale sheet (the feasible-region margin is an absolute 1e-3): a probe beyond a
    corner gets one vertex contact, a probe beside a boundary edge one edge contact, and a probe
    under a cell one interior contact while the cell's diagonal edge, also within the radius,
    is rejected by its divider planes."""
    p, f = scenes.cloth_sheet(3, 2.0, 0.0)
    probes = [([-0.2, 0.2, -0.2], 0), ([0.5, 0.3, -0.2], 1), ([0.3, -0.2, 0.6], 2)]
    for q, want_type in probes:
        q = np.array([q], np.float32)
        meshes = [(p.astype(np.float32), f), (q, np.zeros((0, 3), np.int32))]
        D = oracle.Detector(meshes)
        out = D.detect(np.concatenate([meshes[0][0], q]), 0.5, "brute")
        v = len(p)
        assert len(out["vf_pairs"][out["vf_pairs"][:, 0] == v]) > 0
        c = out["vf_contacts"][out["vf_contacts"][:, 0] == v]
        assert len(c) == 1 and c[0, 1] == want_type, (q, c)
    # the interior probe has two candidate faces (its cell's triangles), one contact
    q = np.array([[0.3, -0.2, 0.6]], np.float32)
    D = oracle.Detector([(p.astype(np.float32), f), (q, np.zeros((0, 3), np.int32))])
    out = D.detect(np.concatenate([p.astype(np.float32), q]), 0.5, "brute")
    assert (out["vf_pairs"][:, 0] == len(p)).sum() >= 2
