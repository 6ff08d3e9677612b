"""Pins of the CPU oracle: reference fixtures, closed forms, invariants, finite differences.

Each check is chosen so that a plausible restatement mistake (dropped term,
wrong sign or index, transposed operand) fails it.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_t

FIXTURES = ["Beam_lowRes", "Beam", "Cube", "bunny_small"]


def categories(colors):
    return [sorted(np.flatnonzero(colors == c).tolist()) for c in range(int(colors.max()) + 1)]


# ---------------------------------------------------------------- colouring
@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_mcs_matches_reference_fixture(oracle, name):
    """The reference's shipped vertex colourings are exact MCS outputs (mcs.cpp:14-168)."""
    pos, tets = load_t(os.path.join(GOLDEN, "coloring", name + ".t"))
    want = json.load(open(os.path.join(GOLDEN, "coloring", name + ".t.vertexColoring.json")))
    got = categories(oracle.color_tet_mesh(len(pos), tets, "mcs"))
    assert got == [sorted(c) for c in want]


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_greedy_is_valid_and_not_mcs_copy(oracle, name):
    pos, tets = load_t(os.path.join(GOLDEN, "coloring", name + ".t"))
    c = oracle.color_tet_mesh(len(pos), tets, "greedy")
    for t in tets:
        for a in range(4):
            for b in range(a + 1, 4):
                assert c[t[a]] != c[t[b]]


def test_greedy_brute_force_small_graphs(oracle):
    """Ordered greedy on tiny graphs equals a brute-force restatement of its rule."""
    rng = np.random.default_rng(7)
    for trial in range(40):
        n = int(rng.integers(2, 9))
        edges = [(a, b) for a in range(n) for b in range(a + 1, n) if rng.random() < 0.45]
        adj = {i: set() for i in range(n)}
        for a, b in edges:
            adj[a].add(b)
            adj[b].add(a)
        deg = {i: len(adj[i]) for i in range(n)}
        col = {}
        while len(col) < n:
            v = min((deg[i], i) for i in range(n) if i not in col)[1]
            used = {col[u] for u in adj[v] if u in col}
            col[v] = min(c for c in range(n + 1) if c not in used)
            for u in adj[v]:
                if u not in col:
                    deg[u] -= 1
        got = oracle.color_graph(n, np.asarray(edges, np.int32).reshape(-1, 2), "greedy") if edges else \
            oracle.color_graph(n, np.zeros((0, 2), np.int32), "greedy")
        assert got.tolist() == [col[i] for i in range(n)], trial


def test_mcs_path_graph_known_answer(oracle):
    """MCS on the path 0-1-2-3-4: order 0,1,2,3,4 (max weight, lowest id) -> colours 0,1,0,1,0."""
    edges = np.array([[0, 1], [1, 2], [2, 3], [3, 4]], np.int32)
    assert oracle.color_graph(5, edges, "mcs").tolist() == [0, 1, 0, 1, 0]


# ---------------------------------------------------------------- rest-pose constants
def test_tet_prepare_closed_form(oracle):
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    dminv, vol, mass = oracle.tet_prepare(pos, [[0, 1, 2, 3]], density=600.0)
    assert np.allclose(dminv[0].reshape(3, 3, order="F"), np.eye(3))
    assert vol[0] == pytest.approx(1.0 / 6.0)
    assert np.allclose(mass, 600.0 / 24.0)


def test_tet_prepare_random(oracle):
    rng = np.random.default_rng(1)
    pos = rng.normal(size=(4, 3))
    dminv, vol, mass = oracle.tet_prepare(pos, [[0, 1, 2, 3]], density=2.0)
    Dm = np.stack([pos[1] - pos[0], pos[2] - pos[0], pos[3] - pos[0]], axis=1)
    assert np.allclose(dminv[0].reshape(3, 3, order="F") @ Dm, np.eye(3), atol=1e-10)
    assert vol[0] == pytest.approx(np.linalg.det(Dm) / 6.0)
    assert mass.sum() == pytest.approx(2.0 * vol[0])


# ---------------------------------------------------------------- neo-Hookean
def nh_energy(x4, dminv, vol, mu, lmbd):
    Dm_inv = dminv.reshape(3, 3, order="F")
    Ds = np.stack([x4[1] - x4[0], x4[2] - x4[0], x4[3] - x4[0]], axis=1)
    F = Ds @ Dm_inv
    a = 1.0 + mu / lmbd
    return vol * (0.5 * mu * (np.sum(F * F) - 3.0) + 0.5 * lmbd * (np.linalg.det(F) - a) ** 2)


@pytest.mark.parametrize("order", [0, 1, 2, 3])
def test_neohookean_force_hessian_finite_difference(oracle, order):
    """Force = -dE/dx_v and Hessian = d2E/dx_v^2 of the stable neo-Hookean energy
    (VBD_NeoHookean.cpp:379-581), checked by central differences in FP64 (the
    reference's validators use eps 1e-6, thresholds 1e-5, VBD_NeoHookean.cpp:1536-1611)."""
    rng = np.random.default_rng(10 + order)
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.1
    dminv, vol, _ = oracle.tet_prepare(rest, [[0, 1, 2, 3]], 1.0)
    mu, lmbd = 1e5, 1e6
    x4 = rest + rng.normal(scale=0.02, size=(4, 3))
    f, H = oracle.neohookean_vertex(x4, dminv[0], vol[0], mu, lmbd, order)
    eps = 1e-7
    g = np.zeros(3)
    Hfd = np.zeros((3, 3))
    for d in range(3):
        xp, xm = x4.copy(), x4.copy()
        xp[order, d] += eps
        xm[order, d] -= eps
        g[d] = (nh_energy(xp, dminv[0], vol[0], mu, lmbd) - nh_energy(xm, dminv[0], vol[0], mu, lmbd)) / (2 * eps)
        fp, _ = oracle.neohookean_vertex(xp, dminv[0], vol[0], mu, lmbd, order)
        fm, _ = oracle.neohookean_vertex(xm, dminv[0], vol[0], mu, lmbd, order)
        Hfd[:, d] = -(fp - fm) / (2 * eps)
    assert np.allclose(-f, g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    # the reference drops nothing from the vertex block: exact second derivative
    assert np.allclose(H, Hfd, rtol=1e-4, atol=1e-5 * np.abs(Hfd).max())
    assert np.allclose(H, H.T, atol=1e-9 * np.abs(H).max())


def test_neohookean_rest_state_invariant(oracle):
    """F = I gives zero force (lambda (1 - a) = -mu cancels mu F m) and an SPD Hessian."""
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.1
    dminv, vol, _ = oracle.tet_prepare(rest, [[0, 1, 2, 3]], 1.0)
    for order in range(4):
        f, H = oracle.neohookean_vertex(rest, dminv[0], vol[0], 1e5, 1e6, order)
        assert np.abs(f).max() < 1e-8
        assert np.all(np.linalg.eigvalsh(H) > 0)


def test_neohookean_hessian_contraction_identity(oracle):
    """The vertex block equals A (mu |m|^2 I + lambda g g^T), g = cof(F) m: the
    det(F) second-derivative terms vanish once contracted with m (x) m. This is the
    identity the HIP kernel relies on."""
    rng = np.random.default_rng(3)
    rest = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.1
    dminv, vol, _ = oracle.tet_prepare(rest, [[0, 1, 2, 3]], 1.0)
    Dinv = dminv[0].reshape(3, 3, order="F")
    B = np.vstack([-Dinv.sum(axis=0), Dinv])
    mu, lmbd = 2e4, 3e5
    for order in range(4):
        x4 = rest + rng.normal(scale=0.03, size=(4, 3))
        _, H = oracle.neohookean_vertex(x4, dminv[0], vol[0], mu, lmbd, order)
        F = np.stack([x4[1] - x4[0], x4[2] - x4[0], x4[3] - x4[0]], axis=1) @ Dinv
        cof = np.linalg.det(F) * np.linalg.inv(F).T
        m = B[order]
        g = cof @ m
        want = vol[0] * (mu * (m @ m) * np.eye(3) + lmbd * np.outer(g, g))
        assert np.allclose(H, want, rtol=1e-8, atol=1e-8 * np.abs(want).max())


# ---------------------------------------------------------------- 3x3 solve
def test_solve3x3_against_lapack(oracle):
    rng = np.random.default_rng(5)
    for _ in range(200):
        A = rng.normal(size=(3, 3))
        H = A @ A.T + 0.1 * np.eye(3)
        b = rng.normal(size=3)
        ok, x = oracle.solve3x3(H, b)
        assert ok
        assert np.allclose(x, np.linalg.solve(H, b), rtol=1e-9, atol=1e-12)
        ok32, x32 = oracle.solve3x3(H.astype(np.float32), b.astype(np.float32), precision=32)
        assert ok32
        assert np.allclose(x32, np.linalg.solve(H, b), rtol=2e-3, atol=1e-5)


def test_solve3x3_nonsymmetric_orientation(oracle):
    """A non-symmetric system pins row/column orientation of the adjugate."""
    H = np.array([[4.0, 1.0, 0.5], [0.2, 3.0, 0.7], [0.3, 0.1, 5.0]])
    b = np.array([1.0, -2.0, 0.5])
    ok, x = oracle.solve3x3(H, b)
    assert ok and np.allclose(x, np.linalg.solve(H, b))


def test_solve3x3_singular_fails(oracle):
    H = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [0.0, 1.0, 1.0]])
    ok, _ = oracle.solve3x3(H, np.ones(3))
    assert not ok
