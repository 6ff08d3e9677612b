"""Pins of the oracle's cloth path (StVK membrane, quadratic bending, cloth step) and of the
product's host-side cloth preparation against it."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import scenes
from gaia_amd.vbd import physics_dict

TRI = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.02, 0.0, 0.09]])


def stvk_energy(x3, dminv4, area, mu, lmbd):
    D = np.asarray(dminv4).reshape(2, 2, order="F")
    Ds = np.stack([x3[1] - x3[0], x3[2] - x3[0]], axis=1)
    F = Ds @ D
    G = 0.5 * (F.T @ F - np.eye(2))
    return area * (mu * np.sum(G * G) + 0.5 * lmbd * np.trace(G) ** 2)


@pytest.mark.parametrize("order", [0, 1, 2])
def test_stvk_force_hessian_finite_difference(oracle, order):
    """f = -dW/dx_v and H = d2W/dx_v2 of W = A (mu |G|^2 + lambda/2 tr(G)^2)
    (VBDTriMeshStVK.cpp:721-911), central differences in FP64 (the reference's own validator
    uses delta 1e-4 and 1% on the global Hessian, Test.cpp:532-579)."""
    rng = np.random.default_rng(20 + order)
    dminv, area, _, _, _ = oracle.cloth_prepare(TRI, [[0, 1, 2]], 1.0)
    mu, lmbd = 3e4, 5e4
    x3 = TRI + rng.normal(scale=0.01, size=(3, 3))
    f, H = oracle.stvk_vertex(x3, dminv[0], area[0], mu, lmbd, order)
    eps = 1e-7
    g = np.zeros(3)
    Hfd = np.zeros((3, 3))
    for d in range(3):
        xp, xm = x3.copy(), x3.copy()
        xp[order, d] += eps
        xm[order, d] -= eps
        g[d] = (stvk_energy(xp, dminv[0], area[0], mu, lmbd) - stvk_energy(xm, dminv[0], area[0], mu, lmbd)) / (2 * eps)
        fp, _ = oracle.stvk_vertex(xp, dminv[0], area[0], mu, lmbd, order)
        fm, _ = oracle.stvk_vertex(xm, dminv[0], area[0], mu, lmbd, order)
        Hfd[:, d] = -(fp - fm) / (2 * eps)
    assert np.allclose(-f, g, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    assert np.allclose(H, Hfd, rtol=1e-4, atol=1e-5 * np.abs(Hfd).max())
    assert np.allclose(H, H.T, atol=1e-9 * np.abs(H).max())


def test_stvk_contraction_identity(oracle):
    """The oracle's 6x6 assembly contracts to A[(m^T S m) I + mu |m|^2 F F^T + (mu+lambda)(Fm)(Fm)^T],
    the closed form the HIP kernel evaluates."""
    rng = np.random.default_rng(2)
    dminv, area, _, _, _ = oracle.cloth_prepare(TRI, [[0, 1, 2]], 1.0)
    D = dminv[0].reshape(2, 2, order="F")
    B = np.vstack([-D.sum(axis=0), D])
    mu, lmbd = 1e4, 2e4
    for order in range(3):
        x3 = TRI + rng.normal(scale=0.02, size=(3, 3))
        _, H = oracle.stvk_vertex(x3, dminv[0], area[0], mu, lmbd, order)
        F = np.stack([x3[1] - x3[0], x3[2] - x3[0]], axis=1) @ D
        G = 0.5 * (F.T @ F - np.eye(2))
        S = 2 * mu * G + lmbd * np.trace(G) * np.eye(2)
        m = B[order]
        want = area[0] * ((m @ S @ m) * np.eye(3) + mu * (m @ m) * F @ F.T + (mu + lmbd) * np.outer(F @ m, F @ m))
        assert np.allclose(H, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())


def test_stvk_rest_state_zero_force(oracle):
    dminv, area, _, _, _ = oracle.cloth_prepare(TRI, [[0, 1, 2]], 1.0)
    for order in range(3):
        f, _ = oracle.stvk_vertex(TRI, dminv[0], area[0], 1e5, 1e5, order)
        assert np.abs(f).max() < 1e-9


def test_cloth_rest_pose_closed_form(oracle):
    """Right triangle legs a, b in a tilted plane: area a b / 2, masses rho A / 3, F = I at rest."""
    a, b = 0.3, 0.2
    R = np.linalg.qr(np.random.default_rng(0).normal(size=(3, 3)))[0]
    x = np.array([[0, 0, 0], [a, 0, 0], [0, b, 0]]) @ R.T + 0.5
    dminv, area, mass, edges, _ = oracle.cloth_prepare(x, [[0, 1, 2]], 2.0)
    assert area[0] == pytest.approx(a * b / 2)
    assert np.allclose(mass, 2.0 * a * b / 6)
    D = dminv[0].reshape(2, 2, order="F")
    F = np.stack([x[1] - x[0], x[2] - x[0]], axis=1) @ D
    assert np.allclose(F.T @ F, np.eye(2))
    assert len(edges) == 3 and np.all(edges[:, 5] == -1)


def test_bending_form_unit_square_known_answer(oracle):
    """Unit square split along the diagonal (v1=(0,0,0), v2=(1,1,0), v12Next=(1,0,0),
    v21Next=(0,1,0)): every cotangent is cot 45 = 1, so K0 = (2, 2, -2, -2), A0 = A1 = 1 and
    Q = 3 K0 K0^T / 2 (VBDTriMeshStVK.cpp:50-88)."""
    x = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 0], [0, 1, 0]], float)
    faces = [[0, 1, 2], [1, 0, 3]]
    _, _, _, edges, Q = oracle.cloth_prepare(x, faces, 1.0)
    e = [i for i in range(len(edges)) if edges[i, 5] != -1]
    assert len(e) == 1
    E = edges[e[0]]
    assert (E[0], E[1], E[2], E[3]) == (0, 1, 2, 3)
    K0 = np.array([2.0, 2.0, -2.0, -2.0])
    assert np.allclose(Q[e[0]], 1.5 * np.outer(K0, K0))


def test_bending_form_invariants(oracle):
    """Random flat sheet: Q X = 0 for every interior edge (flat => no bending force), rows sum
    to zero (translation invariance), Q symmetric PSD of rank one; a bent stencil has energy > 0."""
    rng = np.random.default_rng(1)
    pos, faces = scenes.cloth_sheet(5, 1.0, 0.0)
    pos[:, [0, 2]] += rng.normal(scale=0.02, size=(len(pos), 2))
    _, _, _, edges, Q = oracle.cloth_prepare(pos, faces, 1.0)
    for i in range(len(edges)):
        if edges[i, 5] == -1:
            assert np.all(Q[i] == 0)
            continue
        X = pos[edges[i, :4]]
        assert np.abs(Q[i] @ X).max() < 1e-9 * np.abs(Q[i]).max()
        assert np.abs(Q[i].sum(axis=1)).max() < 1e-9 * np.abs(Q[i]).max()
        assert np.allclose(Q[i], Q[i].T)
        assert np.linalg.matrix_rank(Q[i], tol=1e-9 * np.abs(Q[i]).max()) == 1
        Xb = X.copy()
        Xb[2, 1] += 0.05
        assert np.sum(Xb * (Q[i] @ Xb)) > 0


def test_product_cloth_preparation_matches_oracle(oracle):
    rng = np.random.default_rng(3)
    pos, faces = scenes.cloth_sheet(9, 0.8, 0.3)
    pos = pos + rng.normal(scale=0.01, size=pos.shape)
    d, a, m, e, Q = gaia_amd.cloth_mesh_constants(pos, faces, 0.2)
    od, oa, om, oe, oQ = oracle.cloth_prepare(pos.astype(np.float32).astype(np.float64), faces, 0.2)
    assert np.array_equal(e, oe)
    assert np.allclose(d, od, rtol=1e-5, atol=1e-5 * np.abs(od).max())
    assert np.allclose(a, oa, rtol=1e-5)
    assert np.allclose(m, om, rtol=1e-5)
    assert np.allclose(Q, oQ, rtol=1e-4, atol=1e-5 * np.abs(oQ).max())


@pytest.mark.parametrize("with_bending", [False, True])
def test_cloth_colouring_matches_oracle_and_is_proper(oracle, with_bending):
    pos, faces = scenes.cloth_sheet(12, 1.0, 0.0)
    c = gaia_amd.color_tri_mesh(len(pos), faces, "mcs", with_bending=with_bending)
    assert np.array_equal(c, oracle.color_tri_mesh(len(pos), faces, "mcs", with_bending=with_bending))
    for f in faces:
        assert len({c[f[0]], c[f[1]], c[f[2]]}) == 3
    if with_bending:
        _, _, _, edges, _ = gaia_amd.cloth_mesh_constants(pos, faces, 1.0)
        for E in edges:
            if E[5] != -1:
                assert c[E[2]] != c[E[3]]


def cloth_scene(oracle, physics, material, pos, faces, vel=None, fixed=(), precision=64):
    sc = oracle.ClothScene(physics, precision=precision)
    colors = oracle.color_tri_mesh(len(pos), faces)
    sc.add_mesh(pos, faces, material, colors, fixed=fixed, vel=vel)
    sc.finalize()
    return sc


def test_cloth_free_fall_closed_form(oracle):
    """Init type 0 (inertia) and a rigid translation: membrane and bending forces vanish, so one
    step lands exactly on x0 + dt v0 + dt^2 g."""
    dt = 0.01
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=dt, num_substeps=1, iterations=3, use_plane_ground=0)
    pos, faces = scenes.cloth_sheet(6, 0.5, 1.0)
    v0 = np.tile([0.2, 0.1, -0.3], (len(pos), 1))
    sc = cloth_scene(oracle, p, dict(density=0.2, mu=1e4, lmbd=2e4, bending_stiffness=1e-3, initialization_type=0),
                     pos, faces, vel=v0)
    sc.step(1)
    x, v = sc.get_state()
    dtf = float(np.float32(dt))
    assert np.allclose(x, pos + dtf * v0 + dtf * dtf * np.array([0, -9.8, 0]), atol=1e-12)
    assert np.allclose(v, v0 + dtf * np.array([0, -9.8, 0]), atol=1e-9)


def test_cloth_velocity_damping_not_rescaled(oracle):
    """Cloth applies exponentialVelDamping per substep as given (no pow(., dt); the tet path
    rescales), VBDClothPhysics.cpp:2318-2360."""
    dt = 0.01
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=dt, num_substeps=1, iterations=1, use_plane_ground=0)
    pos, faces = scenes.cloth_sheet(4, 0.5, 1.0)
    v0 = np.tile([1.0, 0.0, 0.5], (len(pos), 1))
    sc = cloth_scene(oracle, p, dict(density=0.2, initialization_type=0, exponential_vel_damping=0.9,
                                     constant_vel_damping=0.05), pos, faces, vel=v0)
    sc.step(1)
    _, v = sc.get_state()
    dtf = float(np.float32(dt))
    raw = v0[0] + dtf * np.array([0, -9.8, 0])
    n = np.linalg.norm(raw)
    assert np.allclose(v, raw * (n * 0.9 - 0.05) / n, rtol=1e-9)


def test_cloth_pinned_corners_stay(oracle):
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1 / 60, num_substeps=1, iterations=5, use_plane_ground=0)
    pos, faces = scenes.cloth_sheet(6, 0.5, 1.0)
    sc = cloth_scene(oracle, p, dict(density=0.2, mu=1e4, lmbd=2e4, bending_stiffness=1e-4), pos, faces,
                     fixed=[0, 5])
    sc.step(5)
    x, v = sc.get_state()
    assert np.array_equal(x[[0, 5]], pos[[0, 5]]) and np.all(v[[0, 5]] == 0)
    assert x[:, 1].min() < 1.0 - 1e-3
