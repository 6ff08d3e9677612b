"""Pins of the oracle's cloth contact model (VBDClothPhysics.cpp:1223-1399, 1500-1573) by
closed-form Newton steps of single free vertices: VF repulsion (quadratic and log branches),
truncation to the conservative bound, IPC friction, and an EE-only configuration."""
import numpy as np
import pytest

from gaia_amd.vbd import physics_dict

DT = float(np.float32(1.0 / 60.0))  # the scene's dt is the FP32 time step, as on the device


def scene(oracle, meshes, iterations=1, **phys):
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=iterations,
                     world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=100.0,
                     contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45, contact_energy_type=0,
                     contact_query_radius=0.02)
    p.update(phys)
    sc = oracle.ClothScene(p, precision=64)
    for m in meshes:
        colors = np.arange(len(m["pos"]), dtype=np.int32)  # every vertex its own colour
        sc.add_mesh(np.asarray(m["pos"], np.float64), m["faces"], m["material"], colors, fixed=m.get("fixed", []),
                    vel=m.get("vel"))
    sc.finalize()
    return sc


BIG = dict(pos=[[-10.0, 0.0, -10.0], [10.0, 0.0, -10.0], [0.0, 0.0, 10.0]], faces=np.array([[0, 2, 1]], np.int32),
           material=dict(density=1.0, mu=0.0, lmbd=0.0), fixed=[0, 1, 2])


def small(d, density=16666.666666666668, vel=None, **mat):
    m = dict(density=density, mu=0.0, lmbd=0.0, bending_stiffness=0.0, initialization_type=2)
    m.update(mat)
    pos = [[0.0, d, 0.0], [0.01, d, 0.0], [0.0, d, 0.01]]
    return dict(pos=pos, faces=np.array([[0, 2, 1]], np.int32), material=m, vel=vel)


def vertex_mass(density):
    return density * 0.5 * 0.01 * 0.01 / 3.0


def test_vf_quadratic_newton_step_is_exact(oracle):
    """Linear contact force (energy type 0) + inertia: one Newton step lands on the fixed point
    height (k r + M d) / (M + k) for every vertex of the small triangle."""
    d, k, r = 0.004, 100.0, 0.01
    sc = scene(oracle, [BIG, small(d)], iterations=3)
    sc.step(1, threads=1)
    x, v = sc.get_state(1)
    M = vertex_mass(16666.666666666668) / DT ** 2
    want = (k * r + M * d) / (M + k)
    np.testing.assert_allclose(x[:, 1], want, rtol=1e-9)
    np.testing.assert_allclose(v[:, 1], (want - d) / DT, rtol=1e-7)
    assert (want - d) < 0.45 * d  # no truncation in this configuration


def test_vf_log_branch_one_newton_step(oracle):
    """Energy type 1 below tau = r / 2: E' = -k2 / d, E'' = k2 / d^2 with k2 = tau^2 k / 2."""
    d, k, r = 0.003, 100.0, 0.01
    sc = scene(oracle, [BIG, small(d)], iterations=1, contact_energy_type=1)
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    M = vertex_mass(16666.666666666668) / DT ** 2
    k2 = 0.5 * (0.5 * r) ** 2 * k
    want = d + (k2 / d) / (M + k2 / d ** 2)
    np.testing.assert_allclose(x[:, 1], want, rtol=1e-9)


def test_truncation_to_the_conservative_bound(oracle):
    """A stiff contact would push the vertex by more than its bound 0.45 d: the step is cut to
    exactly relax * (d - thickness)."""
    d = 0.004
    f32 = np.float32
    sc = scene(oracle, [BIG, small(d)], iterations=1, contact_stiffness=1e5)
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    # the detector works in FP32 (as Embree does): bound = 0.45f * d_f
    np.testing.assert_allclose(x[:, 1], d + float(f32(0.45) * f32(d)), rtol=1e-9)
    # with a thickness the bound shrinks to relax * (d - thickness)
    sc = scene(oracle, [BIG, small(d)], iterations=1, contact_stiffness=1e5, thickness=0.001,
               contact_radius=0.009)
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    np.testing.assert_allclose(x[:, 1], d + float(f32(0.45) * (f32(d) - f32(0.001))), rtol=1e-9)


def test_vf_friction_newton_step(oracle):
    """Sliding along +x at vx with |u| = dt vx > epsU: friction -mu lambda e0 with tangential
    Hessian mu lambda / |u|; the normal direction is as without friction. (The whole step stays
    inside the conservative bound 0.45 d.)"""
    d, k, r, vx = 0.004, 100.0, 0.01, 0.05
    mu_f = 0.3
    big = dict(BIG, material=dict(BIG["material"], friction_dynamic=mu_f, friction_eps_v=0.01))
    sm = small(d, vel=np.tile([vx, 0.0, 0.0], (3, 1)), initialization_type=0, friction_dynamic=mu_f,
               friction_eps_v=0.01)
    sc = scene(oracle, [big, sm], iterations=1)
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    M = vertex_mass(16666.666666666668) / DT ** 2
    lam = k * (r - d)
    u = DT * vx
    x0 = np.array(sm["pos"])[:, 0] + DT * vx
    want_x = x0 - mu_f * lam / (M + mu_f * lam / u)
    np.testing.assert_allclose(x[:, 0], want_x, rtol=1e-9)
    np.testing.assert_allclose(x[:, 1], (k * r + M * d) / (M + k), rtol=1e-9)


def test_ee_only_crossing_edges(oracle):
    """Two triangles whose edges cross at distance d (only EE contacts): the free endpoint of the
    lower edge (weight mu = 1/2) moves down by 0.5 k (r - d) / (M + 0.25 k)."""
    d, k, r = 0.004, 100.0, 0.01
    density = 3000.0
    A = dict(pos=[[-1.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, -1.0, 0.0]], faces=np.array([[0, 1, 2]], np.int32),
             material=dict(density=density, mu=0.0, lmbd=0.0), fixed=[0, 2])
    B = dict(pos=[[0.0, d, -1.0], [0.0, d, 1.0], [0.0, 1.0 + d, 0.0]], faces=np.array([[0, 2, 1]], np.int32),
             material=dict(density=1.0, mu=0.0, lmbd=0.0), fixed=[0, 1, 2])
    sc = scene(oracle, [A, B], iterations=1)
    sc.step(1, threads=1)
    x, _ = sc.get_state(0)
    area = 1.0
    M = density * area / 3.0 / DT ** 2
    np.testing.assert_allclose(x[1], [1.0, -0.5 * k * (r - d) / (M + 0.25 * k), 0.0], atol=1e-12)
    np.testing.assert_allclose(x[[0, 2]], np.array(A["pos"])[[0, 2]], atol=0)


def test_no_contact_beyond_radius(oracle):
    """Vertices farther than the contact radius (but within the query radius) feel nothing."""
    d = 0.015  # contact radius 0.01 < d < query radius 0.02
    sc = scene(oracle, [BIG, small(d)], iterations=2)
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    np.testing.assert_allclose(x[:, 1], d, rtol=0, atol=0)
