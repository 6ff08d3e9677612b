"""Closed-form pins of the oracle's tet VBD step (VBDPhysics.cpp:1580-1632 restated).

Single-tet scenes where one step has a closed form: free fall, ground penalty,
IPC friction in both mollifier branches, velocity damping and material damping.
"""
import numpy as np
import pytest

from gaia_amd.vbd import physics_dict

REST = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.1
TET = np.array([[0, 1, 2, 3]], np.int32)


def single_tet(oracle, physics, material, x0, v0, precision=64, fixed=()):
    sc = oracle.TetScene(physics, precision=precision)
    colors = oracle.color_tet_mesh(4, TET, "mcs")
    sc.add_mesh(x0, TET, material, colors, fixed=fixed, vel=v0)
    sc.finalize()
    return sc


def test_free_fall_one_step(oracle):
    dt = 0.01
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=dt, num_substeps=1, iterations=3, use_plane_ground=0)
    v0 = np.tile([0.3, 0.1, -0.2], (4, 1))
    sc = single_tet(oracle, p, dict(density=1000.0, mu=1e5, lmbd=1e6), REST, v0)
    sc.step(1)
    x, v = sc.get_state()
    dtf = float(np.float32(dt))
    g = np.array([0, -9.8, 0])
    assert np.allclose(x, REST + dtf * v0 + dtf * dtf * g, rtol=0, atol=1e-12)
    assert np.allclose(v, v0 + dtf * g, atol=1e-9)


def test_fixed_point_stays(oracle):
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=0.01, num_substeps=1, iterations=5, use_plane_ground=0)
    sc = single_tet(oracle, p, dict(density=1000.0, mu=1e5, lmbd=1e6), REST, np.zeros((4, 3)), fixed=[2])
    sc.step(3)
    x, v = sc.get_state()
    assert np.array_equal(x[2], REST[2]) and np.all(v[2] == 0)
    assert x[0, 1] < REST[0, 1]


def test_ground_penalty_one_iteration(oracle):
    """Vertex 0 starts 1 cm below y = 0 with everything at rest and no gravity: one
    Newton step on its block gives dx_y = k pen / (m/dt^2 + k) (material ~ 0)."""
    dt, k = 0.01, 1e5
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=dt, num_substeps=1, iterations=1,
                     world_lower=(-1e7, 0.0, -1e7), boundary_collision_stiffness=k)
    x0 = REST + np.array([0.0, 0.05, 0.0])
    x0[0, 1] = -0.01
    mat = dict(density=1000.0, mu=1e-9, lmbd=1e-9)
    sc = single_tet(oracle, p, mat, x0, np.zeros((4, 3)))
    _, _, mass = oracle.tet_prepare(x0, TET, 1000.0)
    sc.step(1)
    x, _ = sc.get_state()
    dtf = float(np.float32(dt))
    m = mass[0]
    want = 0.01 * k / (m / dtf ** 2 + k)
    assert x[0, 1] - x0[0, 1] == pytest.approx(want, rel=1e-7)
    assert np.allclose(x[0, [0, 2]], x0[0, [0, 2]], atol=1e-12)


@pytest.mark.parametrize("unorm", [0.02, 0.004])
def test_ground_friction_both_branches(oracle, unorm):
    """Tangential slide u with penetration: friction -mu lambda f1(|u|) u with
    f1 = 1/|u| above eps_u = eps_v dt, (2 - |u|/eps_u)/eps_u below
    (VBDPhysics.cpp:2563-2592); Hessian mu lambda f1 on the tangent axes."""
    dt, k, muf, epsv = 0.01, 1e5, 0.3, 1.0
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=dt, num_substeps=1, iterations=1,
                     world_lower=(-1e7, 0.0, -1e7), boundary_collision_stiffness=k,
                     boundary_friction_dynamic=muf, boundary_friction_eps_v=epsv)
    dtf = float(np.float32(dt))
    u = np.array([0.6, 0.0, 0.8]) * unorm
    x0 = REST + np.array([0.0, 0.05, 0.0])
    x0[0, 1] = -0.01
    v0 = np.tile(u / dtf, (4, 1))
    mat = dict(density=1000.0, mu=1e-9, lmbd=1e-9)
    sc = single_tet(oracle, p, mat, x0, v0)
    _, _, mass = oracle.tet_prepare(x0, TET, 1000.0)
    sc.step(1)
    x, _ = sc.get_state()
    m = mass[0]
    pen = 0.01
    lam = k * pen
    epsu = np.float32(epsv) * np.float32(dtf)
    un = np.linalg.norm(u)
    f1 = 1.0 / un if un > epsu else (-un / epsu + 2.0) / epsu
    start = x0[0] + dtf * v0[0]
    dxt = -muf * lam * f1 * u / (m / dtf ** 2 + muf * lam * f1)
    want = start + dxt
    want[1] = start[1] + pen * k / (m / dtf ** 2 + k)
    assert np.allclose(x[0], want, rtol=1e-7, atol=1e-10)


def test_velocity_damping(oracle):
    """dampVelocities with exponentialVelDamping e and constantVelDamping c made
    step-invariant (e^dt, c dt; VBDPhysics.cpp:794-807, 2520-2561)."""
    dt = 0.01
    p = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=dt, num_substeps=1, iterations=1, use_plane_ground=0)
    v0 = np.tile([1.0, 0.5, -0.25], (4, 1))
    mat = dict(density=1000.0, mu=1e5, lmbd=1e6, exponential_vel_damping=0.5, constant_vel_damping=0.1)
    sc = single_tet(oracle, p, mat, REST, v0)
    sc.step(1)
    _, v = sc.get_state()
    dtf = float(np.float32(dt))
    raw = v0[0] + dtf * np.array([0, -9.8, 0])
    n = np.linalg.norm(raw)
    want = raw * (n * 0.5 ** dtf - 0.1 * dtf) / n
    assert np.allclose(v, want, rtol=1e-9)


def test_material_damping_first_vertex(oracle):
    """Rigid slide by u with hydrostatic/deviatoric damping: for the first vertex
    solved, f = -D u / dt and H = m/dt^2 + H_el + D/dt with
    D = gamma_h A lambda m m^T + gamma_d A mu |m|^2 I and H_el = A (mu |m|^2 I + lambda m m^T)
    at F = I (VBD_NeoHookean.cpp:562-574)."""
    dt, mu, lmbd, gh, gd = 0.01, 1e5, 1e6, 0.02, 0.05
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=dt, num_substeps=1, iterations=1, use_plane_ground=0)
    dtf = float(np.float32(dt))
    u = np.array([0.001, -0.002, 0.0015])
    v0 = np.tile(u / dtf, (4, 1))
    mat = dict(density=1000.0, mu=mu, lmbd=lmbd, damping_hydrostatic=gh, damping_deviatoric=gd)
    sc = single_tet(oracle, p, mat, REST, v0)
    groups = sc.groups()
    first = int(groups[0][0, 1])
    dminv, vol, mass = oracle.tet_prepare(REST, TET, 1000.0)
    Dinv = dminv[0].reshape(3, 3, order="F")
    B = np.vstack([-Dinv.sum(axis=0), Dinv])
    mv = B[first]
    A = vol[0]
    D = gh * A * lmbd * np.outer(mv, mv) + gd * A * mu * (mv @ mv) * np.eye(3)
    Hel = A * (mu * (mv @ mv) * np.eye(3) + lmbd * np.outer(mv, mv))
    H = mass[first] / dtf ** 2 * np.eye(3) + Hel + D / dtf
    f = -D @ u / dtf
    sc.step(1)
    x, _ = sc.get_state()
    start = REST[first] + u
    assert np.allclose(x[first], start + np.linalg.solve(H, f), rtol=1e-8, atol=1e-12)
