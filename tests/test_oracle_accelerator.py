"""Pins of the oracle's Chebyshev acceleration of the tet iterations (runStepGPU,
VBD/VBDPhysics.cpp:872-906; VBDSolveParallelGroup_applyAcceleration_kernel,
VBDPhysicsCompute.cu:910-944; getAcceleratorOmega, VBDPhysics.cpp:3558-3575).

The omega recurrence is the Chebyshev semi-iterative one (omega_1 = 1, omega_2 = 2 / (2 - rho^2),
omega_k+1 = 4 / (4 - rho^2 omega_k); Wang, "A Chebyshev semi-iterative approach for accelerating
projective and position-based dynamics", 2015), whose limit is 2 / (1 + sqrt(1 - rho^2))."""
import numpy as np
import pytest

from gaia_amd import scenes
from gaia_amd.vbd import physics_dict

DT = float(np.float32(1.0 / 60.0))  # the scene's dt is float32


def omegas(rho, n):
    w, out = 1.0, []
    for k in range(1, n + 1):
        w = 1.0 if k == 1 else (2.0 / (2.0 - rho * rho) if k == 2 else 4.0 / (4.0 - rho * rho * w))
        out.append(w)
    return out


def test_omega_recurrence_tends_to_the_chebyshev_limit():
    for rho in (0.5, 0.93, 0.99):
        w = omegas(rho, 400)
        lim = 2.0 / (1.0 + np.sqrt(1.0 - rho * rho))
        assert abs(w[-1] - lim) < 1e-9
        assert all(a >= b >= lim - 1e-12 for a, b in zip(w[1:], w[2:]))  # decreasing from omega_2


def free_tet_scene(oracle, physics, v0):
    """One tet with negligible moduli: every vertex's Newton step lands on its inertia target."""
    pos = np.array([[0.0, 1.0, 0.0], [0.1, 1.0, 0.0], [0.0, 1.1, 0.0], [0.0, 1.0, 0.1]])
    sc = oracle.TetScene(physics, precision=64)
    sc.add_mesh(pos, np.int32([[0, 1, 2, 3]]), dict(density=1000.0, mu=1e-12, lmbd=1e-12, initialization_type=2),
                np.arange(4, dtype=np.int32), vel=np.tile(v0, (4, 1)))
    sc.finalize()
    return sc, pos


@pytest.mark.parametrize("iters", [1, 2, 5, 10])
def test_free_vertices_follow_the_accelerated_recurrence(oracle, iters):
    """Initialisation type 2 starts the iterations at x0 = x + dt v (gravity left out) while every
    sweep returns y = x0 + dt^2 g; the accelerated iterate is a_k = a_{k-2} + omega_k (y - a_{k-2})
    with a_{-1} = a_0 = x0. Without the accelerator the result is y."""
    rho, g, v0 = 0.9, np.array([0.0, -9.8, 0.0]), np.array([0.3, 0.1, -0.2])
    phys = physics_dict(gravity=tuple(g), time_step=DT, num_substeps=1, iterations=iters, use_accelerator=1,
                        accelerator_rho=rho, world_lower=(-1e7, -1e7, -1e7))
    sc, pos = free_tet_scene(oracle, phys, v0)
    sc.step(1, threads=1)
    x, v = sc.get_state(0)
    x0 = pos + DT * v0
    y = x0 + DT * DT * g
    a = [x0, x0]
    for w in omegas(rho, iters):
        a.append(a[-2] + w * (y - a[-2]))
    np.testing.assert_allclose(x, a[-1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(v, (a[-1] - pos) / DT, rtol=0, atol=1e-9)
    plain, _ = free_tet_scene(oracle, dict(phys, use_accelerator=0), v0)
    plain.step(1, threads=1)
    np.testing.assert_allclose(plain.get_state(0)[0], y, rtol=0, atol=1e-12)


def soft_cube(oracle, **kw):
    """A soft cube with random vertex velocities (the inertia targets strain it)."""
    import gaia_amd

    pos, tets = scenes.cube_tet_mesh(3, 0.1, origin=(0.0, 0.5, 0.0))
    vel = np.random.default_rng(2).normal(scale=0.5, size=pos.shape)
    phys = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 120.0, num_substeps=1,
                        world_lower=(-1e7, -1e7, -1e7), **kw)
    sc = oracle.TetScene(phys, precision=64)
    sc.add_mesh(pos, tets, dict(density=1000.0, mu=2e4, lmbd=1e5), gaia_amd.color_tet_mesh(len(pos), tets, "greedy"),
                vel=vel)
    sc.finalize()
    return sc


def test_rho_zero_is_the_plain_iteration(oracle):
    """rho = 0 makes every omega 1: x <- y_{k-1} + (x - y_{k-1}), the plain sweep up to rounding."""
    a = soft_cube(oracle, iterations=6, use_accelerator=1, accelerator_rho=0.0)
    b = soft_cube(oracle, iterations=6, use_accelerator=0)
    a.step(3, threads=1)
    b.step(3, threads=1)
    np.testing.assert_allclose(a.get_state(0)[0], b.get_state(0)[0], rtol=0, atol=1e-13)


def test_acceleration_keeps_the_fixed_point(oracle):
    """With enough iterations both reach the same implicit-Euler step (the accelerator changes the
    path, not the solution, up to the 1e-10 update threshold)."""
    ref = soft_cube(oracle, iterations=400, use_accelerator=0)
    acc = soft_cube(oracle, iterations=400, use_accelerator=1, accelerator_rho=0.3)
    ref.step(1, threads=8)
    acc.step(1, threads=8)
    np.testing.assert_allclose(acc.get_state(0)[0], ref.get_state(0)[0], rtol=0, atol=1e-8)


def test_acceleration_converges_faster_on_a_stiff_cube(oracle):
    """A stiffer cube at dt = 1/60, where the plain sweep converges slowly: after 10 iterations the
    accelerated iterate (rho = 0.8) is closer to the converged step than the plain one."""
    import gaia_amd

    def cube(**kw):
        pos, tets = scenes.cube_tet_mesh(6, 0.05, origin=(0.0, 0.5, 0.0))
        vel = np.random.default_rng(2).normal(scale=0.5, size=pos.shape)
        phys = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1,
                            world_lower=(-1e7, -1e7, -1e7), **kw)
        sc = oracle.TetScene(phys, precision=64)
        sc.add_mesh(pos, tets, dict(density=1000.0, mu=2e5, lmbd=1e6),
                    gaia_amd.color_tet_mesh(len(pos), tets, "greedy"), vel=vel)
        sc.finalize()
        sc.step(1, threads=8)
        return sc.get_state(0)[0]

    x = cube(iterations=2000, use_accelerator=0)
    e_plain = np.abs(cube(iterations=10, use_accelerator=0) - x).max()
    e_acc = np.abs(cube(iterations=10, use_accelerator=1, accelerator_rho=0.8) - x).max()
    assert e_acc < 0.6 * e_plain
