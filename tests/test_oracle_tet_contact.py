"""Pins of the oracle's tet meshes with cloth-model contact on their boundary surface (the
composition used for scenes of tet bodies in contact; the reference's tet solver has its own
penetration contact, VBDPhysics.cpp:2205-2311, with no contact radius):

- the boundary surface is closed, consistently and outward oriented;
- without contact forces and truncation the tet vertex solve is the tet oracle's, bit for bit;
- a vertex of a soft tet over a fixed triangle takes the closed-form VF Newton step.
"""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import scenes
from gaia_amd.vbd import physics_dict

DT = float(np.float32(1.0 / 60.0))


def contact_physics(**kw):
    p = physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=1,
                     world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=100.0,
                     contact_radius=0.01, thickness=0.0, conservative_step_relaxation=0.45, contact_energy_type=0,
                     contact_query_radius=0.02)
    p.update(kw)
    return p


def test_boundary_surface_closed_and_outward(oracle):
    n, h = 3, 0.1
    pos, tets = scenes.cube_tet_mesh(n, h)
    sc = oracle.ClothScene(contact_physics(), precision=64)
    sc.add_tet_mesh(pos, tets, dict(density=1000.0, mu=1e5, lmbd=1e6), np.arange(len(pos), dtype=np.int32))
    faces = sc.surface(0)
    assert len(faces) == 12 * n * n  # two triangles per boundary square
    # every directed edge appears once, its reverse once: closed and consistently oriented
    directed = {}
    for f in faces:
        for k in range(3):
            e = (int(f[k]), int(f[(k + 1) % 3]))
            directed[e] = directed.get(e, 0) + 1
    assert all(c == 1 for c in directed.values())
    assert all((b, a) in directed for (a, b) in directed)
    # outward: normals point away from the cube's centre
    centre = pos.mean(axis=0)
    a, b, c = pos[faces[:, 0]], pos[faces[:, 1]], pos[faces[:, 2]]
    normal = np.cross(b - a, c - a)
    assert np.all(np.einsum("ij,ij->i", normal, (a + b + c) / 3 - centre) > 0)
    # only boundary vertices are on the surface: (n+1)^3 - (n-1)^3 of them
    assert len(np.unique(faces)) == (n + 1) ** 3 - (n - 1) ** 3


def test_without_contact_forces_equals_tet_oracle(oracle):
    """One vertex per colour (Jacobi = Gauss-Seidel), contact radius far below every distance and
    a query radius large enough that no step reaches the conservative bound: the composition steps
    exactly as the tet oracle."""
    pos, tets = scenes.cube_tet_mesh(2, 0.05, origin=(0.0, 0.1, 0.0))
    rng = np.random.default_rng(3)
    pos = pos + rng.normal(scale=0.003, size=pos.shape)  # a strained start: elastic forces act
    colors = np.arange(len(pos), dtype=np.int32)
    mat = dict(density=1000.0, mu=1e5, lmbd=1e6, damping_hydrostatic=0.01, damping_deviatoric=0.02)
    for precision in (64, 32):
        phys = contact_physics(gravity=(0.0, -9.8, 0.0), iterations=4, contact_radius=1e-7, contact_query_radius=1.0,
                               world_lower=(-1e7, 0.0, -1e7))
        a = oracle.ClothScene(phys, precision=precision)
        a.add_tet_mesh(pos, tets, mat, colors)
        a.finalize()
        b = oracle.TetScene(dict(phys, handle_collision=0), precision=precision)
        b.add_mesh(pos, tets, mat, colors)
        b.finalize()
        for _ in range(3):
            a.step(1, threads=1)
            b.step(1, threads=1)
        xa, va = a.get_state(0)
        xb, vb = b.get_state(0)
        assert np.array_equal(xa, xb) and np.array_equal(va, vb), precision


def test_soft_tet_vertex_over_fixed_triangle_closed_form(oracle):
    """A tet with its apex at distance d above a large fixed triangle (only VF contacts; elastic
    moduli negligible, zero force at rest): one Newton step puts the apex at (k r + M d) / (M + k),
    the other vertices stay."""
    d, k, r = 0.004, 100.0, 0.01
    big = np.array([[-10.0, 0.0, -10.0], [10.0, 0.0, -10.0], [0.0, 0.0, 10.0]])
    L = 0.05
    tet_pos = np.array([[0.0, d, 0.0], [L, d + L, 0.0], [-L, d + L, -L], [0.0, d + L, L]])
    tets = np.array([[0, 1, 2, 3]], np.int32)
    vol = np.linalg.det(np.stack([tet_pos[1] - tet_pos[0], tet_pos[2] - tet_pos[0], tet_pos[3] - tet_pos[0]])) / 6
    assert vol > 0  # positively oriented (rest volume det(Dm) / 6 > 0)
    density = 1000.0 * DT ** 2 * 4 / vol  # lumped mass m with m / dt^2 = 1000
    sc = oracle.ClothScene(contact_physics(), precision=64)
    sc.add_mesh(big, np.array([[0, 2, 1]], np.int32), dict(density=1.0, mu=0.0, lmbd=0.0),
                np.arange(3, dtype=np.int32), fixed=[0, 1, 2])
    sc.add_tet_mesh(tet_pos, tets, dict(density=density, mu=1e-9, lmbd=1e-9), np.arange(4, dtype=np.int32))
    sc.finalize()
    sc.step(1, threads=1)
    x, _ = sc.get_state(1)
    M = density * vol / 4 / DT ** 2
    np.testing.assert_allclose(M, 1000.0, rtol=1e-12)
    np.testing.assert_allclose(x[0], [0.0, (k * r + M * d) / (M + k), 0.0], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(x[1:], tet_pos[1:], rtol=0, atol=1e-12)


def _depth_inside(points, pos, tets):
    """Per point, the largest smallest barycentric coordinate over the tets: > 0 inside a tet."""
    a = pos[tets]
    dm = np.stack([a[:, 1] - a[:, 0], a[:, 2] - a[:, 0], a[:, 3] - a[:, 0]], axis=2)
    lam = np.einsum("tij,tpj->tpi", np.linalg.inv(dm), points[None, :, :] - a[:, 0][:, None, :])
    return np.minimum(lam.min(-1), 1.0 - lam.sum(-1)).max(0)


@pytest.mark.parametrize("energy", [0, 1, 2])
def test_two_cubes_never_interpenetrate(oracle, energy):
    """Invariant of the conservative bound: a cube dropped on a cube on a tilted fixed sheet, 40
    steps in FP64. No vertex of either cube ever lies inside a tet of the other, and no cube
    vertex crosses the sheet's plane."""
    from vbd_cases import build_oracle_mixed, tet_drop_spec

    spec = tet_drop_spec(second_cube=True, steps=40, energy=energy)
    tets = spec["meshes"][1]["tets"]
    sc = build_oracle_mixed(oracle, spec, 64)
    for _ in range(spec["steps"]):
        sc.step(1, threads=8)
        c1, c2 = sc.get_state(1)[0], sc.get_state(2)[0]
        assert _depth_inside(c2, c1, tets).max() < 0 and _depth_inside(c1, c2, tets).max() < 0
        for c in (c1, c2):
            assert ((c[:, 1] - 0.02 * c[:, 0]) / np.hypot(1.0, 0.02)).min() > 0
    # the lower cube ends on the sheet, inside the contact radius
    c1 = sc.get_state(1)[0]
    assert ((c1[:, 1] - 0.02 * c1[:, 0]) / np.hypot(1.0, 0.02)).min() < 0.01
