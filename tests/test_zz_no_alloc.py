"""The allocation-free convergence evaluation (gvbd_scene_evaluate_convergence_noalloc, no
reference counterpart) through the C ABI: the same doubles as gvbd_scene_evaluate_convergence
(VBDPhysics::evaluateConvergenceGPU, VBD/VBDPhysics.cpp:3510-3556) on two calls in a row, the
second on a smaller scene, and again after both scenes have stepped; the precondition error when
the scene's energy records are not resident yet."""
import numpy as np
import pytest

from gaia_amd import VBDPhysics, scenes
from gaia_amd._lib import VBDError
from gaia_amd.vbd import physics_dict

pytestmark = pytest.mark.gpu
KEYS = ("energy", "inertia_energy", "elastic_energy", "avg_force_norm")


def _deformed_cube(ctx, cells, seed):
    """A cube with one fixed corner stepped once, then every vertex moved by 2 mm of noise (so
    all three parts are far from zero)."""
    pos, tets = scenes.cube_tet_mesh(cells, 0.1, origin=(0.0, 0.5, 0.0))
    sim = VBDPhysics(ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1 / 60, num_substeps=1, iterations=3,
                                        world_lower=(-1e7, 0.0, -1e7)))
    sim.addTetMesh(pos, tets, fixed_points=[0], density=1000.0, mu=1e5, lmbd=1e6, damping_hydrostatic=1e-3,
                   damping_deviatoric=1e-3, initialization_type=2)
    sim.initialize()
    sim.runStep()
    x, v = sim.getState(0)
    rng = np.random.default_rng(seed)
    sim.setState(0, x + rng.normal(0.0, 2e-3, x.shape).astype(np.float32), v)
    return sim


def _same(a, b):
    for k in KEYS:
        assert a[k] == b[k], (k, a[k], b[k])
        assert np.isfinite(a[k])


def test_noalloc_equals_the_allocating_entry_point(gpu_ctx):
    big = _deformed_cube(gpu_ctx, 9, 0)     # 1,000 vertices, 3,645 tets
    small = _deformed_cube(gpu_ctx, 3, 1)   # 64 vertices, 135 tets
    print(f"\n[no-alloc convergence] scenes of {big.meshInfo(0)[0]} and {small.meshInfo(0)[0]} vertices", flush=True)
    with pytest.raises(VBDError) as err:
        small.evaluateConvergenceNoAlloc()
    assert "set-up" in str(err.value)

    ref_big = big.evaluateConvergence()      # set-up: allocates the big scene's energy records
    ref_small = small.evaluateConvergence()
    # two calls in a row, the second with the smaller scene
    _same(ref_big, big.evaluateConvergenceNoAlloc())
    _same(ref_small, small.evaluateConvergenceNoAlloc())
    assert ref_big["elastic_energy"] > 0.0 and ref_small["elastic_energy"] > 0.0

    # after a step of each the scratch holds the last call's partials; the results still match
    big.runStep()
    small.runStep()
    got_big = big.evaluateConvergenceNoAlloc()
    got_small = small.evaluateConvergenceNoAlloc()
    _same(big.evaluateConvergence(), got_big)
    _same(small.evaluateConvergence(), got_small)
    big.destroy()
    small.destroy()
