"""Convergence evaluation on the device (gvbd_scene_evaluate_convergence, after
VBDPhysics::evaluateConvergenceGPU, VBD/VBDPhysics.cpp:3510-3556) against the oracle's restatement
on the same state and against closed forms."""
import numpy as np
import pytest

from gaia_amd import VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_gpu, build_oracle, c1_spec, multi_spec

pytestmark = pytest.mark.gpu
DT = float(np.float32(1.0 / 60.0))


def test_closed_forms_on_the_device(gpu_ctx):
    """At rest: zero energy and force (to float rounding). With no iterations the inertia energy
    is dt^2 |g|^2 M / 2 and the average force norm |g| M / N (the cube moves rigidly to
    x0 = x + dt v)."""
    pos, tets = scenes.cube_tet_mesh(2, 0.1, origin=(0.0, 0.5, 0.0))
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=0,
                                        world_lower=(-1e7, -1e7, -1e7)))
    sim.addTetMesh(pos, tets, density=1000.0, mu=2e4, lmbd=1e5, initialization_type=2,
                   velocities=np.tile([0.2, 0.0, -0.1], (len(pos), 1)))
    sim.initialize()
    c = sim.evaluateConvergence()
    # float rounding of 40 per-tet energies of about 0.3 J each, and of their forces
    assert abs(c["energy"]) < 1e-5 and c["avg_force_norm"] < 1e-4
    sim.runStep()
    c = sim.evaluateConvergence()
    mass = 1000.0 * 0.2 ** 3
    np.testing.assert_allclose(c["inertia_energy"], 0.5 * DT ** 2 * 9.8 ** 2 * mass, rtol=1e-4)
    np.testing.assert_allclose(c["avg_force_norm"], 9.8 * mass / len(pos), rtol=1e-4)
    sim.destroy()


@pytest.mark.parametrize("case", ["c1", "multi"])
def test_same_state_as_the_oracle(gpu_ctx, oracle, case):
    """Stepped in lockstep with the FP32 oracle (every step from the oracle's state), then
    evaluated on the device's final state:
    - as it is, near rest, where an FP32 evaluation is ill-conditioned (per-tet energies of order
      A lambda (1 - a)^2 cancel against the rest energy; forces of order 1e-3 N are float noise of
      stresses): the inertia energy equals the FP32 oracle's within 1e-4, the elastic energy and
      the average force norm are as close to the FP64 evaluation as the FP32 oracle is, with margin;
    - with every vertex moved by 3 mm of noise, where all three are well-conditioned: within 1e-3
      of the FP64 evaluation."""
    spec = c1_spec() if case == "c1" else multi_spec()
    spec["physics"] = dict(spec["physics"], iterations=4)
    sim = build_gpu(gpu_ctx, spec)
    o32 = build_oracle(oracle, spec)
    o64 = build_oracle(oracle, spec, 64)
    nm = len(spec["meshes"])
    for _ in range(6):
        for m in range(nm):
            st = o32.get_state(m)
            sim.setState(m, *st)
            o64.set_state(m, *st)
        sim.runStep()
        o32.step(1, threads=8)
        o64.step(1, threads=8)
    rng = np.random.default_rng(1)
    for moved in (False, True):
        for m in range(nm):
            x, v = sim.getState(m)
            if moved:
                x = (x + rng.normal(scale=0.003, size=x.shape)).astype(np.float32)
                sim.setState(m, x, v)
            o32.set_state(m, x, v)
            o64.set_state(m, x, v)
        g, a, b = sim.evaluateConvergence(), o32.evaluate_convergence(), o64.evaluate_convergence()
        print(f"\n[convergence {case}, moved {moved}] gpu {g}\n  fp32 {a}\n  fp64 {b}", flush=True)
        if moved:
            for k in g:
                np.testing.assert_allclose(g[k], b[k], rtol=1e-3, err_msg=k)
        else:
            np.testing.assert_allclose(g["inertia_energy"], a["inertia_energy"], rtol=1e-4)
            for k in ("elastic_energy", "avg_force_norm"):
                assert abs(g[k] - b[k]) <= max(3.0 * abs(a[k] - b[k]), 1e-4 * abs(b[k])), k
    sim.destroy()
