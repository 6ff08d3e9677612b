"""No-gravity zone of the velocity update on the device (hasNoGravZone / noGravZoneThreshold /
maxVelocityMagnitude, Materials/Materials.h:34-35; VBD/VBDPhysics.cpp:2529-2555 for tet meshes,
VBDCloth/VBDClothPhysics.cpp:2325-2351 for cloth): per-step parity with the FP32 oracle through a
crossing of the zone boundary, and the Models.json fields read by the loader."""
import json
import os

import numpy as np
import pytest

from gaia_amd import VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_gpu, build_oracle, build_oracle_cloth, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5
ZONE = dict(has_no_grav_zone=1, no_grav_zone_threshold=0.6, max_velocity_magnitude=0.8)


def tet_zone_spec():
    """A damped cube thrown downwards from 1 m: above y = 0.6 its speed is capped at 0.8 m/s (no
    damping); vertices below the threshold are damped; it lands on the ground at y = 0."""
    pos, tets = scenes.cube_tet_mesh(4, 0.05, origin=(0.0, 0.9, 0.0))
    mat = dict(density=1000.0, mu=1e5, lmbd=1e6, exponential_vel_damping=0.8, constant_vel_damping=0.05, **ZONE)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=8,
                           world_lower=(-1e7, 0.0, -1e7))
    vel = np.tile([0.3, -2.5, 0.0], (len(pos), 1))
    return dict(physics=physics, steps=40, meshes=[dict(pos=pos, tets=tets, material=mat, fixed=(), vel=vel,
                                                        coloring="greedy")])


def cloth_zone_spec():
    """A sheet tossed upwards: while its vertices rise faster than 0.6 m/s in y their speed is capped
    at 0.8 m/s; once they slow down the usual damping applies. 25 frames: the flight before the
    sheet lands flat on the ground (frame 27), where the FP32 and FP64 oracles part by ~0.1."""
    pos, faces = scenes.cloth_sheet(12, 0.5, 0.5)
    mu, lm = scenes.stvk_lame(1e4, 0.3)
    mat = dict(density=0.2, mu=mu, lmbd=lm, bending_stiffness=1e-4, initialization_type=0,
               exponential_vel_damping=0.99, constant_vel_damping=0.001, **ZONE)
    physics = physics_dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=8,
                           world_lower=(-1e7, 0.0, -1e7))
    vel = np.tile([0.1, 3.0, 0.0], (len(pos), 1))
    return dict(physics=physics, steps=25, meshes=[dict(pos=pos, faces=faces, material=mat, fixed=[], vel=vel,
                                                        coloring="mcs")])


@pytest.mark.parametrize("kind", ["tet", "cloth"])
def test_zone_per_step_parity(gpu_ctx, oracle, kind):
    spec = tet_zone_spec() if kind == "tet" else cloth_zone_spec()
    sim = build_gpu(gpu_ctx, spec)
    osc = (build_oracle if kind == "tet" else build_oracle_cloth)(oracle, spec)
    capped = 0
    for k in range(spec["steps"]):
        sim.setState(0, *osc.get_state(0))
        sim.runStep()
        osc.step(1, threads=4)
        x, v = sim.getState(0)
        xr, vr = osc.get_state(0)
        assert rel_err(x, xr) < TOL, f"step {k}"
        # v = (x - x_prev) / dt: a velocity difference is a position difference over dt, so it is
        # held to the position tolerance times max |x| / dt (the sheet's velocities are small
        # against that scale, so a relative bound on them alone would ask for sub-rounding positions)
        dt = spec["physics"]["time_step"] / spec["physics"]["num_substeps"]
        assert np.abs(v - vr).max() * dt < TOL * np.abs(xr).max(), f"step {k}"
        capped += int(np.sum(np.abs(np.linalg.norm(vr, axis=1) - 0.8) < 1e-6))
    print(f"\n[no-grav zone, {kind}] vertex-steps at the speed cap: {capped}", flush=True)
    assert capped > 0
    sim.destroy()


def test_zone_fields_loaded_from_models_json(gpu_ctx, tmp_path):
    spec = tet_zone_spec()
    m = spec["meshes"][0]
    physics = dict(gravity=(0.0, -9.8, 0.0), time_step=1.0 / 60.0, num_substeps=1, iterations=8,
                   world_lower=(-1e7, 0.0, -1e7), world_upper=(1e7, 1e7, 1e7))
    mat = dict(density=1000.0, mu=1e5, lmbd=1e6)
    models, params = scenes.write_scene_json(str(tmp_path), "cube", m["pos"], m["tets"], physics, mat)
    doc = json.load(open(models))
    doc["Models"][0].update(hasNoGravZone=True, noGravZoneThreshold=0.6, maxVelocityMagnitude=0.8,
                            initialVelocity=[0.0, -2.5, 0.0])
    json.dump(doc, open(models, "w"))
    a = VBDPhysics(gpu_ctx)
    a.loadRunningparameters(models, params, repo_root=str(tmp_path))
    a.initialize()
    b = VBDPhysics(gpu_ctx)
    b.setPhysicsParams(**physics)
    b.addTetMesh(m["pos"], m["tets"], velocities=np.tile([0.0, -2.5, 0.0], (len(m["pos"]), 1)),
                 coloring="mcs", **mat, **ZONE)
    b.initialize()
    for _ in range(6):
        a.runStep()
        b.runStep()
    (xa, va), (xb, vb) = a.getState(0), b.getState(0)
    assert rel_err(xa, xb) < 1e-6 and rel_err(va, vb) < 1e-6
    assert np.abs(np.linalg.norm(va, axis=1) - 0.8).min() < 1e-6  # the cap was applied
    a.destroy()
    b.destroy()
