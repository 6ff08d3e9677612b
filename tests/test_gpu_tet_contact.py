"""GPU parity of tet bodies in contact: the tet vertex solve with cloth-model VF/EE contact,
friction and conservative-bound truncation on each tet mesh's boundary surface (the composition
for scenes of tet bodies; the reference's tet solver uses penetration contact instead). The HIP
path through the C ABI against the FP32 oracle, per step from identical state, 1e-5
normwise-relative."""
import numpy as np
import pytest

import gaia_amd
from gaia_amd import VBDPhysics, scenes
from gaia_amd.vbd import physics_dict
from vbd_cases import build_gpu, build_oracle_mixed, rel_err, tet_drop_spec

pytestmark = pytest.mark.gpu
TOL = 1e-5
THREADS = 8
DT = float(np.float32(1.0 / 60.0))


def parity(gpu_ctx, oracle, spec, with_fp64=False):
    """Per step from identical state: the HIP path against the FP32 oracle (and, when asked, the
    FP32 oracle against the FP64 one)."""
    sim = build_gpu(gpu_ctx, spec)
    o32 = build_oracle_mixed(oracle, spec)
    o64 = build_oracle_mixed(oracle, spec, 64) if with_fp64 else None
    nm = len(spec["meshes"])
    errs, errs64 = [], []
    for _ in range(spec["steps"]):
        for m in range(nm):
            sim.setState(m, *o32.get_state(m))
            if o64 is not None:
                o64.set_state(m, *o32.get_state(m))
        sim.runStep()
        o32.step(1, threads=THREADS)
        if o64 is not None:
            o64.step(1, threads=THREADS)
            errs64.append(max(rel_err(o32.get_state(m)[0], o64.get_state(m)[0]) for m in range(nm)))
        errs.append(max(rel_err(sim.getState(m)[0], o32.get_state(m)[0]) for m in range(nm)))
    sim.destroy()
    return np.array(errs), np.array(errs64)


@pytest.mark.parametrize("energy", [0, 1, 2])
def test_cube_on_fixed_sheet_per_step_parity(gpu_ctx, oracle, energy):
    errs, _ = parity(gpu_ctx, oracle, tet_drop_spec(energy=energy))
    print(f"\n[tet contact, energy {energy}] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: {errs.max():.3e}"


@pytest.mark.parametrize("energy", [0, 1, 2])
def test_two_cubes_per_step_parity(gpu_ctx, oracle, energy):
    """A cube dropped on a cube resting on the sheet: tet-tet contact between the two boundary
    surfaces, with truncation on the impact steps."""
    errs, _ = parity(gpu_ctx, oracle, tet_drop_spec(second_cube=True, steps=25, energy=energy))
    print(f"\n[tet contact, two cubes, energy {energy}] worst per-step rel err {errs.max():.3e}", flush=True)
    assert errs.max() < TOL, f"step {errs.argmax()}: {errs.max():.3e}"


def test_many_bodies_error_distribution(gpu_ctx, oracle):
    """The tet-contact workload at 3^3 cells per body (32 cubes and balls, 2,048 vertices) over 80
    frames: bodies fall, hit the ground and each other. Stiff contact (k = 1e4 against
    m / dt^2 of about 5e3) makes the impact steps ill-conditioned (the FP32 and FP64 oracles part
    by up to ~2e-4 there), so the per-step errors are checked against that spread: the median
    within 2x of the FP32-vs-FP64 median (or 1e-6), and the worst step within 4x of the worst
    FP32-vs-FP64 step (or 1e-5)."""
    bodies = scenes.c2_bodies(n=3)
    spec = dict(physics=physics_dict(**scenes.c2_physics()), steps=80,
                meshes=[dict(pos=b["pos"], tets=b["tets"], material=b["material"], fixed=[], vel=None,
                             coloring="greedy") for b in bodies])
    errs, errs64 = parity(gpu_ctx, oracle, spec, with_fp64=True)
    print(f"\n[tet contact, 32 bodies] gpu-vs-fp32 median {np.median(errs):.2e} max {errs.max():.2e}; "
          f"fp32-vs-fp64 median {np.median(errs64):.2e} max {errs64.max():.2e}", flush=True)
    assert np.median(errs) <= max(2.0 * np.median(errs64), 1e-6)
    assert errs.max() <= max(4.0 * errs64.max(), TOL)


def test_soft_tet_apex_closed_form(gpu_ctx):
    """The oracle pin on the device: a soft tet's apex at d over a large fixed triangle moves to
    (k r + M d) / (M + k) in one Newton step."""
    d, k, r, L = 0.004, 100.0, 0.01, 0.05
    big = np.array([[-10.0, 0.0, -10.0], [10.0, 0.0, -10.0], [0.0, 0.0, 10.0]])
    tet_pos = np.array([[0.0, d, 0.0], [L, d + L, 0.0], [-L, d + L, -L], [0.0, d + L, L]])
    vol = np.linalg.det(np.stack([tet_pos[1] - tet_pos[0], tet_pos[2] - tet_pos[0], tet_pos[3] - tet_pos[0]])) / 6
    density = 1000.0 * DT ** 2 * 4 / vol
    sim = VBDPhysics(gpu_ctx)
    sim.setPhysicsParams(**physics_dict(gravity=(0.0, 0.0, 0.0), time_step=DT, num_substeps=1, iterations=1,
                                        world_lower=(-1e7, -1e7, -1e7), handle_collision=1, contact_stiffness=k,
                                        contact_radius=r, thickness=0.0, conservative_step_relaxation=0.45,
                                        contact_energy_type=0, contact_query_radius=0.02))
    sim.addClothMesh(big, np.array([[0, 2, 1]], np.int32), fixed_points=[0, 1, 2],
                     vertex_colors=np.arange(3, dtype=np.int32), density=1.0, mu=0.0, lmbd=0.0)
    sim.addTetMesh(tet_pos, np.array([[0, 1, 2, 3]], np.int32), vertex_colors=np.arange(4, dtype=np.int32),
                   density=density, mu=1e-9, lmbd=1e-9)
    sim.initialize()
    sim.runStep()
    x, _ = sim.getState(1)
    want = tet_pos.copy()
    want[0, 1] = (k * r + 1000.0 * d) / (1000.0 + k)
    assert rel_err(x, want) < TOL
    sim.destroy()


def test_contact_scene_without_contacts_matches_plain_scene(gpu_ctx):
    """One vertex per colour, a contact radius below every distance and a query radius that keeps
    every step inside its bound: the contact path (boundary-surface detection, positionsNext and
    commit) steps as the plain tet path. The two are separately compiled kernel variants (the
    plain one carries no contact code), so they agree to rounding, not bit for bit (the oracle's
    composition equals its tet solver bit for bit)."""
    pos, tets = scenes.cube_tet_mesh(2, 0.05, origin=(0.0, 0.1, 0.0))
    pos = pos + np.random.default_rng(3).normal(scale=0.003, size=pos.shape)
    colors = np.arange(len(pos), dtype=np.int32)
    mat = dict(density=1000.0, mu=1e5, lmbd=1e6, damping_hydrostatic=0.01, damping_deviatoric=0.02)
    out = []
    for contact in (1, 0):
        sim = VBDPhysics(gpu_ctx)
        sim.setPhysicsParams(**physics_dict(gravity=(0.0, -9.8, 0.0), time_step=DT, num_substeps=1, iterations=4,
                                            world_lower=(-1e7, 0.0, -1e7), handle_collision=contact,
                                            contact_radius=1e-7, contact_query_radius=1.0))
        sim.addTetMesh(pos, tets, vertex_colors=colors, **mat)
        sim.initialize()
        sim.runSteps(3)
        out.append(sim.getState(0))
        sim.destroy()
    assert rel_err(out[0][0], out[1][0]) < 1e-6
    # velocities are position differences over dt: compared on the positions' scale
    assert np.abs(out[0][1] - out[1][1]).max() * DT / np.abs(out[1][0]).max() < 1e-6


def test_tet_contact_json_default_off(gpu_ctx, tmp_path):
    """A tet-only scene loaded without handleCollision keeps the tet framework's default (off): it
    has no contact counters (gvbd_scene_contact_stats exists only for scenes with collision
    handling)."""
    pos, tets = scenes.cube_tet_mesh(2, 0.1, origin=(0.0, 0.5, 0.0))
    models, params = scenes.write_scene_json(str(tmp_path), "cube", pos, tets,
                                             physics_dict(iterations=2, num_substeps=1),
                                             dict(density=1000.0, mu=1e5, lmbd=1e6))
    sim = VBDPhysics(gpu_ctx)
    sim.loadRunningparameters(models, params, repo_root=str(tmp_path))
    sim.initialize()
    sim.recordVBDSolveGraph(True)
    sim.runStep()
    with pytest.raises(gaia_amd.VBDError, match="no collision handling"):
        sim.contactStats()
    sim.destroy()


def test_contact_stats_count_frames_and_detections(gpu_ctx):
    """gvbd_scene_contact_stats: frames and iterations stepped with collision handling, at least
    one detection pass per frame, detection time measured; scenes without collision handling
    have no such counters."""
    import gaia_amd

    spec = tet_drop_spec(steps=6)
    sim = build_gpu(gpu_ctx, spec)
    sim.runSteps(6)
    st = sim.contactStats()
    iters = spec["physics"]["iterations"]
    assert st["frames"] == 6 and st["iterations"] == 6 * iters
    assert 6 <= st["detections"] <= 6 * iters and st["rebuilds"] >= 1
    assert st["detect_ms"] > 0 and st["vf_pairs"] >= 0 and st["ee_pairs"] >= 0
    sim.destroy()
    pos, tets = scenes.cube_tet_mesh(2, 0.1, origin=(0.0, 0.5, 0.0))
    plain = VBDPhysics(gpu_ctx)
    plain.setPhysicsParams(handle_collision=0)
    plain.addTetMesh(pos, tets)
    plain.initialize()
    with pytest.raises(gaia_amd.VBDError, match="no collision handling"):
        plain.contactStats()
    plain.destroy()
