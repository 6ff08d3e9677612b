"""CPU-side checks of the product library: the C ABI loads and exports every
declared symbol, and its host-side preparation (colouring, rest-pose constants)
matches the reference fixtures and the oracle. No GPU calls."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import gaia_amd
from gaia_amd import scenes
from gaia_amd.vbd import tet_mesh_constants
from conftest import GOLDEN, load_t

FIXTURES = ["Beam_lowRes", "Beam", "Cube", "bunny_small"]


def test_library_exports_every_declared_symbol():
    lib = gaia_amd.lib()
    declared = gaia_amd.header_functions()
    assert len(declared) >= 30
    for name in declared:
        assert hasattr(lib, name), name
    # and the dynamic symbol table has them as text symbols
    out = subprocess.run(["nm", "-D", "--defined-only", gaia_amd.LIB_PATH], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared) <= exported
    # nothing else: the C++ internals (and the standard library's instantiations) stay hidden
    everything = {l.split()[-1] for l in out.splitlines() if len(l.split()) >= 3}
    assert everything == set(declared), sorted(everything - set(declared))[:5]
    # ctypes binding covers every declared function
    assert sorted(declared) == sorted(gaia_amd._lib._SIGNATURES)


def test_abi_struct_layout_matches_header():
    from gaia_amd._lib import ClothMeshDesc, PhysicsParams, Schedule, TetMeshDesc

    a, b, c, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert gaia_amd.lib().gvbd_abi_struct_sizes(ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                                ctypes.byref(d)) == 0
    assert ctypes.sizeof(PhysicsParams) == a.value == 200
    assert PhysicsParams.tet_contact_model.offset == 136 and PhysicsParams.dcd_bvh_rebuild_steps.offset == 196
    assert ctypes.sizeof(TetMeshDesc) == b.value
    assert ctypes.sizeof(ClothMeshDesc) == c.value
    assert ctypes.sizeof(Schedule) == d.value == 32
    assert TetMeshDesc.positions.offset == 8 and TetMeshDesc.tets.offset == 40


@pytest.mark.parametrize("name", FIXTURES)
def test_product_mcs_matches_reference_fixture(name):
    pos, tets = load_t(os.path.join(GOLDEN, "coloring", name + ".t"))
    want = json.load(open(os.path.join(GOLDEN, "coloring", name + ".t.vertexColoring.json")))
    c = gaia_amd.color_tet_mesh(len(pos), tets, "mcs")
    got = [sorted(np.flatnonzero(c == k).tolist()) for k in range(c.max() + 1)]
    assert got == [sorted(x) for x in want]


@pytest.mark.parametrize("name", FIXTURES)
def test_product_greedy_matches_oracle(oracle, name):
    pos, tets = load_t(os.path.join(GOLDEN, "coloring", name + ".t"))
    assert np.array_equal(gaia_amd.color_tet_mesh(len(pos), tets, "greedy"),
                          oracle.color_tet_mesh(len(pos), tets, "greedy"))


def test_soft_cube_colouring_sizes(oracle):
    """10^3-cell cube with the alternating 5-tet split: greedy colours of sizes
    665, 216, 126, 138, 110, 49, 19, 8 and MCS 341, 330, 330, 330."""
    pos, tets = scenes.cube_tet_mesh(10, 0.1)
    g = gaia_amd.color_tet_mesh(len(pos), tets, "greedy")
    m = gaia_amd.color_tet_mesh(len(pos), tets, "mcs")
    assert np.bincount(g).tolist() == [665, 216, 126, 138, 110, 49, 19, 8]
    assert np.bincount(m).tolist() == [341, 330, 330, 330]
    assert np.array_equal(g, oracle.color_tet_mesh(len(pos), tets, "greedy"))
    assert np.array_equal(m, oracle.color_tet_mesh(len(pos), tets, "mcs"))
    inc = np.bincount(tets.ravel())
    assert inc.max() == 32 and abs(inc.mean() - 15.03) < 0.01


def test_rest_pose_constants_match_oracle(oracle):
    pos, tets = scenes.cube_tet_mesh(4, 0.1)
    rng = np.random.default_rng(0)
    pos = pos + rng.normal(scale=0.005, size=pos.shape)
    dminv, vol, mass = tet_mesh_constants(pos, tets, 1000.0)
    odm, ovol, omass = oracle.tet_prepare(pos.astype(np.float32), tets, 1000.0, precision=32)
    assert np.allclose(dminv, odm, rtol=1e-6, atol=1e-6 * np.abs(odm).max())
    assert np.allclose(vol, ovol, rtol=1e-6)
    assert np.allclose(mass, omass, rtol=1e-6)
    assert np.all(vol > 0)


def test_errors_are_reported_not_raised_in_c():
    with pytest.raises(gaia_amd.VBDError, match="out of range"):
        gaia_amd.color_tet_mesh(3, np.array([[0, 1, 2, 5]]))
    with pytest.raises(gaia_amd.VBDError, match="degenerate"):
        tet_mesh_constants(np.zeros((4, 3)), np.array([[0, 1, 2, 3]]), 1.0)


def test_tet_contact_workload_bodies():
    """The tet-contact workload's generator: 16 cubes and 16 balls of 8,640 positively oriented
    tets, inside the 2 m box and at least 3 cm apart from each other (rotations included)."""
    from scipy.spatial import cKDTree

    bodies = scenes.c2_bodies()
    assert [b["kind"] for b in bodies].count("cube") == 16 and len(bodies) == 32
    allp = np.concatenate([b["pos"] for b in bodies])
    label = np.concatenate([np.full(len(b["pos"]), i) for i, b in enumerate(bodies)])
    pairs = cKDTree(allp).query_pairs(0.03, output_type="ndarray")
    assert not np.any(label[pairs[:, 0]] != label[pairs[:, 1]])
    assert allp[:, 1].min() > 0 and np.abs(allp[:, [0, 2]]).max() < 1.0
    for b in bodies[:2]:
        a = b["pos"][b["tets"]]
        vol = np.einsum("ij,ij->i", np.cross(a[:, 1] - a[:, 0], a[:, 2] - a[:, 0]), a[:, 3] - a[:, 0])
        assert len(b["tets"]) == 8640 and vol.min() > 0
    mus = np.array([b["material"]["mu"] for b in bodies])
    assert mus.min() >= 1e4 and mus.max() <= 1e6
    assert all(b["material"]["lmbd"] == 10 * b["material"]["mu"] for b in bodies)


def test_deformer_and_attachment_descriptor_layouts():
    """ctypes mirrors of gvbd_deformer_desc / gvbd_attachment_desc match the C layout the library
    asserts at compile time (96 and 40 bytes on LP64)."""
    import ctypes

    from gaia_amd._lib import AttachmentDesc, DeformerDesc

    assert ctypes.sizeof(DeformerDesc) == 96 and DeformerDesc.angular_velocity.offset == 48
    assert DeformerDesc.meshes.offset == 72
    assert ctypes.sizeof(AttachmentDesc) == 40 and AttachmentDesc.bary.offset == 24
