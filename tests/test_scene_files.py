"""Scene files in the reference's formats, host side (no device): recovery states
(PhysicsState::toJson / fromJson, Framework/BasePhysicsFramework.h:145-319) in text and binary
form, and the ASCII PLY writer (GAIA::saveAsPly, IO/FileIO.cpp:8-50), through the C ABI."""
import json
import os

import numpy as np
import pytest

import gaia_amd
from conftest import load_t

REF_KNOT = "/root/reference/Data/mesh_models/ModelsDifferentRes/knots/OverhandKnot_5.t"


def meshes(seed=0):
    rng = np.random.default_rng(seed)
    pos = [rng.normal(size=(n, 3)).astype(np.float32) for n in (5, 1, 7)]
    vel = [rng.normal(size=p.shape).astype(np.float32) for p in pos]
    return pos, vel


@pytest.mark.parametrize("binary", [False, True])
def test_state_round_trip_bitwise(tmp_path, binary):
    pos, vel = meshes()
    path = str(tmp_path / "A00000012.json")
    gaia_amd.state_write(path, pos, vel, frame_id=12, cur_time=0.2, binary=binary)
    fid, t, p2, v2 = gaia_amd.state_read(path)
    assert fid == 12 and t == 0.2
    assert len(p2) == 3
    for a, b in zip(pos + vel, p2 + v2):
        assert np.array_equal(a, b)
    if binary:
        # float32 xyz of every mesh back to back, beside the JSON
        assert os.path.getsize(tmp_path / "A00000012.pos") == 4 * 3 * 13
        raw = np.fromfile(tmp_path / "A00000012.vel", dtype=np.float32).reshape(-1, 3)
        assert np.array_equal(raw, np.concatenate(vel))


def test_state_json_layout(tmp_path):
    """The keys and nesting PhysicsState::toJson writes (nlohmann::json objects keep keys sorted)."""
    pos, vel = meshes(1)
    gaia_amd.state_write(str(tmp_path / "t.json"), pos, vel, frame_id=3, cur_time=0.05)
    d = json.load(open(tmp_path / "t.json"))
    assert list(d) == ["curTime", "format", "frameId", "meshesState"]
    assert d["format"] == "text" and d["frameId"] == 3 and d["curTime"] == 0.05
    assert list(d["meshesState"][0]) == ["position", "velocities"]
    assert np.array_equal(np.float32(d["meshesState"][2]["velocities"]), vel[2])
    gaia_amd.state_write(str(tmp_path / "b.json"), pos, vel, frame_id=3, cur_time=0.05, binary=True)
    d = json.load(open(tmp_path / "b.json"))
    assert list(d) == ["curTime", "format", "frameId", "meshesState", "nMeshes"]
    assert d["format"] == "binary" and d["nMeshes"] == 3
    assert d["meshesState"] == {"nVerts": [5, 1, 7], "position_file": "b.pos", "velocities_file": "b.vel"}


def test_state_without_velocities_reads_zeros(tmp_path):
    """The layout of the reference's shipped initial-state files (positions only, keys curTime,
    frameId, meshesState; no "format", so text)."""
    (tmp_path / "s.json").write_text(json.dumps({"curTime": 0.0, "frameId": 0, "meshesState": [
        {"position": [[0.15, 0.0, 0.0], [0.15087604988156325, -0.008957261668892621, 1.2490009027033013e-18]]}]}))
    fid, t, p, v = gaia_amd.state_read(str(tmp_path / "s.json"))
    assert fid == 0 and t == 0.0
    assert np.array_equal(p[0], np.float32([[0.15, 0.0, 0.0], [0.15087604988156325, -0.008957261668892621,
                                                               1.2490009027033013e-18]]))
    assert np.array_equal(v[0], np.zeros((2, 3), np.float32))


def test_state_errors(tmp_path):
    with pytest.raises(gaia_amd.VBDError, match="cannot open"):
        gaia_amd.state_read(str(tmp_path / "missing.json"))
    (tmp_path / "bad.json").write_text('{"curTime": 0, "frameId": 0, "format": "xml", "meshesState": []}')
    with pytest.raises(gaia_amd.VBDError, match="unknown format"):
        gaia_amd.state_read(str(tmp_path / "bad.json"))
    (tmp_path / "short.json").write_text(json.dumps({"curTime": 0, "frameId": 0, "format": "binary", "nMeshes": 1,
                                                    "meshesState": {"nVerts": [4], "position_file": "short.pos",
                                                                    "velocities_file": "short.vel"}}))
    np.zeros(3, np.float32).tofile(tmp_path / "short.pos")
    np.zeros(3, np.float32).tofile(tmp_path / "short.vel")
    with pytest.raises(gaia_amd.VBDError, match="too short"):
        gaia_amd.state_read(str(tmp_path / "short.json"))


def test_ply_text_matches_reference_format(tmp_path):
    """GAIA::saveAsPly's exact text: header (with the face element's trailing space), "%f %f %f"
    vertices, "3 %d %d %d" faces; no face element for a point set."""
    pos = np.float32([[0.0, 1.5, -2.25], [1e-7, 3.0, 0.125], [2.0, -0.5, 1.0]])
    gaia_amd.write_ply(str(tmp_path / "a.ply"), pos, np.int32([[0, 2, 1]]))
    want = ("ply\nformat ascii 1.0\ncomment Mocap generated\nelement vertex 3\n"
            "property float x\nproperty float y\nproperty float z\n"
            "element face 1 \nproperty list uchar int vertex_indices\nend_header\n"
            + "".join("%f %f %f\n" % tuple(float(c) for c in p) for p in pos) + "3 0 2 1\n")
    assert (tmp_path / "a.ply").read_text() == want
    gaia_amd.write_ply(str(tmp_path / "b.ply"), pos[:1], np.zeros((0, 3), np.int32))
    assert "element face" not in (tmp_path / "b.ply").read_text()
    with pytest.raises(gaia_amd.VBDError, match="out of range"):
        gaia_amd.write_ply(str(tmp_path / "c.ply"), pos, np.int32([[0, 1, 3]]))


@pytest.mark.skipif(not os.path.exists(REF_KNOT), reason="reference data not present")
def test_reads_the_reference_initial_state_file():
    """The reference's shipped initial state of the 5-resolution overhand knot: one mesh with the
    vertex count of its .t file, positions equal to the JSON's numbers in float32."""
    fid, t, p, v = gaia_amd.state_read(REF_KNOT + ".initialState.json")
    pos, _ = load_t(REF_KNOT)
    assert fid == 0 and t == 0.0 and len(p) == 1 and p[0].shape == (len(pos), 3)
    d = json.load(open(REF_KNOT + ".initialState.json"))
    assert np.array_equal(p[0], np.float32(d["meshesState"][0]["position"]))
    assert not np.any(v[0])
