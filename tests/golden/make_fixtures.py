"""Copies the reference's known-answer colouring fixtures into tests/golden/coloring.

The reference holds no test suite; its only known-answer data for this path are
tet meshes shipped with their MCS vertex colourings (Data/mesh_models/**,
written by the GraphColoring tool, Simulator/GraphColoring/GraphColoring.cpp).
These four are data files (meshes + colour categories), copied verbatim.
Run in the build container only (the reference is not present on the GPU box).
"""
import os
import shutil

REF = "/root/reference/Data/mesh_models"
HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = {
    "Beam_lowRes": "UnitTest/Beam/Beam_lowRes.t",
    "Beam": "UnitTest/Beam/Beam.t",
    "Cube": "ModelsDifferentRes/Cube/Cube.t",
    "bunny_small": "t/bunny_small.t",
}

if __name__ == "__main__":
    out = os.path.join(HERE, "coloring")
    os.makedirs(out, exist_ok=True)
    for name, rel in FIXTURES.items():
        shutil.copyfile(os.path.join(REF, rel), os.path.join(out, name + ".t"))
        shutil.copyfile(os.path.join(REF, rel + ".vertexColoring.json"),
                        os.path.join(out, name + ".t.vertexColoring.json"))
        print(name)
