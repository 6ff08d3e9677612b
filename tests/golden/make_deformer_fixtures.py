"""Writes tests/golden/deformers/reference_deformers.json: the "Deformers" arrays of the
reference's twisting (S01) and knot-pulling (S02) experiment parameter files, as data (the
selected vertex lists and motion parameters), for the schema tests in tests/test_oracle_deformers.py.

Run from the repository root in a container that holds the reference at /root/reference."""
import json
import os

ROOT = "/root/reference/Simulator/VBDCloth/ParameterGen/Parameters"
FILES = {
    "S01_Twisting": "S01_Twisting/C50_Twist_ContactR0.2_CStif100000.0_bd20_damping3e-05_relaxiation0.42/Parameters.json",
    "S02_KnotPulling": "S02_KnotPulling/Knot_ContactR0.2_CStif1000000.0_bd50_damping1e-07_relaxiation0.42_pull10/"
                       "Parameters.json",
}

out = {}
for key, rel in FILES.items():
    with open(os.path.join(ROOT, rel)) as f:
        out[key] = json.load(f)["Deformers"]
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "deformers", "reference_deformers.json")
with open(dst, "w") as f:
    json.dump(out, f)
print("wrote", dst)
