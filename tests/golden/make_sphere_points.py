"""Writes sphere_points_15.json: the reference's fixed specular directions
(assets/shaders/voxelLighting.comp:23), the only constant table the reference
holds for this path. Reads the shader as text; run in a tree with the reference.
"""
import json
import re
import sys
from pathlib import Path

SRC = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/assets/shaders/voxelLighting.comp")
line = next(l for l in SRC.read_text().splitlines() if "spherePoints[15]" in l)
vals = [float(v) for v in re.findall(r"-?\d+\.\d+", line)]
assert len(vals) == 45
out = Path(__file__).with_name("sphere_points_15.json")
out.write_text(json.dumps({"source": "assets/shaders/voxelLighting.comp:23", "points": [vals[i:i + 3] for i in range(0, 45, 3)]}, indent=1) + "\n")
print(out)
