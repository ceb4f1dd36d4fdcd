"""Writes the data fixtures the reference's own tests hold, as plain data.

Run in a container where the reference checkout is present:
    python tests/golden/make_fixtures.py /root/reference
Outputs (committed):
    teapot.npz            Tests/Helpers/TeapotModel.h arrays (positions, normals,
                          uvs, indices) - the mesh of the reference's renderer tests
    material_layout.json  (name, byte offset) of every MaterialConstants member in
                          Tests/Assets/TextFiles/DefaultMaterialUniformBuffer.slang,
                          the expected-output fixture of TestUniformBuffer.cpp:42-80
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _array(text, name, dtype):
    m = re.search(r"std::array<[^>]+>\s+%s\s*\{\s*\{(.*?)\}\s*\}" % name, text, re.S)
    vals = [v.strip().rstrip("f") for v in m.group(1).split(",") if v.strip()]
    return np.array([float(v) for v in vals], dtype=dtype)


def main(ref):
    text = open(os.path.join(ref, "Tests/Helpers/TeapotModel.h")).read()
    pos = _array(text, "gTeapotVertices", np.float32).reshape(-1, 3)
    nrm = _array(text, "gTeapotNormals", np.float32).reshape(-1, 3)
    uv = _array(text, "gTeapotUVs", np.float32).reshape(-1, 2)
    idx = _array(text, "gTeapotIndices", np.float64).astype(np.uint32).reshape(-1, 3)
    np.savez(os.path.join(HERE, "teapot.npz"), positions=pos, normals=nrm, uvs=uv, indices=idx)
    layout = []
    src = open(os.path.join(ref, "Tests/Assets/TextFiles/DefaultMaterialUniformBuffer.slang")).read()
    for m in re.finditer(r"^\s*(float|float2|float3|int)\s+(\w+);\s*//\s*Offset:(\d+)\s+Property:(\w+)", src, re.M):
        layout.append({"type": m.group(1), "member": m.group(2), "offset": int(m.group(3)),
                       "property": m.group(4)})
    size = None
    with open(os.path.join(HERE, "material_layout.json"), "w") as f:
        json.dump({"members": layout, "size": size}, f, indent=1)
    print(pos.shape, nrm.shape, uv.shape, idx.shape, len(layout))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
