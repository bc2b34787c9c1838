"""Writes cornell_box.json: the reference's default scene (the Cornell box
that TracerAPI.initDefaultMesh builds, src/cl/ui/fx/main/TracerAPI.java:
205-371) as data: the OBJ geometry it hands to OBJParser.readString
(vertices, vertex normals, polygon corners v//n, one group per `o` line)
and the material settings it then applies per group (setDiffuse /
setEmitter, CMaterial2.java:26-37, on top of CSurfaceParameter2's defaults,
CSurfaceParameter2.java:41-55).

The geometry is read from the Java string literal and parsed here into
numbers; nothing of the Java code is kept.  Run once in a container that has
the reference at /root/reference:  python tests/golden/make_cornell.py
"""
import json
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/cl/ui/fx/main/TracerAPI.java"


def main():
    text = open(SRC).read()
    start = text.index("String cube")
    end = text.index(";", start)
    obj = "".join(re.findall(r'"((?:[^"\\]|\\.)*)"', text[start:end])).replace("\\n", "\n")
    v, vn, polys, groups = [], [], [], []
    for line in obj.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "o":
            groups.append(t[1])
        elif t[0] == "v":
            v.append([float(x) for x in t[1:4]])
        elif t[0] == "vn":
            vn.append([float(x) for x in t[1:4]])
        elif t[0] == "f":
            corners = [[int(c.split("/")[0]), int(c.split("/")[2])] for c in t[1:]]
            polys.append({"group": len(groups) - 1, "corners": corners})
    tail = text[end:text.index("bvhBuild", end)]
    names = dict(re.findall(r"CMaterial2 (\w+)\s*=\s*materialsc\.get\((\d+)\)", tail))
    mats = {}
    for var, idx in re.findall(r"CMaterial2 (\w+)\s*=\s*materialsc\.get\((\d+)\)", tail):
        mats[int(idx)] = {}
    for var, r, g, b in re.findall(r"^\s*(\w+)\.setDiffuse\(([-\d.f]+),\s*([-\d.f]+),\s*([-\d.f]+)\)", tail, re.M):
        mats[int(names[var])]["diffuse"] = [float(x.rstrip("f")) for x in (r, g, b)]
    for var, r, g, b in re.findall(r"^\s*(\w+)\.setEmitter\(([-\d.f]+),\s*([-\d.f]+),\s*([-\d.f]+)\)", tail, re.M):
        mats[int(names[var])]["emitter"] = [float(x.rstrip("f")) for x in (r, g, b)]
    out = {
        "source": "TracerAPI.java:205-371 (initDefaultMesh)",
        "groups": groups,
        "v": v,
        "vn": vn,
        "polygons": polys,
        "materials": {str(k): mats[k] for k in sorted(mats)},
        "defaults": {"diffuse": [0.95, 0.95, 0.95], "diffuse_param_x": 1.0, "glossy": [0.95, 0.95, 0.95],
                     "glossy_param": [0.0, 0.0, 0.0], "mirror": [0.95, 0.95, 0.95],
                     "emission": [1.0, 1.0, 1.0], "emission_param": [0.0, 0.0, 0.0],
                     "set_emitter_param": [1.0, 15.0, 0.0]},
    }
    with open(os.path.join(HERE, "cornell_box.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"{len(v)} vertices, {len(vn)} normals, {len(polys)} polygons, {len(groups)} groups, "
          f"materials {sorted(mats)}")


if __name__ == "__main__":
    main()
