#!/usr/bin/env python3
"""Writes the small fixtures the CPU tests use from data files and test inputs
held in the reference tree (read as text, never executed):

  stripes_graph.json   the 200 graph vertices of datasets/stripes/.../graph.obj
                       and the two node blocks of LoadDeformScript0
                       (GaussianView.cpp:2517-2518; pinned type 0, dragged type 1)
  helper_tests.json    the input matrices of RunTestVKVt (helper.cpp:869-920)
                       and the quaternion + 48 SH inputs of Test_SH_Rotation
                       (helper.cpp:925-928)
  pinocchio_deform.txt datasets/pinocchio/.../deform.txt verbatim (recorded
                       deformation session: node list and edit operations)

Run once in a container that has the reference at REF (default /root/reference);
the outputs are committed and nothing at test time reads the reference.
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
GV = os.path.join(REF, "src/projects/gaussianviewer/renderer")


def stripes_graph():
    path = os.path.join(REF, "datasets/stripes/point_cloud/iteration_30000/graph.obj")
    verts, faces = [], []
    with open(path) as f:
        for line in f:
            t = line.split()
            if t and t[0] == "v":
                verts.append([float(v) for v in t[1:4]])
            elif t and t[0] == "f":
                faces.append([int(v.split("/")[0]) - 1 for v in t[1:]])
    with open(os.path.join(GV, "GaussianView.cpp")) as f:
        src = f.read()
    body = src[src.index("void sibr::GaussianView::LoadDeformScript0()"):]
    blocks = re.findall(r"vector<unsigned int> block(\d) = \{([0-9, ]+)\};", body[:2000])
    blk = {int(k): [int(v) for v in s.split(",")] for k, s in blocks}
    return {"source": "datasets/stripes/point_cloud/iteration_30000/graph.obj; blocks of LoadDeformScript0 "
                      "(GaussianView.cpp:2517-2518), types {0, 1}",
            "nodes": verts, "faces": faces, "block1": blk[1], "block2": blk[2], "block_types": [0, 1]}


def helper_tests():
    with open(os.path.join(GV, "helper.cpp")) as f:
        src = f.read()
    vk = src[src.index("RunTestVKVt"):]
    vk = vk[: vk.index("void Test_SH_Rotation")]
    def mats(name):
        out = []
        for m in re.finditer(r"^\s*" + name + r"\s*<<\s*([-0-9.,\s]+);", vk, re.M):
            out.append([float(v) for v in m.group(1).replace("\n", " ").split(",")])
        return out
    tests = [{"mat": a, "R0": b} for a, b in zip(mats("mat"), mats("R0"))]
    sh = src[src.index("void Test_SH_Rotation"):]
    q = [float(v) for v in re.search(r"origin_q\(([^)]+)\)", sh).group(1).split(",")]
    shs = [float(v) for v in re.search(r"shs_in = \{([^}]+)\}", sh).group(1).split(",")]
    return {"source": "helper.cpp RunTestVKVt (:869-920) and Test_SH_Rotation (:925-928) inputs",
            "vkvt": tests, "sh_rotation": {"q_wxyz": q, "shs_in": shs}}


def main():
    with open(os.path.join(OUT, "stripes_graph.json"), "w") as f:
        json.dump(stripes_graph(), f)
    with open(os.path.join(OUT, "helper_tests.json"), "w") as f:
        json.dump(helper_tests(), f, indent=1)
    shutil.copyfile(os.path.join(REF, "datasets/pinocchio/point_cloud/iteration_30000/deform.txt"),
                    os.path.join(OUT, "pinocchio_deform.txt"))


if __name__ == "__main__":
    main()
