"""A stand-in for the reference's Cornell model files, written by the tests.

The files have the names and layout that rsh_amd.objfile.cornell_scene reads
(Cube.obj, the five wall MTL files, sphere/sphere.obj, sphere/gold.mtl): a
unit cube of 12 triangles with per-face ``vn`` records, as in the reference's
Cube.obj, and a UV sphere of radius 1 with smooth normals (the reference's
sphere.obj has 720 triangles; this one is smaller unless asked otherwise).
Material values are made up; every file names its material "cube", as the
reference's do.
"""
import math
import os

WALL_KD = {"silver.mtl": (0.5, 0.5, 0.5), "gold.mtl": (0.75, 0.6, 0.23), "ruby.mtl": (0.61, 0.04, 0.04),
           "saphire.mtl": (0.0, 0.51, 0.5), "emerald.mtl": (0.08, 0.61, 0.08)}


def cube_obj():
    v = [(x, y, z) for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)]
    idx = {p: i + 1 for i, p in enumerate(v)}
    lines = ["mtllib cube.mtl", "o cube"] + ["v %r %r %r" % p for p in v]
    faces = []
    for axis in range(3):
        for s in (-0.5, 0.5):
            # the four corners of the face in counter-clockwise order seen from outside
            a, b = [k for k in range(3) if k != axis]
            corners = []
            for u, w in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
                p = [0.0, 0.0, 0.0]
                p[axis], p[a], p[b] = s, u, w
                corners.append(tuple(p))
            n = [0.0, 0.0, 0.0]
            n[axis] = 1.0 if s > 0 else -1.0
            # order so that the face normal points outwards
            e1 = [corners[1][k] - corners[0][k] for k in range(3)]
            e2 = [corners[2][k] - corners[0][k] for k in range(3)]
            c = (e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0])
            if sum(c[k] * n[k] for k in range(3)) < 0:
                corners = corners[::-1]
            faces.append((corners, tuple(n)))
    lines.append("usemtl cube")
    for k, (corners, n) in enumerate(faces):
        lines.append("vn %r %r %r" % n)                   # normal k + 1
        i = [idx[p] for p in corners]
        for tri in ((i[0], i[1], i[2]), (i[0], i[2], i[3])):
            lines.append("f " + " ".join("%d//%d" % (t, k + 1) for t in tri))
    return "\n".join(lines) + "\n"


def sphere_obj(slices=12, stacks=7):
    verts = [(0.0, 1.0, 0.0)]
    for i in range(1, stacks):
        th = math.pi * i / stacks
        for j in range(slices):
            ph = 2 * math.pi * j / slices
            verts.append((math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)))
    verts.append((0.0, -1.0, 0.0))
    ring = lambda i, j: 2 + (i - 1) * slices + (j % slices)      # 1-based
    bottom = len(verts)
    tris = []
    for j in range(slices):
        tris.append((1, ring(1, j + 1), ring(1, j)))
        tris.append((bottom, ring(stacks - 1, j), ring(stacks - 1, j + 1)))
    for i in range(1, stacks - 1):
        for j in range(slices):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j), ring(i + 1, j + 1)
            tris += [(a, b, d), (a, d, c)]
    lines = ["mtllib sphere.mtl", "o sphere"] + ["v %r %r %r" % p for p in verts]
    lines += ["vn %r %r %r" % p for p in verts]          # unit sphere: normal = position
    lines += ["f " + " ".join("%d//%d" % (t, t) for t in tri) for tri in tris]
    return "\n".join(lines) + "\n", len(tris)


def mtl(name, kd, ks=(0.5, 0.5, 0.5), ns=102.0):
    return "newmtl %s\nNs %r\nKd %r %r %r\nKs %r %r %r\n" % (name, ns, *kd, *ks)


def write_cornell_models(path, slices=12, stacks=7):
    """Write the files under `path`; returns the sphere's triangle count."""
    os.makedirs(os.path.join(path, "sphere"), exist_ok=True)
    with open(os.path.join(path, "Cube.obj"), "w") as f:
        f.write(cube_obj())
    for name, kd in WALL_KD.items():
        with open(os.path.join(path, name), "w") as f:
            f.write(mtl("cube", kd))
    text, n = sphere_obj(slices, stacks)
    with open(os.path.join(path, "sphere", "sphere.obj"), "w") as f:
        f.write(text)
    with open(os.path.join(path, "sphere", "gold.mtl"), "w") as f:
        f.write(mtl("sphere", WALL_KD["gold.mtl"], (0.63, 0.56, 0.37), 51.2))
    return n
