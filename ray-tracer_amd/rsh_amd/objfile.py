"""Wavefront OBJ / MTL scenes (the reference's OBJ_Reader.cpp:29-381 and
Material loading), and the reference's thesis office scene set-up.

An OBJ file is read as the reference reads it: one object per ``o`` group, in
file order; ``v`` / ``vn`` / ``vt`` records; ``f`` faces whose indices are
1-based over the whole file (the reference subtracts the previous objects'
counts because it re-reads per object, which is the same thing); the object's
material is the ``usemtl`` inside its group.  The reference reads three
vertices per face; polygons with more are split here into a fan (a
triangle-only file, such as the reference's office.obj, loads identically).
Vertex normals come from ``vn`` when the face names them, otherwise from the
face itself.  MTL files give ``Kd``, ``Ks`` and ``Ns`` per ``newmtl``.

Every object is placed by ``p * scale + position`` (the reference's Transform
with no rotation, main.cpp:1986-1988 for the office) plus ``offset`` (this
library's scenes keep surfaces away from the world origin, which the
reference's shading takes for background, quirk Q13).  Object spheres are the
smallest enclosing balls (rsh_object_spheres).
"""
from __future__ import annotations

import math
import os

import numpy as np

from .scenes import Camera, Scene, camera_basis


def parse_mtl(path):
    """{name: (Kd[3], Ks[3], Ns)} of an MTL file."""
    mats, cur = {}, None
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            if t[0] == "newmtl":
                cur = t[1] if len(t) > 1 else ""
                mats[cur] = [[0.8, 0.8, 0.8], [0.0, 0.0, 0.0], 1.0]
            elif cur is not None and t[0] == "Kd":
                mats[cur][0] = [float(x) for x in t[1:4]]
            elif cur is not None and t[0] == "Ks":
                mats[cur][1] = [float(x) for x in t[1:4]]
            elif cur is not None and t[0] == "Ns":
                mats[cur][2] = float(t[1])
    return {k: (tuple(v[0]), tuple(v[1]), float(v[2])) for k, v in mats.items()}


def _index(tok, n):
    """1-based OBJ index (negative: relative to the end) -> 0-based."""
    i = int(tok)
    return i - 1 if i > 0 else n + i


def read_obj(path):
    """Objects of an OBJ file: list of (name, material, tris [n, 3, 3],
    normals [n, 3, 3] or None per triangle), plus the MTL file names."""
    V, N = [], []
    objects, mtllibs = [], []
    cur = None

    def new(name):
        nonlocal cur
        cur = {"name": name, "mat": None, "tris": [], "nrm": []}
        objects.append(cur)

    with open(path) as f:
        for line in f:
            t = line.split()
            if not t:
                continue
            k = t[0]
            if k == "v":
                V.append((float(t[1]), float(t[2]), float(t[3])))
            elif k == "vn":
                N.append((float(t[1]), float(t[2]), float(t[3])))
            elif k == "o":
                new(t[1] if len(t) > 1 else "")
            elif k == "usemtl":
                if cur is None:
                    new("")
                if cur["mat"] is None:
                    cur["mat"] = t[1] if len(t) > 1 else ""
            elif k == "mtllib":
                mtllibs.extend(t[1:])
            elif k == "f":
                if cur is None:
                    new("")
                idx = []
                for tok in t[1:]:
                    parts = tok.split("/")
                    vi = _index(parts[0], len(V))
                    ni = _index(parts[2], len(N)) if len(parts) >= 3 and parts[2] else -1
                    idx.append((vi, ni))
                for j in range(1, len(idx) - 1):          # fan
                    tri = (idx[0], idx[j], idx[j + 1])
                    cur["tris"].append([V[v] for v, _ in tri])
                    cur["nrm"].append([N[n] for _, n in tri] if all(n >= 0 for _, n in tri) else None)
    return [o for o in objects if o["tris"]], mtllibs


def load_obj(path, position=(0.0, 0.0, 0.0), scale=1.0, offset=(0.0, 0.0, 0.0), lights=None,
             reflective=lambda name, kd, ks, ns: False, name=None) -> Scene:
    """Scene of an OBJ file (its MTL libraries are read from the same folder).
    lights: list of (position, colour, (diffuse, specular intensity)) in the
    file's placed coordinates (offset is added to them)."""
    from . import object_spheres
    objs, libs = read_obj(path)
    mats = {}
    for lib in libs:
        p = os.path.join(os.path.dirname(path), lib)
        if os.path.exists(p):
            mats.update(parse_mtl(p))
    names = []
    for o in objs:
        m = o["mat"] if o["mat"] in mats else None
        if m not in names:
            names.append(m)
    mat_d, mat_s = [], []
    for m in names:
        kd, ks, ns = mats.get(m, ((0.8, 0.8, 0.8), (0.0, 0.0, 0.0), 1.0))
        mat_d.append((*kd, 1.0 if reflective(m, kd, ks, ns) else 0.0))
        mat_s.append((*ks, ns))
    shift = np.asarray(position, np.float64) + np.asarray(offset, np.float64)
    pos, nrm, mat, ranges = [], [], [], []
    count = 0
    for o in objs:
        tris = np.asarray(o["tris"], np.float64) * scale + shift
        n = tris.shape[0]
        face = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
        face /= np.maximum(np.linalg.norm(face, axis=1, keepdims=True), 1e-30)
        nn = np.repeat(face[:, None, :], 3, axis=1)
        for i, vn in enumerate(o["nrm"]):
            if vn is not None:
                v = np.asarray(vn, np.float64)
                nn[i] = v / np.maximum(np.linalg.norm(v, axis=1, keepdims=True), 1e-30)
        pos.append(tris.reshape(-1, 3))
        nrm.append(nn.reshape(-1, 3))
        mat.append(np.full(3 * n, names.index(o["mat"] if o["mat"] in mats else None), np.int32))
        ranges.append((count, count + n - 1))
        count += n
    P = np.concatenate(pos).astype(np.float32)
    Nv = np.concatenate(nrm).astype(np.float32)
    tri_pos = np.ascontiguousarray(np.hstack([P, np.ones((len(P), 1), np.float32)]))
    lights = lights or [((0.0, 10.0, 0.0), (1.0, 1.0, 1.0), (0.75, 0.75))]
    lo, hi = P.min(0), P.max(0)
    pad = 0.01 * float((hi - lo).max())
    return Scene(
        name=name or os.path.splitext(os.path.basename(path))[0],
        tri_pos=tri_pos,
        tri_nrm=np.ascontiguousarray(np.hstack([Nv, np.zeros((len(Nv), 1), np.float32)])),
        tri_mat=np.concatenate(mat).astype(np.int32),
        mat_diffuse=np.array(mat_d, np.float32),
        mat_specular=np.array(mat_s, np.float32),
        object_spheres=object_spheres(tri_pos, np.array(ranges, np.uint32)),
        light_pos=np.array([(*(np.asarray(l[0], np.float64) + np.asarray(offset)), 1.0) for l in lights], np.float32),
        light_col=np.array([(*l[1], 1.0) for l in lights], np.float32),
        light_int=np.array([l[2] for l in lights], np.float32),
        bounds_min=(lo - pad).astype(np.float32),
        bounds_max=(hi + pad).astype(np.float32),
    )


# The reference's office scene (main.cpp:1811-1833 light, 1912-1923 camera,
# 1975-1990 objects): office.obj placed at (0, -15, 0), scale 10, one white
# light at (-15, 15, 7.5) with intensities 0.75 / 0.75.
OFFICE_OFFSET = (50.0, 50.0, 50.0)


def office_scene(obj_path, offset=OFFICE_OFFSET) -> Scene:
    return load_obj(obj_path, position=(0.0, -15.0, 0.0), scale=10.0, offset=offset,
                    lights=[((-15.0, 15.0, 7.5), (1.0, 1.0, 1.0), (0.75, 0.75))], name="office")


def orbit_camera(target, zoom, latitude, longitude, width, height, offset=(0.0, 0.0, 0.0), fov=60.0) -> Camera:
    """Camera::update's orbit (Camera.cpp:94-96, CAMERA_RADIUS 25, Camera.h:40):
    eye = target + zoom * 25 * (sin(lat - 90) cos(lon), cos(lat - 90), -sin(lat - 90) sin(lon))."""
    r = zoom * 25.0
    la, lo = math.radians(latitude - 90.0), math.radians(longitude)
    t = np.asarray(target, np.float64) + np.asarray(offset, np.float64)
    eye = t + r * np.array([math.sin(la) * math.cos(lo), math.cos(la), -math.sin(la) * math.sin(lo)])
    return camera_basis(eye, t, (0.0, 1.0, 0.0), fov, width, height)


def office_camera(width=1024, height=1024, offset=OFFICE_OFFSET) -> Camera:
    return orbit_camera((-4.50, 7.25, 1.50), 0.5, 30.5, 215.0, width, height, offset)
