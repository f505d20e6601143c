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


def first_material(path):
    """(Kd, Ks, Ns) of the first material of an MTL file (the reference's
    Cornell walls each name their own file, whose material is always called
    "cube")."""
    m = parse_mtl(path)
    return next(iter(m.values())) if m else ((0.8, 0.8, 0.8), (0.0, 0.0, 0.0), 1.0)


def _assemble(objects, materials, lights, offset, name) -> Scene:
    """Scene from placed objects.  objects: (tris [n, 3, 3], normals [n, 3, 3],
    material index) in world coordinates before `offset`; materials: (Kd, Ks,
    Ns, reflective) per index.  Normals are normalised here."""
    from . import object_spheres
    off = np.asarray(offset, np.float64)
    pos, nrm, mat, ranges = [], [], [], []
    count = 0
    for tris, nn, m in objects:
        n = tris.shape[0]
        nn = nn / np.maximum(np.linalg.norm(nn, axis=2, keepdims=True), 1e-30)
        pos.append((tris + off).reshape(-1, 3))
        nrm.append(nn.reshape(-1, 3))
        mat.append(np.full(3 * n, m, np.int32))
        ranges.append((count, count + n - 1))
        count += n
    P = np.concatenate(pos).astype(np.float32)
    Nv = np.concatenate(nrm).astype(np.float32)
    tri_pos = np.ascontiguousarray(np.hstack([P, np.ones((len(P), 1), np.float32)]))
    lo, hi = P.min(0), P.max(0)
    pad = 0.01 * float((hi - lo).max())
    return Scene(
        name=name,
        tri_pos=tri_pos,
        tri_nrm=np.ascontiguousarray(np.hstack([Nv, np.zeros((len(Nv), 1), np.float32)])),
        tri_mat=np.concatenate(mat).astype(np.int32),
        mat_diffuse=np.array([(*kd, 1.0 if r else 0.0) for kd, ks, ns, r in materials], np.float32),
        mat_specular=np.array([(*ks, ns) for kd, ks, ns, r in materials], np.float32),
        object_spheres=object_spheres(tri_pos, np.array(ranges, np.uint32)),
        light_pos=np.array([(*(np.asarray(l[0], np.float64) + off), 1.0) for l in lights], np.float32),
        light_col=np.array([(*l[1], 1.0) for l in lights], np.float32),
        light_int=np.array([l[2] for l in lights], np.float32),
        bounds_min=(lo - pad).astype(np.float32),
        bounds_max=(hi + pad).astype(np.float32),
    )


def _normals(o, tris):
    """Per-vertex normals of an object read by read_obj: `vn` where the face
    names them, else the face normal of the (unplaced) triangle."""
    face = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    nn = np.repeat(face[:, None, :], 3, axis=1)
    for i, vn in enumerate(o["nrm"]):
        if vn is not None:
            nn[i] = np.asarray(vn, np.float64)
    return nn


def load_obj(path, position=(0.0, 0.0, 0.0), scale=1.0, offset=(0.0, 0.0, 0.0), lights=None,
             reflective=lambda name, kd, ks, ns: False, name=None) -> Scene:
    """Scene of an OBJ file (its MTL libraries are read from the same folder).
    lights: list of (position, colour, (diffuse, specular intensity)) in the
    file's placed coordinates (offset is added to them)."""
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
    materials = []
    for m in names:
        kd, ks, ns = mats.get(m, ((0.8, 0.8, 0.8), (0.0, 0.0, 0.0), 1.0))
        materials.append((kd, ks, ns, reflective(m, kd, ks, ns)))
    shift = np.asarray(position, np.float64)
    placed = []
    for o in objs:
        raw = np.asarray(o["tris"], np.float64)
        placed.append((raw * scale + shift, _normals(o, raw), names.index(o["mat"] if o["mat"] in mats else None)))
    lights = lights or [((0.0, 10.0, 0.0), (1.0, 1.0, 1.0), (0.75, 0.75))]
    return _assemble(placed, materials, lights, offset, name or os.path.splitext(os.path.basename(path))[0])


def model_matrix(position, rotation, scale):
    """The reference's Transform model matrix (Transform.cpp:40-46,124-138):
    translate(position) * rotate(q) * scale(scale), with q the product of the
    rotations by rotation[0], [1], [2] degrees about x, y and z, in that
    order.  A unit quaternion (w, x, y, z) rotates by the usual matrix
    [[1-2(yy+zz), 2(xy-zw), 2(xz+yw)], [2(xy+zw), 1-2(xx+zz), 2(yz-xw)],
    [2(xz-yw), 2(yz+xw), 1-2(xx+yy)]] (Matrix.cpp:10-40 builds its transpose
    and quaternionRotate transposes it back).  Computed in double."""
    def quat(deg, axis):
        a = math.radians(deg) / 2.0
        return np.array([math.cos(a), *(math.sin(a) * np.asarray(axis, np.float64))])

    def qmul(p, q):
        w1, x1, y1, z1 = p
        w2, x2, y2, z2 = q
        return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])

    q = qmul(qmul(quat(rotation[0], (1, 0, 0)), quat(rotation[1], (0, 1, 0))), quat(rotation[2], (0, 0, 1)))
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    R[np.abs(R) < 1e-12] = 0.0          # exact quarter turns
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(scale, np.float64)[None, :3]
    M[:3, 3] = np.asarray(position, np.float64)[:3]
    return M


def place(o, M):
    """Triangles and normals of an object read by read_obj, moved by the 4x4
    model matrix M (normals by its inverse transpose)."""
    raw = np.asarray(o["tris"], np.float64)
    tris = raw @ M[:3, :3].T + M[:3, 3]
    nn = _normals(o, raw) @ np.linalg.inv(M[:3, :3])       # (M^-T n) as row vectors
    return tris, nn


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


# The reference's Cornell box (main.cpp:1824-1838 light, 1927-1940 camera,
# 2020-2103 objects): six walls, each the unit cube of Cube.obj scaled to a
# 0.5-thick slab, turned a quarter about z (left, right) or x (front, back)
# and placed on a face of a 200 x 100 x 200 room centred at the origin, with
# one material file per wall; a gold sphere (sphere/sphere.obj, radius 1)
# scaled by 25 at (0, -12.5, 0); one light of colour 0.8 at (0, 45, 0) with
# intensities 0.75 / 0.75.  7 objects, 792 triangles.
CORNELL_WALLS = (  # position, rotation (degrees about x, y, z), scale, material file
    ((0.0, 50.0, 0.0), (0.0, 0.0, 0.0), (200.0, 0.5, 200.0), "silver.mtl"),      # top
    ((0.0, -50.0, 0.0), (0.0, 0.0, 0.0), (200.0, 0.5, 200.0), "silver.mtl"),     # bottom
    ((-100.0, 0.0, 0.0), (0.0, 0.0, 90.0), (100.0, 0.5, 200.0), "gold.mtl"),     # left
    ((100.0, 0.0, 0.0), (0.0, 0.0, 90.0), (100.0, 0.5, 200.0), "ruby.mtl"),      # right
    ((0.0, 0.0, 100.0), (90.0, 0.0, 0.0), (200.0, 0.5, 100.0), "saphire.mtl"),   # front
    ((0.0, 0.0, -100.0), (90.0, 0.0, 0.0), (200.0, 0.5, 100.0), "emerald.mtl"),  # back
)


def _model_file(models, *names):
    """First of `names` that exists in the models folder (the reference's
    code names "cube.obj" for the file shipped as Cube.obj)."""
    for n in names:
        p = os.path.join(models, n)
        if os.path.exists(p):
            return p
    raise FileNotFoundError(os.path.join(models, names[0]))


def cornell_scene(models, offset=OFFICE_OFFSET, reflective=lambda k: False) -> Scene:
    """The reference's Cornell box from its models folder (Cube.obj, the five
    wall MTL files, sphere/sphere.obj and sphere/gold.mtl).  reflective(k):
    whether object k (walls 0-5 in the order above, the sphere 6) mirrors,
    for the configs' reflect gate; the reference's own rule reflects from
    every covered pixel whatever the material."""
    cube = read_obj(_model_file(models, "cube.obj", "Cube.obj"))[0]
    sphere = read_obj(_model_file(models, os.path.join("sphere", "sphere.obj")))[0]
    objects, materials = [], []
    for k, (pos, rot, scl, mtl) in enumerate(CORNELL_WALLS):
        for o in cube:
            objects.append((*place(o, model_matrix(pos, rot, scl)), k))
        materials.append((*first_material(_model_file(models, mtl)), reflective(k)))
    for o in sphere:
        objects.append((*place(o, model_matrix((0.0, -12.5, 0.0), (0.0, 0.0, 0.0), (25.0, 25.0, 25.0))), 6))
    materials.append((*first_material(_model_file(models, os.path.join("sphere", "gold.mtl"))), reflective(6)))
    return _assemble(objects, materials, [((0.0, 45.0, 0.0), (0.8, 0.8, 0.8), (0.75, 0.75))], offset, "cornell")


def cornell_camera(width=512, height=512, offset=OFFICE_OFFSET) -> Camera:
    return orbit_camera((4.35, -24.35, 1.50), 4.5, 0.5, 225.0, width, height, offset)
