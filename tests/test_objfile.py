"""OBJ / MTL scenes (rsh_amd.objfile; reference OBJ_Reader.cpp:29-381).

A small OBJ written here checks the reading rules: objects per ``o`` group,
1-based and negative indices, ``vn`` normals against face normals, polygons
split into fans, the per-object ``usemtl``, the MTL fields and the placement
(scale, position, offset).  The reference's own office.obj is read when the
reference is present: the thesis scene's 299 objects and 36,308 triangles,
and a low-resolution oracle frame through the reference camera.

The Cornell box is assembled from per-object transforms (the reference's
Transform: translate * quaternion rotation * scale, checked against scipy's
rotations) over a stand-in set of model files written here, and, where the
reference is present, from its own files, whose 512x512 frame is compared
with the ray and candidate counts of the reference's recorded Cornell run
(output/tests/configuration-n8/test-crsh-d2-n8-cornell.txt).
"""
import os

import numpy as np
import pytest

from rsh_amd import objfile

MODELS = "/root/reference/projects/CUDA-Engine/models"
OFFICE = MODELS + "/office.obj"
CORNELL_LOG = "/root/reference/output/tests/configuration-n8/test-crsh-d2-n8-cornell.txt"

OBJ = """\
# two objects
mtllib t.mtl
o quad
v 0 0 0
v 2 0 0
v 2 1 0
v 0 1 0
usemtl red
f 1 2 3 4
o tri
v 0 0 1
v 1 0 1
v 0 1 1
vn 0 0 2
usemtl blue
f -3//1 -2//1 -1//1
f 5 6 7
"""

MTL = """\
newmtl red
Kd 1 0 0
Ks 0.5 0.5 0.5
Ns 20
newmtl blue
Kd 0 0 1
"""


@pytest.fixture()
def small(tmp_path):
    (tmp_path / "t.obj").write_text(OBJ)
    (tmp_path / "t.mtl").write_text(MTL)
    return str(tmp_path / "t.obj")


def test_reading_rules(small):
    objs, libs = objfile.read_obj(small)
    assert libs == ["t.mtl"]
    assert [o["name"] for o in objs] == ["quad", "tri"]
    assert [o["mat"] for o in objs] == ["red", "blue"]
    assert len(objs[0]["tris"]) == 2                      # the quad as a fan
    assert objs[0]["tris"][1] == [(0, 0, 0), (2, 1, 0), (0, 1, 0)]
    assert objs[1]["tris"][0] == [(0, 0, 1), (1, 0, 1), (0, 1, 1)]       # negative indices
    assert objs[1]["nrm"][0] == [(0, 0, 2)] * 3 and objs[1]["nrm"][1] is None
    m = objfile.parse_mtl(small[:-3] + "mtl")
    assert m["red"] == ((1.0, 0.0, 0.0), (0.5, 0.5, 0.5), 20.0)
    assert m["blue"][0] == (0.0, 0.0, 1.0)


def test_scene_layout(small):
    s = objfile.load_obj(small, position=(1, 2, 3), scale=10.0, offset=(50, 50, 50),
                         reflective=lambda name, kd, ks, ns: name == "blue")
    assert s.triangle_total == 4 and s.object_total == 2
    assert np.array_equal(s.object_spheres[:, 4:6], [[0, 1], [2, 3]])
    v = s.tri_pos.reshape(-1, 3, 4)
    assert np.allclose(v[0, 1, :3], np.array([2, 0, 0]) * 10 + [51, 52, 53])
    n = s.tri_nrm.reshape(-1, 3, 4)
    assert np.allclose(n[0, :, :3], [0, 0, 1])            # face normal of the quad
    assert np.allclose(n[2, :, :3], [0, 0, 1])            # vn, normalised
    assert np.allclose(s.mat_diffuse[:, 3], [0, 1])       # blue reflective
    assert s.tri_mat.tolist() == [0] * 6 + [1] * 6
    for k, (a, b) in enumerate(s.object_spheres[:, 4:6].astype(int)):
        pts = v[a:b + 1, :, :3].reshape(-1, 3)
        assert (np.linalg.norm(pts - s.object_spheres[k, :3], axis=1) <= s.object_spheres[k, 3] * (1 + 1e-6)).all()


@pytest.mark.skipif(not os.path.exists(OFFICE), reason="the reference's office.obj is not present")
def test_reference_office_scene():
    from oracle import pyoracle
    from rsh_amd import config as cf
    s = objfile.office_scene(OFFICE)
    assert s.object_total == 299 and s.triangle_total == 36308
    cam = objfile.office_camera(64, 64)
    fr = pyoracle.render(s, cam, cf.config1(64, 64).with_scene_bounds(s), threads=0)
    assert 0.3 < (np.asarray(fr.pixels) != 0).mean() < 1.0    # the room fills much of the view


def test_model_matrix():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(5)
    for _ in range(20):
        pos, rot, scl = rng.uniform(-50, 50, 3), rng.uniform(-180, 180, 3), rng.uniform(0.1, 10, 3)
        M = objfile.model_matrix(pos, rot, scl)
        R = Rotation.from_euler("XYZ", rot, degrees=True).as_matrix()      # Rx @ Ry @ Rz
        p = rng.normal(size=3)
        assert np.allclose(M[:3, :3] @ p + M[:3, 3], R @ (scl * p) + pos)
        assert np.array_equal(M[3], [0, 0, 0, 1])
    # quarter turns are exact: about z, x -> y; about x, y -> z
    assert np.array_equal(objfile.model_matrix((0, 0, 0), (0, 0, 90), (1, 1, 1))[:3, :3] @ [1, 0, 0], [0, 1, 0])
    assert np.array_equal(objfile.model_matrix((0, 0, 0), (90, 0, 0), (1, 1, 1))[:3, :3] @ [0, 1, 0], [0, 0, 1])


def test_cornell_layout(tmp_path):
    import cornell_models
    n = cornell_models.write_cornell_models(str(tmp_path))
    s = objfile.cornell_scene(str(tmp_path), offset=(0, 0, 0), reflective=lambda k: k == 6)
    assert s.object_total == 7 and s.triangle_total == 6 * 12 + n
    v = s.tri_pos.reshape(-1, 3, 4)[:, :, :3]
    boxes = [((-100, 49.75, -100), (100, 50.25, 100)), ((-100, -50.25, -100), (100, -49.75, 100)),
             ((-100.25, -50, -100), (-99.75, 50, 100)), ((99.75, -50, -100), (100.25, 50, 100)),
             ((-100, -50, 99.75), (100, 50, 100.25)), ((-100, -50, -100.25), (100, 50, -99.75))]
    rg = s.object_spheres[:, 4:6].astype(int)
    for k, (lo, hi) in enumerate(boxes):
        p = v[rg[k, 0]:rg[k, 1] + 1].reshape(-1, 3)
        assert np.allclose(p.min(0), lo) and np.allclose(p.max(0), hi), k
    p = v[rg[6, 0]:rg[6, 1] + 1].reshape(-1, 3)
    assert np.allclose(np.linalg.norm(p - [0, -12.5, 0], axis=1), 25, rtol=1e-6)
    # each wall has its own file's material; only the sphere is reflective here
    kd = [tuple(np.round(s.mat_diffuse[s.tri_mat[3 * a], :3], 6)) for a in rg[:, 0]]
    assert kd[:6] == [cornell_models.WALL_KD[m] for m in ("silver.mtl", "silver.mtl", "gold.mtl", "ruby.mtl",
                                                           "saphire.mtl", "emerald.mtl")]
    assert [s.mat_diffuse[s.tri_mat[3 * a], 3] for a in rg[:, 0]] == [0] * 6 + [1]
    # normals: the walls' inner faces point into the room
    nr = s.tri_nrm.reshape(-1, 3, 4)[:, :, :3]
    for k, (axis, at) in enumerate(((1, 49.75), (1, -49.75), (0, -99.75), (0, 99.75), (2, 99.75), (2, -99.75))):
        t = np.arange(rg[k, 0], rg[k, 1] + 1)
        face = t[np.all(np.isclose(v[t, :, axis], at), axis=1)]
        assert len(face) == 2, k
        want = np.zeros(3)
        want[axis] = -np.sign(at)
        assert np.allclose(nr[face], want), k
    assert np.allclose(s.light_pos[0, :3], [0, 45, 0]) and np.allclose(s.light_col[0, :3], 0.8)
    cam = objfile.cornell_camera(8, 8, offset=(0, 0, 0))
    eye = np.asarray(cam.eye, np.float64)
    assert (np.abs(eye) < [100, 50, 100]).all()                # the camera is inside the room


@pytest.mark.skipif(not (os.path.exists(MODELS + "/Cube.obj") and os.path.exists(CORNELL_LOG)),
                    reason="the reference's Cornell model files and recorded run are not present")
def test_reference_cornell_scene():
    """The reference's recorded Cornell run at 512x512 (REF keys, depth 2,
    8 rays per leaf): 242,015 shadow rays, then 262,144 rays in each of two
    reflection passes (every pixel sees a wall), with 1,394,752 / 1,493,384 /
    1,792,462 leaf candidate pairs.  The pinhole camera here stands in for the
    reference's raster, so the counts are close, not equal."""
    import dataclasses
    from oracle import pyoracle
    from rsh_amd import config as cf
    log = open(CORNELL_LOG).read().split("Iteration")
    rays = [int(b.split("Ray Total:")[1].split()[0]) for b in log[1:]]
    pairs = [int(b.split("[Final 0] Connected Hit Total:")[1].split()[0]) for b in log[1:]]
    assert rays == [242015, 262144, 262144]
    s = objfile.cornell_scene(MODELS)
    assert s.object_total == 7 and s.triangle_total == 792
    c = dataclasses.replace(cf.config1(512, 512), reflect_gate=0, bounces=2, bounce_shadows=0,
                            key_mode=cf.KEY_REF, cone_mode=cf.CONE_REF).with_scene_bounds(s)
    fr = pyoracle.render(s, objfile.cornell_camera(512, 512), c, threads=0)
    got = [p for p in fr.passes if p["kind"] != 0]
    assert [p["kind"] for p in got] == [1, 2, 2]
    assert abs(got[0]["rays"] / rays[0] - 1) < 0.03
    assert [p["rays"] for p in got[1:]] == rays[1:]
    for p, ref in zip(got, pairs):
        assert abs(p["pairs_leaf"] / ref - 1) < 0.1, (p["pairs_leaf"], ref)
