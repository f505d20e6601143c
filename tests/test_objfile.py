"""OBJ / MTL scenes (rsh_amd.objfile; reference OBJ_Reader.cpp:29-381).

A small OBJ written here checks the reading rules: objects per ``o`` group,
1-based and negative indices, ``vn`` normals against face normals, polygons
split into fans, the per-object ``usemtl``, the MTL fields and the placement
(scale, position, offset).  The reference's own office.obj is read when the
reference is present: the thesis scene's 299 objects and 36,308 triangles,
and a low-resolution oracle frame through the reference camera.
"""
import os

import numpy as np
import pytest

from rsh_amd import objfile

OFFICE = "/root/reference/projects/CUDA-Engine/models/office.obj"

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
