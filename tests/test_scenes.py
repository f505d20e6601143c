"""Synthetic scene generators: layouts, counts and bounding spheres."""
import numpy as np
import pytest


def check_scene(s):
    T = s.triangle_total
    assert s.tri_pos.shape == (3 * T, 4) and s.tri_nrm.shape == (3 * T, 4)
    assert s.tri_pos.dtype == np.float32 and s.tri_mat.dtype == np.int32
    # objects tile [0, T) in order
    first = s.object_spheres[:, 4].astype(np.int64)
    last = s.object_spheres[:, 5].astype(np.int64)
    assert first[0] == 0 and last[-1] == T - 1
    assert np.all(first[1:] == last[:-1] + 1)
    # every object's sphere contains its vertices
    for k in range(s.object_total):
        v = s.tri_pos[3 * first[k]:3 * (last[k] + 1), :3].astype(np.float64)
        c = s.object_spheres[k, :3].astype(np.float64)
        assert np.linalg.norm(v - c, axis=1).max() <= s.object_spheres[k, 3]
    # one material per object, ids in range
    assert s.tri_mat.min() >= 0 and s.tri_mat.max() < s.mat_diffuse.shape[0]
    # no surface at the world origin (it reads as "no fragment")
    assert np.linalg.norm(s.tri_pos[:, :3], axis=1).min() > 1.0
    nrm = np.linalg.norm(s.tri_nrm[:, :3], axis=1)
    np.testing.assert_allclose(nrm, 1.0, atol=1e-5)


def test_config0_scene(rsh):
    s = rsh.scenes.config0_scene()
    assert s.triangle_total == 12 + 2 * 1024
    assert s.light_total == 1
    check_scene(s)


def test_config1_scene(rsh):
    s = rsh.scenes.config1_scene()
    assert s.triangle_total == 12 + 2500 * 12
    assert s.light_total == 2
    refl = s.mat_diffuse[:, 3] > 0
    assert 0.3 < refl.mean() < 0.5
    check_scene(s)


def test_config2_scene(rsh):
    s = rsh.scenes.config2_scene()
    assert abs(s.triangle_total - 262_144) / 262_144 < 0.01
    refl = s.mat_diffuse[:, 3] > 0
    assert 0.2 < refl.mean() < 0.4
    check_scene(s)


def test_scenes_are_seeded(rsh):
    a, b = rsh.scenes.config0_scene(), rsh.scenes.config0_scene()
    np.testing.assert_array_equal(a.tri_pos, b.tri_pos)


def test_camera_basis_matches_reference_formula(rsh):
    cam = rsh.scenes.camera_basis([0, 0, 0], [0, 0, 10], [0, 1, 0], 60.0, 200, 100)
    assert np.linalg.norm(cam.direction) == pytest.approx(10, rel=1e-6)
    # Camera.cpp:114-115: "half" height = 2 * distance * tan(fov/2)
    assert np.linalg.norm(cam.up) == pytest.approx(2 * 10 * np.tan(np.radians(30)), rel=1e-5)
    assert np.linalg.norm(cam.right) == pytest.approx(2 * 2 * 10 * np.tan(np.radians(30)), rel=1e-5)
    assert abs(np.dot(cam.right, cam.up)) < 1e-4
