"""Native engine (include/rsh.h) against the oracle frame, pass by pass.

Each case renders the same seeded scene and camera on the GPU and with the
CPU oracle and compares live-ray counts, packet counts, every pass's per-slot
hit records, the primary G-buffer and the image.  The bar is the parity
target (hit IDs identical on >= 99.99% of rays, images within 1/255); the
implementation is expected to be bit-exact and the strict cases assert that.
"""
import dataclasses
import sys

import numpy as np
import pytest

from gpu_common import compare_frames

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def engine(rsh):
    e = rsh.Engine(0)
    yield e
    e.close()


def run_case(engine, oracle, rsh, scene, cam, cfg, strict=True):
    cfg = cfg.with_scene_bounds(scene) if cfg.key_mode == 1 else cfg
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.KEEP_PASSES)
    ofr = oracle.render(scene, cam, cfg, threads=0)
    rep = compare_frames(engine, ofr, cfg, strict=strict)
    print(cfg.width, cfg.height, rep, file=sys.stderr)
    return rep


def c0(rsh, w):
    return rsh.scenes.config0_scene(), rsh.scenes.config0_camera(w, w)


@pytest.mark.parametrize("w", [64, 128, 512])
def test_config0_default(engine, oracle, rsh, w):
    scene, cam = c0(rsh, w)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config0(w, w))


@pytest.mark.parametrize("variant", [
    dict(algorithm=1),                       # RAH: unsorted hierarchy, no object culling
    dict(key_mode=0),                        # reference spherical keys
    dict(hit_mode=1),                        # reference 32-bit hit packing
    dict(cone_mode=0),                       # reference cone union (NaN nodes)
    dict(depth=3, branch=4),
    dict(depth=1),
    dict(leaf=16, branch=4),
    dict(leaf=4, branch=16),
    dict(leaf=32, branch=2),
    dict(leaf=2, branch=32),
    dict(traversal_kernel=1),                # generic kernel on the 64-ray shape
    dict(traversal_kernel=1, algorithm=1),
    dict(algorithm=1, bounces=2, bounce_shadows=1),
    dict(traversal_chunk=1),                 # one cluster per work item, merged by atomicMin
    dict(traversal_chunk=3),                 # items that start and end inside objects
    dict(traversal_chunk=3, algorithm=1),
    dict(traversal_chunk=1, bounces=2, bounce_shadows=1),
    dict(traversal_chunk=1 << 30),           # one item per top node
    dict(traversal_splits=3, traversal_kernel=1),
    dict(traversal_splits=5, traversal_kernel=1, bounces=2, bounce_shadows=1),
    dict(traversal_splits=4, depth=3, branch=4),
    dict(key_layout=1),
    dict(key_layout=2),                      # bit-interleaved cell / direction key
    dict(traversal_prefilter=1),             # without the ray pre-tests
    dict(traversal_prefilter=1, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=1, leaf=4, branch=16, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=1, traversal_chunk=1, hit_mode=1),
    dict(traversal_prefilter=2, bounces=2, bounce_shadows=1),   # triangle pre-test only
    dict(traversal_prefilter=3),             # and cluster ray masks on every pass
    dict(traversal_prefilter=3, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, leaf=4, branch=16, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, leaf=32, branch=2),
    dict(traversal_prefilter=3, traversal_chunk=1, hit_mode=1),
    dict(triangle_batch=512),
    dict(bounces=2, bounce_shadows=1),
])
def test_config0_variants(engine, oracle, rsh, variant):
    scene, cam = c0(rsh, 96)
    cfg = dataclasses.replace(rsh.config.config0(96, 96), **variant)
    run_case(engine, oracle, rsh, scene, cam, cfg)


def test_config1_small(engine, oracle, rsh):
    scene = rsh.scenes.config1_scene(boxes=300)
    cam = rsh.scenes.config1_camera(128, 128)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config1(128, 128))


def test_config2_small(engine, oracle, rsh):
    scene = rsh.scenes.config2_scene(grid=3, col_slices=16, col_stacks=12, tiles=8, tile_sub=2, arch_around=6,
                                     arch_along=8)
    cam = rsh.scenes.config2_camera(160, 90)
    run_case(engine, oracle, rsh, scene, cam, rsh.config.config2(160, 90))


@pytest.fixture(scope="module")
def atrium(rsh):
    return rsh.scenes.config2_scene()


def test_config2_full_scene_low_res(engine, oracle, rsh, atrium):
    """The full 262k-triangle atrium: many top nodes whose 64 shadow rays are
    all occluded part-way through a batch of candidate blocks (regression:
    the survivor queue must not grow past its capacity after the wave exits)."""
    cam = rsh.scenes.config2_camera(192, 108)
    run_case(engine, oracle, rsh, atrium, cam, rsh.config.config2(192, 108))


def test_config2_full_frame_matches_oracle(engine, oracle, rsh, atrium):
    """The bench's workload at its full size: the 1920x1080 atrium frame, 3
    bounces with shadow rays at every hit (4.45M rays), with every engine
    path the timed frames take (front-to-back camera pass in two launches,
    reach-pruned shadow items, cluster ray masks, 256-cluster items), against
    the oracle pass by pass: live rays, packets, per-slot hit records, the
    G-buffer and the image, all bit-exact."""
    cam = rsh.scenes.config2_camera(1920, 1080)
    run_case(engine, oracle, rsh, atrium, cam, rsh.config.config2(1920, 1080))


@pytest.mark.parametrize("variant", [
    dict(),
    dict(algorithm=1),
    dict(traversal_prefilter=1),
    dict(traversal_prefilter=1, bounces=2, bounce_shadows=1),
    dict(traversal_prefilter=3, bounces=2, bounce_shadows=1),   # counting keeps every ray
    dict(traversal_chunk=1, hit_mode=1),
    dict(depth=3, branch=4),                 # generic kernel
    dict(leaf=16, branch=4, bounces=2, bounce_shadows=1),
])
def test_traversal_counts_match_oracle(engine, oracle, rsh, variant):
    """Candidate (top node, triangle) pairs, passing (leaf, triangle) pairs and
    ray-triangle tests equal the oracle's for every closest-hit pass (the
    cluster pre-test and the work split must not change the candidate sets).
    Shadow passes stop early once every ray of a wave is occluded, so their
    leaf pairs and tests may only be fewer."""
    scene, cam = c0(rsh, 96)
    cfg = dataclasses.replace(rsh.config.config0(96, 96), **variant).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.COUNTERS)
    gp = engine.passes()
    op = oracle.render(scene, cam, cfg, threads=0, keep_hits=False).passes
    assert len(gp) == len(op)
    for g, o in zip(gp, op):
        assert g["pairs_top"] == o["pairs_top"], (g, o)
        if g["kind"] == 1:
            assert g["pairs_leaf"] <= o["pairs_leaf"] and g["ray_tests"] <= o["ray_tests"], (g, o)
        else:
            assert (g["pairs_leaf"], g["ray_tests"]) == (o["pairs_leaf"], o["ray_tests"]), (g, o)
        assert g["tri_tests"] <= g["pairs_top"]
        assert g["mt_tests"] <= g["ray_tests"]


def test_config2_counts_match_oracle(engine, oracle, rsh):
    """Same on an atrium slice, where clusters do cull."""
    scene = rsh.scenes.config2_scene(grid=3, col_slices=16, col_stacks=12, tiles=8, tile_sub=2, arch_around=6,
                                     arch_along=8)
    cam = rsh.scenes.config2_camera(160, 90)
    cfg = rsh.config.config2(160, 90).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.COUNTERS)
    gp = engine.passes()
    op = oracle.render(scene, cam, cfg, threads=0, keep_hits=False).passes
    for g, o in zip(gp, op):
        assert g["pairs_top"] == o["pairs_top"]
        if g["kind"] != 1:
            assert (g["pairs_leaf"], g["ray_tests"]) == (o["pairs_leaf"], o["ray_tests"])
    assert sum(g["tri_tests"] for g in gp) < sum(g["pairs_top"] for g in gp)


def test_render_twice_is_deterministic(engine, rsh):
    scene, cam = c0(rsh, 128)
    cfg = rsh.config.config0(128, 128).with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg)
    a = engine.pixels()
    engine.render(cam, cfg)
    np.testing.assert_array_equal(a, engine.pixels())


def test_invalid_arguments_are_rejected(engine, rsh):
    scene, cam = c0(rsh, 32)
    engine.load_scene(scene)
    bad = dataclasses.replace(rsh.config.config0(32, 32), leaf=0)
    with pytest.raises(rsh.RshError):
        engine.render(cam, bad)
    bad = dataclasses.replace(rsh.config.config0(32, 32), cell_bits=8, dir_bits=8)
    with pytest.raises(rsh.RshError):
        engine.render(cam, bad)
    broken = dataclasses.replace(scene, object_spheres=scene.object_spheres.copy())
    broken.object_spheres[1, 4] += 1.0        # gap in the object tiling
    with pytest.raises(rsh.RshError):
        engine.load_scene(broken)


def _object_transforms(scene, seed):
    """Model / normal / scale matrices per object (row-major 4x4): the room
    stays, every other object turns about its sphere's centre, scales and
    moves a little (the reference's per-frame update, main.cpp:1131-1196)."""
    rng = np.random.default_rng(seed)
    K = scene.object_total
    model, normal, scale = (np.tile(np.eye(4), (K, 1, 1)) for _ in range(3))
    for k in range(1, K):
        c = scene.object_spheres[k, :3].astype(np.float64)
        a = rng.normal(size=3)
        a /= np.linalg.norm(a)
        th = rng.uniform(0.2, np.pi)
        X = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        R = np.eye(3) + np.sin(th) * X + (1 - np.cos(th)) * X @ X
        s = rng.uniform(0.8, 1.2)
        M3 = R * s
        model[k, :3, :3] = M3
        model[k, :3, 3] = c + rng.uniform(-3.0, 3.0, 3) - M3 @ c
        normal[k, :3, :3] = np.linalg.inv(M3).T
        scale[k] = np.diag([s, s, s, 1.0])
    return [m.astype(np.float32).reshape(K, 16) for m in (model, normal, scale)]


def _moved_scene(oracle, scene, model, normal, scale):
    """The scene after the oracle's update wrappers (RayTracer.cu:448-570)."""
    T, K = scene.triangle_total, scene.object_total
    obj = np.repeat(scene.triangle_object(), 3).astype(np.int32)
    sph4 = np.zeros((2 * K, 4), np.float32)
    sph4[0::2, :3] = scene.object_spheres[:, :3]
    sph4[0::2, 3] = 1.0
    sph4[1::2, :3] = scene.object_spheres[:, 3:6]
    sph4[1::2, 3] = 1.0
    oracle.call("oracle_bind_original", scene.tri_pos, scene.tri_nrm, obj, sph4)
    pos, nrm = np.empty_like(scene.tri_pos), np.empty_like(scene.tri_nrm)
    oracle.call("oracle_TriangleUpdateWrapper", model, normal, T, pos, nrm)
    sph = np.empty((K, 6), np.float32)
    oracle.call("oracle_BoundingSphereUpdateWrapper", model, scale, K, sph)
    return dataclasses.replace(scene, tri_pos=pos, tri_nrm=nrm, object_spheres=sph)


def test_moved_objects_match_oracle(engine, oracle, rsh):
    # two updates in a row, each applied to the scene as loaded
    scene, cam = c0(rsh, 96)
    cfg = dataclasses.replace(rsh.config.config0(96, 96), bounces=2, bounce_shadows=1)
    engine.load_scene(scene)
    for seed in (1, 2):
        model, normal, scale = _object_transforms(scene, seed)
        engine.update_objects(model, normal, scale)
        engine.render(cam, cfg, flags=engine.KEEP_PASSES)
        ofr = oracle.render(_moved_scene(oracle, scene, model, normal, scale), cam, cfg, threads=0)
        print(seed, compare_frames(engine, ofr, cfg, strict=True), file=sys.stderr)


def test_minimal_object_spheres_match_oracle(engine, oracle, rsh):
    # scene with the smallest enclosing balls as object spheres (rsh_object_spheres)
    scene = rsh.scenes.config1_scene(boxes=300)
    rg = scene.object_spheres[:, 4:6].astype(np.uint32)
    tight = dataclasses.replace(scene, object_spheres=rsh.object_spheres(scene.tri_pos, rg))
    cam = rsh.scenes.config1_camera(128, 128)
    run_case(engine, oracle, rsh, tight, cam, rsh.config.config1(128, 128))


def test_pixel_buffers_grow_after_a_two_light_scene(rsh):
    # A two-light scene leaves slot capacity (pixels x lights) for a later
    # frame with more pixels: the per-pixel buffers must still grow.
    e = rsh.Engine(0)
    office = rsh.scenes.config1_scene(boxes=50)
    assert office.light_total == 2
    e.load_scene(office)
    e.render(rsh.scenes.config1_camera(128, 128), rsh.config.config1(128, 128))
    scene, cam = rsh.scenes.config0_scene(), rsh.scenes.config0_camera(181, 181)   # 32761 pixels <= 32768 slots
    cfg = rsh.config.config0(181, 181)
    e.load_scene(scene)
    e.render(cam, cfg)
    got = e.pixels().copy()
    e.close()
    fresh = rsh.Engine(0)
    fresh.load_scene(scene)
    fresh.render(cam, cfg)
    np.testing.assert_array_equal(got, fresh.pixels())
    fresh.close()


def test_obj_scene_matches_oracle(engine, oracle, rsh, tmp_path):
    # the 512 scene written as an OBJ (per-vertex normals, one material per
    # object) and read back through rsh_amd.objfile
    from rsh_amd import objfile
    scene, cam = c0(rsh, 96)
    v = scene.tri_pos.reshape(-1, 3, 4)[:, :, :3]
    n = scene.tri_nrm.reshape(-1, 3, 4)[:, :, :3]
    lines, mtl = ["mtllib s.mtl"], []
    nv = 0
    for k, (a, b) in enumerate(scene.object_spheres[:, 4:6].astype(int)):
        lines.append(f"o obj{k}")
        lines.append(f"usemtl m{k}")
        kd = scene.mat_diffuse[scene.tri_mat[3 * a]]
        ks = scene.mat_specular[scene.tri_mat[3 * a]]
        mtl += [f"newmtl m{k}", "Kd %r %r %r" % tuple(map(float, kd[:3])), "Ks %r %r %r" % tuple(map(float, ks[:3])),
                "Ns %r" % float(ks[3])]
        for t in range(a, b + 1):
            for j in range(3):
                lines.append("v %r %r %r" % tuple(map(float, v[t, j])))
                lines.append("vn %r %r %r" % tuple(map(float, n[t, j])))
            lines.append("f %d//%d %d//%d %d//%d" % (nv + 1, nv + 1, nv + 2, nv + 2, nv + 3, nv + 3))
            nv += 3
    (tmp_path / "s.obj").write_text("\n".join(lines) + "\n")
    (tmp_path / "s.mtl").write_text("\n".join(mtl) + "\n")
    lights = [(tuple(map(float, scene.light_pos[i, :3])), tuple(map(float, scene.light_col[i, :3])),
               tuple(map(float, scene.light_int[i]))) for i in range(scene.light_total)]
    s2 = objfile.load_obj(str(tmp_path / "s.obj"), lights=lights,
                          reflective=lambda name, kd, ks, ns: scene.mat_diffuse[scene.tri_mat[
                              3 * int(scene.object_spheres[int(name[1:]), 4])], 3] > 0)
    assert s2.triangle_total == scene.triangle_total and s2.object_total == scene.object_total
    assert np.array_equal(s2.tri_pos[:, :3], scene.tri_pos[:, :3])
    run_case(engine, oracle, rsh, s2, cam, rsh.config.config0(96, 96))


@pytest.mark.parametrize("case", ["c0", "c0_ref_keys", "c1_bounce_shadows"])
def test_soft_shadows_match_oracle(engine, oracle, rsh, case):
    # soft shadows: 16 rays per shadow slot, shading by the grid weights of
    # the open cells; one light, reference keys, and two lights with shadow
    # passes at every bounce
    import dataclasses
    if case == "c1_bounce_shadows":
        scene, cam = rsh.scenes.config1_scene(), rsh.scenes.config1_camera(80, 64)
        cfg = rsh.config.config1(80, 64)
    else:
        scene, cam = c0(rsh, 96)
        cfg = rsh.config.config0(96, 96, key_mode=0 if case == "c0_ref_keys" else 1)
    rep = run_case(engine, oracle, rsh, scene, cam, dataclasses.replace(cfg, soft_shadows=1))
    st = engine.stats
    assert all(st.pass_slots[i] == 16 * cfg.width * cfg.height * scene.light_total
               for i in range(st.pass_count) if st.pass_kind[i] == 1)
    assert rep["pixels_exact"]


@pytest.mark.parametrize("key_mode,gate", [(0, 0), (1, 1)])
def test_cornell_layout_matches_oracle(engine, oracle, rsh, tmp_path, key_mode, gate):
    # the Cornell box assembled from per-object transforms (rsh_amd.objfile)
    # over stand-in model files: reference keys with every covered pixel
    # reflecting (the reference's rule), then octahedral keys with the right
    # wall and the sphere as the only mirrors
    import dataclasses
    import cornell_models
    from rsh_amd import objfile
    cornell_models.write_cornell_models(str(tmp_path))
    scene = objfile.cornell_scene(str(tmp_path), reflective=lambda k: k in (3, 6))
    cfg = dataclasses.replace(rsh.config.config1(128, 96), reflect_gate=gate, key_mode=key_mode)
    run_case(engine, oracle, rsh, scene, objfile.cornell_camera(128, 96), cfg)


@pytest.mark.parametrize("which,variant", [("c0", dict()), ("c0", dict(key_mode=0)), ("c0", dict(algorithm=1)),
                                           ("c0", dict(soft_shadows=1)), ("atrium", dict())])
def test_pass_keys_order_packets_nodes(engine, oracle, rsh, which, variant):
    """Per pass, the native engine's hash keys (every live slot), the stable
    sort permutation (slot of each sorted ray), the packets (chunk first
    positions and keys) and the hierarchy nodes, against the oracle bit for
    bit: the 512 scene at 128 x 128 and a 192 x 108 slice of the
    262k-triangle atrium frame (3 bounces with shadows)."""
    if which == "c0":
        scene, cam = c0(rsh, 128)
        cfg = dataclasses.replace(rsh.config.config0(128, 128), **variant)
    else:
        scene, cam = rsh.scenes.config2_scene(), rsh.scenes.config2_camera(192, 108)
        cfg = dataclasses.replace(rsh.config.config2(192, 108), **variant)
    cfg = cfg.with_scene_bounds(scene)
    engine.load_scene(scene)
    engine.render(cam, cfg, flags=engine.KEEP_ORDER)
    ofr = oracle.render(scene, cam, cfg, keep_order=True, keep_hits=False)
    assert engine.stats.pass_count == len(ofr.passes)
    checked_keys = checked_chunks = 0
    for i, op in enumerate(ofr.passes):
        assert engine.stats.pass_rays[i] == op["rays"] and engine.stats.pass_chunks[i] == op["chunks"], i
        if op["rays"] == 0:
            continue
        g = engine.pass_order(i)
        np.testing.assert_array_equal(g["order"], op["order"], err_msg=f"sort permutation, pass {i}")
        nn = g["nodes"].shape[0]
        np.testing.assert_array_equal(g["nodes"].view(np.uint32), op["nodes"][:nn].view(np.uint32),
                                      err_msg=f"hierarchy nodes, pass {i}")
        if "keys" in g:
            np.testing.assert_array_equal(g["flags"], op["flags"], err_msg=f"slot flags, pass {i}")
            live = op["flags"] == 0
            np.testing.assert_array_equal(g["keys"][live], op["keys"][live], err_msg=f"keys, pass {i}")
            checked_keys += 1
        if "bases" in g:
            np.testing.assert_array_equal(g["bases"], op["bases"], err_msg=f"packet bases, pass {i}")
            np.testing.assert_array_equal(g["ckeys"], op["ckeys"], err_msg=f"packet keys, pass {i}")
            checked_chunks += 1
    assert checked_keys >= len(ofr.passes) - 1
    if cfg.algorithm == 0:
        assert checked_chunks >= len(ofr.passes) - 1
