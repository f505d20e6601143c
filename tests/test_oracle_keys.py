"""Hash keys of the oracle: the reference's spherical bins and the octahedral key.

Reference expressions: projects/CUDA-Engine/source/RayTracer.cu:147-168
(CreateShadowRayIndex) and :170-198 (CreateReflectionRayIndex).  They are
evaluated here independently with numpy float32 arithmetic and numpy's own
arctan2/arccos; bins may only differ where acos lands within a few ulp of a
bin edge.
"""
import ctypes as C

import numpy as np
import pytest

HALF_PI = np.float32(np.float32(3.1415927) / np.float32(2.0))


def sat_u32(x):
    x = np.asarray(x, np.float32)
    out = np.where(np.isnan(x) | (x <= 0), 0.0, np.minimum(np.trunc(x.astype(np.float64)), 4294967295.0))
    return out.astype(np.uint64)


def ref_shadow_key(d, l):
    az = sat_u32((np.arctan2(d[:, 0], d[:, 0]).astype(np.float32) + HALF_PI) * np.float32(5.0))
    pol = sat_u32(np.arccos(d[:, 2]).astype(np.float32) * np.float32(5.0))
    k = (np.uint64(l + 1) << np.uint64(30)) + (np.minimum(az, 31) << np.uint64(15)) + np.minimum(pol, 31) + 1
    return (k & 0xFFFFFFFF).astype(np.uint32), np.arccos(d[:, 2].astype(np.float64)) * 5.0


def ref_reflection_key(o, d):
    dist = np.sqrt(((o[:, 0] * o[:, 0] + o[:, 1] * o[:, 1]) + o[:, 2] * o[:, 2]).astype(np.float32))
    k = np.minimum(sat_u32((np.arctan2(o[:, 0], o[:, 0]).astype(np.float32) + HALF_PI) * np.float32(20.0)), 128)
    with np.errstate(invalid="ignore", divide="ignore"):
        k = (k << np.uint64(7)) | np.minimum(sat_u32(np.arccos((o[:, 2] / dist).astype(np.float32)).astype(np.float32)
                                                     * np.float32(20.0)), 128)
    k = (k << np.uint64(7)) | np.minimum(sat_u32((np.arctan2(d[:, 0], d[:, 0]).astype(np.float32) + HALF_PI)
                                                 * np.float32(5.0)), 30)
    k = (k << np.uint64(5)) | np.minimum(sat_u32(np.arccos(d[:, 2]).astype(np.float32) * np.float32(5.0)), 30)
    return ((k + 1) & 0xFFFFFFFF).astype(np.uint32)


def unit_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return d.astype(np.float32)


def keys_of(fn, a, b=None, extra=None):
    out = []
    for i in range(a.shape[0]):
        pa = np.ascontiguousarray(a[i], np.float32)
        args = [pa.ctypes.data_as(C.POINTER(C.c_float))]
        if b is not None:
            pb = np.ascontiguousarray(b[i], np.float32)
            args.append(pb.ctypes.data_as(C.POINTER(C.c_float)))
        if extra is not None:
            args.append(extra)
        out.append(fn(*args))
    return np.array(out, np.uint32)


def test_atan2_diagonal_bins(oracle):
    """atan2(v, v) takes four values, so the azimuth bin is 11 (v>0), 7 (+0) or 0."""
    L = oracle.lib()
    o = np.zeros((1, 3), np.float32)
    for v, want in [(0.3, 11), (1e-30, 11), (0.0, 7), (-0.0, 0), (-0.2, 0)]:
        d = np.array([[v, 0.6, 0.8]], np.float32)
        k = keys_of(L.oracle_shadow_key_ref, o, d, 0)[0]
        assert (k - 1) >> 15 & 31 == want, (v, k)
        want_np = int(sat_u32((np.arctan2(np.float32(v), np.float32(v)) + HALF_PI) * np.float32(5.0)))
        assert want == want_np
    assert keys_of(L.oracle_shadow_key_ref, o, np.array([[0.0, 0.0, 1.0]], np.float32), 0)[0] == \
        (1 << 30) + (7 << 15) + 0 + 1
    # light index in the top two bits
    assert keys_of(L.oracle_shadow_key_ref, o, np.array([[0.0, 0.0, -1.0]], np.float32), 1)[0] >> 30 == 2


def test_shadow_keys_match_numpy(oracle):
    rng = np.random.default_rng(5)
    d = unit_dirs(rng, 20000)
    d[:50, 0] = 0.0
    d[50:100, 0] = -0.0
    d[100:110, 2] = 1.0
    d[110:120, 2] = -1.0
    o = rng.uniform(50, 150, size=d.shape).astype(np.float32)
    got = keys_of(oracle.lib().oracle_shadow_key_ref, o, d, 0)
    want, pol = ref_shadow_key(d, 0)
    diff = got != want
    # any disagreement must sit on a polar bin edge (acos * 5 within 1e-5 of an integer)
    near_edge = np.abs(pol - np.round(pol)) < 1e-5
    assert np.all(near_edge[diff]), np.nonzero(diff & ~near_edge)[0][:10]
    assert diff.mean() < 1e-3


def test_reflection_keys_match_numpy(oracle):
    rng = np.random.default_rng(6)
    d = unit_dirs(rng, 20000)
    o = rng.uniform(-150, 150, size=d.shape).astype(np.float32)
    o[:20] = 0.0                     # origin at 0: acos(NaN) bin -> 0
    o[20:40, 0] = -0.0
    got = keys_of(oracle.lib().oracle_reflection_key_ref, o, d)
    want = ref_reflection_key(o, d)
    assert (got != want).mean() < 2e-3
    assert got[0] == want[0]


def oct_config(oracle, cell_bits=5, dir_bits=6, layout=0, lo=(50.0, 50.0, 50.0), ext=100.0):
    c = oracle.OConfig()
    c.cell_bits, c.dir_bits, c.key_layout = cell_bits, dir_bits, layout
    c.scene_min[:] = list(lo)
    c.cell_scale[:] = [float(np.float32(np.float32(1 << cell_bits) / np.float32(ext)))] * 3
    return c


def oct_decode(u, v, db):
    n = float(1 << db)
    px = (u + 0.5) / n * 2 - 1
    py = (v + 0.5) / n * 2 - 1
    pz = 1 - np.abs(px) - np.abs(py)
    neg = pz < 0
    qx = np.where(neg, (1 - np.abs(py)) * np.sign(px), px)
    qy = np.where(neg, (1 - np.abs(px)) * np.sign(py), py)
    d = np.stack([qx, qy, pz], 1)
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("layout", [0, 1])
def test_oct_key_layout_and_decode(oracle, layout):
    """Field layout is [type:2][cell z|y|x][u|v] (or [type][u|v][cell]); the
    decoded cell contains the origin and the decoded direction lies within
    the octahedral bin of the direction."""
    L = oracle.lib()
    cfg = oct_config(oracle, layout=layout)
    rng = np.random.default_rng(7)
    n = 4000
    d = unit_dirs(rng, n)
    o = rng.uniform(50, 150, size=(n, 3)).astype(np.float32)
    for t in (0, 1, 2, 3):
        got = keys_of(lambda a, b, tt: L.oracle_key_oct(C.byref(cfg), a, b, tt), o, d, t)
        assert np.all(got >> 27 == t)
        if layout == 0:
            cell, dirb = (got >> 12) & 0x7FFF, got & 0xFFF
        else:
            dirb, cell = (got >> 15) & 0xFFF, got & 0x7FFF
        cx, cy, cz = cell & 31, (cell >> 5) & 31, (cell >> 10) & 31
        ci = np.stack([cx, cy, cz], 1).astype(np.float64)
        lo = 50 + ci * (100 / 32)
        assert np.all(o >= lo - 1e-3) and np.all(o <= lo + 100 / 32 + 1e-3)
        u, v = (dirb >> 6) & 63, dirb & 63
        dec = oct_decode(u, v, 6)
        ang = np.degrees(np.arccos(np.clip((dec * d).sum(1), -1, 1)))
        assert ang.max() < 6.0, ang.max()   # octahedral bins of 64x64 span at most a few degrees


def test_oct_key_edges(oracle):
    L = oracle.lib()
    cfg = oct_config(oracle)
    o = np.array([[0.0, 0.0, 0.0], [1e9, 1e9, 1e9], [100, 100, 100]], np.float32)   # saturating cells
    d = np.array([[0, 0, 1], [0, 0, -1], [0, 0, 0]], np.float32)                    # poles, zero vector
    got = keys_of(lambda a, b, tt: L.oracle_key_oct(C.byref(cfg), a, b, tt), o, d, 1)
    assert (got[0] >> 12) & 0x7FFF == 0
    assert (got[1] >> 12) & 0x7FFF == 0x7FFF
    assert got[0] & 0xFFF == (32 << 6) | 32          # +z maps to the centre
    assert got[2] & 0xFFF == 0                       # NaN direction -> bin 0


def test_oct_interleaved_layout_is_a_bit_shuffle_of_layout0(oracle, rsh):
    """Layout 2 interleaves the z, y, x cell bits and u, v direction bits of
    layout 0's fields, most significant level first, under the same type."""
    import dataclasses
    rng = np.random.default_rng(21)
    o = rng.uniform(-40, 160, size=(300, 3)).astype(np.float32)
    d = unit_dirs(rng, 300)
    for cb, db in ((4, 6), (5, 5), (3, 7), (6, 3)):
        base = dataclasses.replace(rsh.config.FrameConfig(), cell_bits=cb, dir_bits=db, scene_min=(0.0, 0.0, 0.0),
                                   cell_scale=(float(np.float32((1 << cb) / 120.0)),) * 3)
        L = oracle.lib()
        for t in range(4):
            c0 = oracle.make_config(dataclasses.replace(base, key_layout=0))
            c2 = oracle.make_config(dataclasses.replace(base, key_layout=2))
            k0 = keys_of(lambda a, b: L.oracle_key_oct(C.byref(c0), a, b, t), o, d)
            k2 = keys_of(lambda a, b: L.oracle_key_oct(C.byref(c2), a, b, t), o, d)
            nb = 3 * cb + 2 * db
            assert np.all((k0 >> nb) == t) and np.all((k2 >> nb) == t)
            v = (k0 & ((1 << db) - 1)).astype(np.int64)
            u = ((k0 >> db) & ((1 << db) - 1)).astype(np.int64)
            cell = (k0 >> (2 * db)).astype(np.int64)
            cx, cy, cz = cell & ((1 << cb) - 1), (cell >> cb) & ((1 << cb) - 1), cell >> (2 * cb)
            want = np.zeros(len(k0), np.int64)
            for i in range(max(cb, db) - 1, -1, -1):
                if i < cb:
                    want = (want << 3) | (((cz >> i) & 1) << 2) | (((cy >> i) & 1) << 1) | ((cx >> i) & 1)
                if i < db:
                    want = (want << 2) | (((u >> i) & 1) << 1) | ((v >> i) & 1)
            np.testing.assert_array_equal(k2 & ((1 << nb) - 1), want)
