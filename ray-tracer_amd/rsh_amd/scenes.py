"""Seeded synthetic scenes for the five workloads.

A scene is plain numpy arrays in the engine's upload layout
(projects/CUDA-Engine/source/main.cpp:2375-2470 and :2596-2656):

* ``tri_pos``   float32 [3T, 4]  per-vertex position (w = 1)
* ``tri_nrm``   float32 [3T, 4]  per-vertex normal (w = 0)
* ``tri_mat``   int32   [3T]     material id per vertex (one per object)
* ``mat_diffuse``  float32 [M, 4]  (Kd.rgb, reflectivity: 1 mirror, 0 not)
* ``mat_specular`` float32 [M, 4]  (Ks.rgb, Ns)
* ``object_spheres`` float32 [K, 6]  center | (radius, firstTri, lastTri);
  objects tile the triangle range in order
* ``light_pos`` [L, 4], ``light_col`` [L, 4], ``light_int`` [L, 2]

Geometry is generated in unit coordinates and mapped to world space by
``WORLD_SCALE * unit + WORLD_OFFSET``: the reference's absolute constants
(ray offset 0.01, RayTracer.cu:34; degenerate-triangle threshold 0.01,
RayTracer.cu:314; attenuation coefficients RayTracer.cu:1667) assume scenes
about a hundred units across, and the offset keeps every surface away from
the world origin, which the reference reads as "no fragment" (RayTracer.cu:708).

Material colours are the values of the reference's .mtl files
(projects/CUDA-Engine/models/{Cube,gold,silver,ruby,emerald,saphire}.mtl).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

WORLD_SCALE = 100.0
WORLD_OFFSET = 50.0
SEED_BASE = 0xC0FFEE

# (Kd, Ks, Ns) from projects/CUDA-Engine/models/*.mtl
PALETTE = {
    "cube": ((0.784314, 0.784314, 0.784314), (0.0, 0.0, 0.0), 153.0),
    "gold": ((0.75164, 0.60648, 0.22648), (0.628281, 0.555802, 0.366065), 102.0),
    "silver": ((0.50754, 0.50754, 0.50754), (0.508273, 0.508273, 0.508273), 102.0),
    "ruby": ((0.61424, 0.04136, 0.04136), (0.727811, 0.626959, 0.626959), 153.0),
    "emerald": ((0.07568, 0.61424, 0.07568), (0.633, 0.727811, 0.633), 153.0),
    "saphire": ((0.0, 0.50980392, 0.50196078), (0.50196078, 0.50196078, 0.50196078), 63.75),
}
PALETTE_NAMES = list(PALETTE)


@dataclasses.dataclass
class Camera:
    eye: np.ndarray
    direction: np.ndarray
    right: np.ndarray
    up: np.ndarray


@dataclasses.dataclass
class Scene:
    name: str
    tri_pos: np.ndarray
    tri_nrm: np.ndarray
    tri_mat: np.ndarray
    mat_diffuse: np.ndarray
    mat_specular: np.ndarray
    object_spheres: np.ndarray
    light_pos: np.ndarray
    light_col: np.ndarray
    light_int: np.ndarray
    bounds_min: np.ndarray
    bounds_max: np.ndarray

    @property
    def triangle_total(self) -> int:
        return self.tri_pos.shape[0] // 3

    @property
    def object_total(self) -> int:
        return self.object_spheres.shape[0]

    @property
    def light_total(self) -> int:
        return self.light_pos.shape[0]

    def triangle_object(self) -> np.ndarray:
        """Object index of every triangle (objects tile the range in order)."""
        out = np.empty(self.triangle_total, np.int32)
        for k, row in enumerate(self.object_spheres):
            out[int(row[4]):int(row[5]) + 1] = k
        return out


def camera_basis(eye, target, up, fov_deg: float, width: int, height: int) -> Camera:
    """Pinhole basis of Camera::loadView (projects/CUDA-Engine/source/Camera.cpp:108-133).

    ``direction`` has the eye-target distance as length, ``right`` and ``up``
    span the full image plane at that distance (the reference names them
    half-width/half-height but multiplies tan(fov/2) by 2, Camera.cpp:114-115).
    """
    eye = np.asarray(eye, np.float64)
    target = np.asarray(target, np.float64)
    up = np.asarray(up, np.float64)
    distance = np.linalg.norm(target - eye)
    theta = math.radians(fov_deg * 0.5)
    half_height = 2.0 * distance * math.tan(theta)
    half_width = half_height * (width / height)
    direction = (target - eye) / distance * distance
    right = np.cross(direction, up)
    right = right / np.linalg.norm(right) * half_width
    upv = np.cross(right, direction)
    upv = upv / np.linalg.norm(upv) * half_height
    f = lambda v: np.asarray(v, np.float32)
    return Camera(f(eye), f(direction), f(right), f(upv))


def world(p) -> np.ndarray:
    return np.asarray(p, np.float64) * WORLD_SCALE + WORLD_OFFSET


class _Builder:
    """Accumulates objects (triangle soup + per-vertex normals + one material)."""

    def __init__(self):
        self.pos, self.nrm, self.mat, self.spheres = [], [], [], []
        self.materials_d, self.materials_s = [], []
        self.count = 0

    def material(self, name: str, reflective: bool) -> int:
        kd, ks, ns = PALETTE[name]
        self.materials_d.append((*kd, 1.0 if reflective else 0.0))
        self.materials_s.append((*ks, ns))
        return len(self.materials_d) - 1

    def add(self, tris: np.ndarray, nrms: np.ndarray, material: int):
        """tris, nrms: float64 [n, 3, 3] in unit coordinates."""
        n = tris.shape[0]
        wp = world(tris.reshape(-1, 3))
        self.pos.append(wp)
        nn = nrms.reshape(-1, 3)
        nn = nn / np.linalg.norm(nn, axis=1, keepdims=True)
        self.nrm.append(nn)
        self.mat.append(np.full(n * 3, material, np.int32))
        lo, hi = wp.min(0), wp.max(0)
        center = 0.5 * (lo + hi)
        radius = float(np.sqrt(((wp - center) ** 2).sum(1)).max()) * (1.0 + 1e-5) + 1e-4
        self.spheres.append((*center, radius, self.count, self.count + n - 1))
        self.count += n

    def build(self, name, lights, bounds_min, bounds_max) -> Scene:
        pos = np.concatenate(self.pos).astype(np.float32)
        nrm = np.concatenate(self.nrm).astype(np.float32)
        one = np.ones((pos.shape[0], 1), np.float32)
        zero = np.zeros((pos.shape[0], 1), np.float32)
        lp = np.array([(*world(l[0]), 1.0) for l in lights], np.float32)
        lc = np.array([(*l[1], 1.0) for l in lights], np.float32)
        li = np.array([l[2] for l in lights], np.float32)
        return Scene(
            name=name,
            tri_pos=np.ascontiguousarray(np.hstack([pos, one])),
            tri_nrm=np.ascontiguousarray(np.hstack([nrm, zero])),
            tri_mat=np.concatenate(self.mat).astype(np.int32),
            mat_diffuse=np.array(self.materials_d, np.float32),
            mat_specular=np.array(self.materials_s, np.float32),
            object_spheres=np.array(self.spheres, np.float32),
            light_pos=lp, light_col=lc, light_int=li,
            bounds_min=world(bounds_min).astype(np.float32),
            bounds_max=world(bounds_max).astype(np.float32),
        )


# ------------------------------------------------------------------ shapes --
def _quad(a, b, c, d):
    return [(a, b, c), (a, c, d)]


def box_triangles(lo, hi, inward=False):
    """12 triangles of an axis-aligned box with flat normals."""
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = lambda x, y, z: np.array([x, y, z])
    faces = [
        ((v(x0, y0, z0), v(x0, y1, z0), v(x0, y1, z1), v(x0, y0, z1)), (-1, 0, 0)),
        ((v(x1, y0, z0), v(x1, y0, z1), v(x1, y1, z1), v(x1, y1, z0)), (1, 0, 0)),
        ((v(x0, y0, z0), v(x0, y0, z1), v(x1, y0, z1), v(x1, y0, z0)), (0, -1, 0)),
        ((v(x0, y1, z0), v(x1, y1, z0), v(x1, y1, z1), v(x0, y1, z1)), (0, 1, 0)),
        ((v(x0, y0, z0), v(x1, y0, z0), v(x1, y1, z0), v(x0, y1, z0)), (0, 0, -1)),
        ((v(x0, y0, z1), v(x0, y1, z1), v(x1, y1, z1), v(x1, y0, z1)), (0, 0, 1)),
    ]
    tris, nrms = [], []
    for quad, n in faces:
        n = np.array(n, np.float64) * (-1.0 if inward else 1.0)
        for t in _quad(*quad):
            tris.append(t)
            nrms.append((n, n, n))
    return np.array(tris), np.array(nrms)


def uv_sphere(center, radius, slices=32, stacks=17):
    """UV sphere: 2*slices*(stacks-2) + 2*slices triangles (1024 for 32 x 17)."""
    c = np.asarray(center, np.float64)
    theta = np.linspace(0.0, math.pi, stacks + 1)
    phi = np.linspace(0.0, 2.0 * math.pi, slices + 1)

    def p(i, j):
        n = np.array([math.sin(theta[i]) * math.cos(phi[j]), math.cos(theta[i]),
                      math.sin(theta[i]) * math.sin(phi[j])])
        return c + radius * n, n

    tris, nrms = [], []
    for i in range(stacks):
        for j in range(slices):
            (a, na), (b, nb) = p(i, j), p(i + 1, j)
            (cc, nc), (d, nd) = p(i + 1, j + 1), p(i, j + 1)
            if i != 0:
                tris.append((a, b, d)); nrms.append((na, nb, nd))
            if i != stacks - 1:
                tris.append((b, cc, d)); nrms.append((nb, nc, nd))
    return np.array(tris), np.array(nrms)


def cylinder(base, radius, height, slices, stacks):
    """Open vertical cylinder (side wall only): 2*slices*stacks triangles."""
    b = np.asarray(base, np.float64)
    tris, nrms = [], []
    for k in range(stacks):
        y0, y1 = height * k / stacks, height * (k + 1) / stacks
        for j in range(slices):
            a0, a1 = 2 * math.pi * j / slices, 2 * math.pi * (j + 1) / slices
            n0 = np.array([math.cos(a0), 0.0, math.sin(a0)])
            n1 = np.array([math.cos(a1), 0.0, math.sin(a1)])
            p00, p01 = b + radius * n0 + [0, y0, 0], b + radius * n1 + [0, y0, 0]
            p10, p11 = b + radius * n0 + [0, y1, 0], b + radius * n1 + [0, y1, 0]
            tris.append((p00, p10, p11)); nrms.append((n0, n0, n1))
            tris.append((p00, p11, p01)); nrms.append((n0, n1, n1))
    return np.array(tris), np.array(nrms)


def arch(p0, p1, height, tube, around, along):
    """Semicircular tube from p0 to p1 (both at the same height), 2*around*along tris."""
    p0, p1 = np.asarray(p0, np.float64), np.asarray(p1, np.float64)
    mid = 0.5 * (p0 + p1)
    axis = p1 - p0
    span = np.linalg.norm(axis)
    ax = axis / span
    upv = np.array([0.0, 1.0, 0.0])
    side = np.cross(ax, upv)
    r_big = 0.5 * span

    def frame(s):
        a = math.pi * s
        centre = mid - ax * r_big * math.cos(a) + upv * height * math.sin(a)
        tangent = ax * r_big * math.sin(a) + upv * height * math.cos(a)
        tangent /= np.linalg.norm(tangent)
        normal = np.cross(side, tangent)
        return centre, normal, side

    tris, nrms = [], []
    for i in range(along):
        c0, n0, s0 = frame(i / along)
        c1, n1, s1 = frame((i + 1) / along)
        for j in range(around):
            b0, b1 = 2 * math.pi * j / around, 2 * math.pi * (j + 1) / around
            d00 = n0 * math.cos(b0) + s0 * math.sin(b0)
            d01 = n0 * math.cos(b1) + s0 * math.sin(b1)
            d10 = n1 * math.cos(b0) + s1 * math.sin(b0)
            d11 = n1 * math.cos(b1) + s1 * math.sin(b1)
            q00, q01 = c0 + tube * d00, c0 + tube * d01
            q10, q11 = c1 + tube * d10, c1 + tube * d11
            tris.append((q00, q10, q11)); nrms.append((d00, d10, d11))
            tris.append((q00, q11, q01)); nrms.append((d00, d11, d01))
    return np.array(tris), np.array(nrms)


def floor_tile(x0, z0, size, sub, y=0.0):
    tris, nrms = [], []
    n = np.array([0.0, 1.0, 0.0])
    step = size / sub
    for i in range(sub):
        for j in range(sub):
            a = np.array([x0 + i * step, y, z0 + j * step])
            b = a + [step, 0, 0]
            c = a + [step, 0, step]
            d = a + [0, 0, step]
            tris.append((a, d, c)); nrms.append((n, n, n))
            tris.append((a, c, b)); nrms.append((n, n, n))
    return np.array(tris), np.array(nrms)


# ------------------------------------------------------------------ configs --
def _white_light(pos, di=0.75, si=0.75):
    return (pos, (1.0, 1.0, 1.0), (di, si))


def config0_scene(seed: int = SEED_BASE + 0, sphere_slices=32, sphere_stacks=17) -> Scene:
    """512x512 plain case: 12-triangle room + 2 UV spheres of 1024 triangles."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    room_m = b.material("cube", True)
    b.add(*box_triangles((0, 0, 0), (1, 1, 1), inward=True), room_m)
    for name in ("gold", "silver"):
        r = rng.uniform(0.1, 0.3)
        c = rng.uniform(r, 1.0 - r, size=3)
        c[2] = max(c[2], 0.35)              # keep the spheres in front of the camera
        b.add(*uv_sphere(c, r, sphere_slices, sphere_stacks), b.material(name, True))
    return b.build("config0", [_white_light((0.5, 0.9, 0.5))], (0, 0, 0), (1, 1, 1))


def config0_camera(width=512, height=512) -> Camera:
    return camera_basis(world((0.5, 0.5, 0.05)), world((0.5, 0.5, 1.0)), (0, 1, 0), 60.0, width, height)


def config1_scene(seed: int = SEED_BASE + 1, boxes: int = 2500, reflective=0.4) -> Scene:
    """Office scale: room + seeded axis-aligned boxes, 2 point lights."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.add(*box_triangles((0, 0, 0), (1, 1, 1), inward=True), b.material("cube", False))
    eye = np.array([0.5, 0.5, 0.02])
    lights = [np.array([0.3, 0.95, 0.5]), np.array([0.7, 0.95, 0.6])]
    placed = 0
    while placed < boxes:
        side = np.exp(rng.uniform(math.log(0.02), math.log(0.3), size=3)) * np.array([1.0, 1.0, 1.0])
        side = side * 0.35                   # keeps the room from being filled solid
        lo = rng.uniform(0.0, 1.0 - side)
        hi = lo + side
        centre = 0.5 * (lo + hi)
        if np.linalg.norm(centre - eye) < 0.25 or any(np.linalg.norm(centre - l) < 0.08 for l in lights):
            continue
        name = PALETTE_NAMES[int(rng.integers(len(PALETTE_NAMES)))]
        b.add(*box_triangles(lo, hi), b.material(name, bool(rng.random() < reflective)))
        placed += 1
    return b.build("config1", [_white_light(l) for l in lights], (0, 0, 0), (1, 1, 1))


def config1_camera(width=1024, height=1024) -> Camera:
    return camera_basis(world((0.5, 0.5, 0.02)), world((0.5, 0.45, 1.0)), (0, 1, 0), 60.0, width, height)


def config2_scene(seed: int = SEED_BASE + 2, reflective=0.3, grid=6, col_slices=48, col_stacks=54,
                  tiles=32, tile_sub=4, arch_around=12, arch_along=30) -> Scene:
    """Atrium scale: tiled floor, a grid of tessellated columns joined by arches."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.add(*box_triangles((0, 0, 0), (1, 1, 1), inward=True), b.material("cube", False))

    def mat():
        name = PALETTE_NAMES[int(rng.integers(len(PALETTE_NAMES)))]
        return b.material(name, bool(rng.random() < reflective))

    size = 1.0 / tiles
    for i in range(tiles):
        for j in range(tiles):
            b.add(*floor_tile(i * size, j * size, size, tile_sub, y=0.002), mat())
    col_h, col_r = 0.62, 0.025
    xs = np.linspace(0.14, 0.86, grid)
    for x in xs:
        for z in xs:
            b.add(*cylinder((x, 0.0, z + 0.1), col_r, col_h, col_slices, col_stacks), mat())
    for i in range(grid):
        for j in range(grid - 1):
            b.add(*arch((xs[j], col_h, xs[i] + 0.1), (xs[j + 1], col_h, xs[i] + 0.1), 0.09, 0.012,
                        arch_around, arch_along), mat())
            b.add(*arch((xs[i], col_h, xs[j] + 0.1), (xs[i], col_h, xs[j + 1] + 0.1), 0.09, 0.012,
                        arch_around, arch_along), mat())
    return b.build("config2", [_white_light((0.5, 0.95, 0.55))], (0, 0, 0), (1, 1, 1))


def config2_camera(width=1920, height=1080) -> Camera:
    return camera_basis(world((0.5, 0.35, 0.01)), world((0.5, 0.3, 1.0)), (0, 1, 0), 60.0, width, height)


def soup_scene(seed: int = SEED_BASE + 4, triangles: int = 100_000, per_object: int = 250) -> Scene:
    """Random triangle soup (serving workload's fourth resident scene)."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.add(*box_triangles((0, 0, 0), (1, 1, 1), inward=True), b.material("cube", False))
    left = triangles - 12
    while left > 0:
        n = min(per_object, left)
        centre = rng.uniform(0.1, 0.9, size=3)
        v = centre + rng.normal(0.0, 0.03, size=(n, 3, 3))
        v = np.clip(v, 0.01, 0.99)
        nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
        nrm /= np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-12)
        nrms = np.repeat(nrm[:, None, :], 3, axis=1)
        name = PALETTE_NAMES[int(rng.integers(len(PALETTE_NAMES)))]
        b.add(v, nrms, b.material(name, bool(rng.random() < 0.3)))
        left -= n
    return b.build("soup", [_white_light((0.5, 0.95, 0.5))], (0, 0, 0), (1, 1, 1))


def random_view(rng: np.random.Generator, width: int, height: int) -> Camera:
    """Random camera pose inside the unit room looking at a random interior point."""
    eye = rng.uniform(0.15, 0.85, size=3)
    target = rng.uniform(0.2, 0.8, size=3)
    while np.linalg.norm(target - eye) < 0.2:
        target = rng.uniform(0.2, 0.8, size=3)
    return camera_basis(world(eye), world(target), (0, 1, 0), 60.0, width, height)


def batch_views(n: int = 64, seed: int = SEED_BASE + 3, lo: int = 128, hi: int = 1024):
    """Batched-views workload: n square views of side uniform in [lo, hi]
    (E[pixels] = 398,677 for 128..1024) at random poses in the unit room.
    Returns [(camera, width, height)]."""
    rng = np.random.default_rng(seed)
    views = []
    for _ in range(n):
        side = int(rng.integers(lo, hi + 1))
        views.append((random_view(rng, side, side), side, side))
    return views
