"""Synthetic ray queries against a scene: camera, shadow and cosine-diffuse rays.

Rays are float32 [n, 6] (origin xyz, direction xyz, the reference's ray
layout).  Shadow queries go from a light to a random surface point, as the
frame's shadow rays do (RayTracer.cu:682-740), with tmax = |P - light| -
2 * 0.01; diffuse queries leave a random surface point in a cosine-weighted
direction about the triangle's face normal.  Surface points are drawn
uniformly over the triangles by area.
"""
from __future__ import annotations

import numpy as np

from . import scenes

EPS = np.float32(0.01)
PRIMARY, SHADOW, DIFFUSE = 0, 1, 2
KIND_OF = {PRIMARY: 0, SHADOW: 1, DIFFUSE: 0}   # pipeline kind: closest hit, any hit, closest hit


def _surface_points(scene, n, rng):
    v = scene.tri_pos.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]), axis=1)
    t = rng.choice(len(v), size=n, p=area / area.sum())
    r1, r2 = rng.random(n), rng.random(n)
    s = np.sqrt(r1)
    b0, b1, b2 = 1.0 - s, s * (1.0 - r2), s * r2
    P = b0[:, None] * v[t, 0] + b1[:, None] * v[t, 1] + b2[:, None] * v[t, 2]
    N = np.cross(v[t, 1] - v[t, 0], v[t, 2] - v[t, 0])
    N /= np.maximum(np.linalg.norm(N, axis=1, keepdims=True), 1e-30)
    return P, N


def primary_rays(scene, n, rng):
    cam = scenes.random_view(rng, 256, 256)
    a, c = rng.random(n) - 0.5, rng.random(n) - 0.5
    d = np.asarray(cam.direction) + np.outer(a, cam.right) + np.outer(c, cam.up)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(np.asarray(cam.eye, np.float64), d.shape)
    return np.concatenate([o, d], 1).astype(np.float32), None


def shadow_rays(scene, n, rng):
    P, _ = _surface_points(scene, n, rng)
    L = scene.light_pos[:, :3].astype(np.float64)[rng.integers(len(scene.light_pos), size=n)]
    d = P - L
    dist = np.linalg.norm(d, axis=1)
    d /= np.maximum(dist[:, None], 1e-30)
    rays = np.concatenate([L, d], 1).astype(np.float32)
    tmax = (np.linalg.norm(P.astype(np.float32) - L.astype(np.float32), axis=1) - 2 * EPS).astype(np.float32)
    return rays, tmax


def diffuse_rays(scene, n, rng):
    P, N = _surface_points(scene, n, rng)
    N *= np.where(rng.random(n) < 0.5, 1.0, -1.0)[:, None]        # either side of the surface
    # cosine-weighted hemisphere about N
    r1, r2 = rng.random(n), rng.random(n)
    phi, rr = 2 * np.pi * r1, np.sqrt(r2)
    local = np.stack([rr * np.cos(phi), rr * np.sin(phi), np.sqrt(1.0 - r2)], 1)
    a = np.where(np.abs(N[:, 0:1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    T = np.cross(N, a)
    T /= np.linalg.norm(T, axis=1, keepdims=True)
    B = np.cross(N, T)
    d = local[:, 0:1] * T + local[:, 1:2] * B + local[:, 2:3] * N
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([P, d], 1).astype(np.float32), None


GENERATORS = {PRIMARY: primary_rays, SHADOW: shadow_rays, DIFFUSE: diffuse_rays}


def make_query(scene, mix_kind, n, rng):
    """(rays [n, 6], tmax or None, pipeline kind) for a primary / shadow / diffuse query."""
    rays, tmax = GENERATORS[mix_kind](scene, n, rng)
    return rays, tmax, KIND_OF[mix_kind]
