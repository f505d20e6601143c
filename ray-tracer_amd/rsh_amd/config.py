"""Frame configuration shared by the engine API and its callers.

Fields mirror ``rsh_frame_config`` in include/rsh.h.  Defaults reproduce the
reference's compile-time constants (projects/CUDA-Engine/source/Constants.h:9-29):
depth 2, subdivision 8 (split here into ``leaf`` and ``branch``), triangle
batch 5000.  The octahedral key holds 2 type + 3 x 5 cell + 2 x 6 direction
bits with the cell and direction bits interleaved (layout 2), which groups
incoherent bounce rays into tighter hierarchy nodes than the concatenated
layouts.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .scenes import Scene

CRSH, RAH = 0, 1
KEY_REF, KEY_OCT = 0, 1
HIT_EXACT, HIT_REF = 0, 1
CONE_REF, CONE_GUARDED = 0, 1


@dataclasses.dataclass
class FrameConfig:
    width: int = 512
    height: int = 512
    algorithm: int = CRSH
    key_mode: int = KEY_OCT
    hit_mode: int = HIT_EXACT
    cone_mode: int = CONE_GUARDED
    leaf: int = 8
    branch: int = 8
    depth: int = 2
    triangle_batch: int = 5000
    bounces: int = 1
    reflect_gate: int = 1
    bounce_shadows: int = 0
    primary_rays: int = 1
    cell_bits: int = 5
    dir_bits: int = 6
    key_layout: int = 2             # 0 cell|dir, 1 dir|cell, 2 bits interleaved (default: best grouping)
    scene_min: tuple = (0.0, 0.0, 0.0)
    cell_scale: tuple = (1.0, 1.0, 1.0)
    segment_shift: int = 0
    initial_radius: float = 0.0
    initial_spread: float = 0.0
    traversal_kernel: int = 0       # engine-only: 0 automatic, 1 generic kernel
    traversal_splits: int = 0       # engine-only, generic kernel: waves per top node, 0 = fill the GPU
    traversal_chunk: int = 0        # engine-only, 64-ray kernel: candidate clusters per work item, 0 = per pass
    traversal_prefilter: int = 0    # engine-only, 64-ray kernel: 0 auto, 1 no pre-test, 2 triangle, 3 + cluster masks
    soft_shadows: int = 0           # 1: 16 shadow rays per slot from a 4x4 grid about the light

    def with_scene_bounds(self, scene: Scene) -> "FrameConfig":
        """Cell grid over the scene AABB: 2^cell_bits cells per axis."""
        lo = scene.bounds_min.astype(np.float32)
        ext = (scene.bounds_max - scene.bounds_min).astype(np.float32)
        scale = np.float32(1 << self.cell_bits) / ext
        return dataclasses.replace(self, scene_min=tuple(float(v) for v in lo),
                                   cell_scale=tuple(float(np.float32(v)) for v in scale))

    def hash_bits(self) -> int:
        return 2 + 3 * self.cell_bits + 2 * self.dir_bits


def config0(width=512, height=512, **kw) -> FrameConfig:
    """512x512, 1 shadow ray + 1 mirror bounce per hit, OCT 3x5 + 2x6 + 2 key."""
    return FrameConfig(width=width, height=height, bounces=1, bounce_shadows=0, **kw)


def config1(width=1024, height=1024, **kw) -> FrameConfig:
    """1024x1024, shadow rays per light at every hit, 2 gated mirror bounces."""
    return FrameConfig(width=width, height=height, bounces=2, bounce_shadows=1, **kw)


def config2(width=1920, height=1080, **kw) -> FrameConfig:
    """1920x1080 atrium, 3 gated bounces plus shadow rays."""
    return FrameConfig(width=width, height=height, bounces=3, bounce_shadows=1, **kw)
