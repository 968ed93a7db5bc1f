"""Registered environments on the GPU backend.

* ``abstract.test``, ``abstract.center_rect_2d``, ``abstract.center_rect_10d`` — the reference's
  abstract environments (abstract_env.py:381-491) with the same scenes, radii, start poses,
  resolution 0.01 and tolerance 0.01.
* ``mrcc.planar_disks`` (C1), ``mrcc.planar_boxes`` (C2), ``mrcc.quad_arms`` (C3),
  ``mrcc.mixed_arms`` (C5) — the benchmark scenes of this engine (seed 0).
"""
from __future__ import annotations

import numpy as np

from . import scenes
from .backend import MrccEnvironment
from .configuration import NpConfiguration
from .planning_env import register


@register("abstract.test")
class AbstractTestEnv(MrccEnvironment):
    """abstract_env.py:381-421: two disks r=0.1, Sphere((0,0),0.2), Rectangle((0,0.4),(0.5,0.5))."""

    def __init__(self):
        super().__init__(scenes.abstract_test_scene(), robots=["a1", "a2"],
                         start_pos=NpConfiguration.from_list([[-0.8, 0], [0.8, 0]]))


@register([("abstract.center_rect_2d", {"n": 2}), ("abstract.center_rect_10d", {"n": 10})])
class AbstractCenterRectEnv(MrccEnvironment):
    """abstract_env.py:443-491: two disks r=0.1 in n dimensions around a centred box of size 0.5."""

    def __init__(self, n=2):
        r1 = np.zeros(n)
        r1[0] = -0.8
        r2 = np.zeros(n)
        r2[0] = 0.8
        super().__init__(scenes.center_rect_scene(n), robots=["a1", "a2"],
                         start_pos=NpConfiguration.from_list([r1, r2]))


@register("mrcc.planar_disks")
class PlanarDisksEnv(MrccEnvironment):
    def __init__(self, seed: int = 0):
        super().__init__(scenes.planar_disks(seed), collision_resolution=0.01)


@register("mrcc.planar_boxes")
class PlanarBoxesEnv(MrccEnvironment):
    def __init__(self, seed: int = 0):
        super().__init__(scenes.planar_boxes(seed), collision_resolution=0.01)


@register("mrcc.quad_arms")
class QuadArmsEnv(MrccEnvironment):
    def __init__(self, seed: int = 0):
        super().__init__(scenes.quad_arms(seed), collision_resolution=0.05)


@register("mrcc.mixed_arms")
class MixedArmsEnv(MrccEnvironment):
    def __init__(self, seed: int = 0, n_modes: int = 64):
        super().__init__(scenes.mixed_arms(seed, n_modes=n_modes), collision_resolution=0.05)
