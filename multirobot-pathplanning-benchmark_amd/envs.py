"""Registered environments on the GPU backend.

* ``abstract.test``, ``abstract.center_rect_2d``, ``abstract.center_rect_10d`` — the reference's
  abstract environments (abstract_env.py:381-491) with the same scenes, radii, start poses,
  resolution 0.01 and tolerance 0.01.
* ``mrcc.planar_disks`` (C1), ``mrcc.planar_boxes`` (C2), ``mrcc.quad_arms`` (C3),
  ``mrcc.mixed_arms`` (C5) — the benchmark scenes of this engine (seed 0).
* ``mujoco.four_ur10`` — the reference's four-UR10e mujoco world (mujoco_env.py:1160-1204):
  robots, start pose, resolution 0.01 and tolerance 0.01, on the spherized model that
  ``mjcf.load_mjcf_arms`` builds from the world's MJCF file.  The file is not part of this
  package: pass its path, or set ``MRCC_MJCF_FOUR_UR10``.
* ``mujoco.swap`` — mujoco_env.py:892-933: two cylinders on slide x, slide y and hinge yaw joints
  (``mj_two_dim.xml``; path or ``MRCC_MJCF_TWO_DIM``), start (0, -1, 0), (0, 1, 0), resolution
  0.01, tolerance 0.
* ``pinocchio.middle_obs`` — pinocchio_env.py:664-707: the planar URDF world
  ``middle_obstacle_two_agents.urdf`` (path or ``MRCC_URDF_MIDDLE_OBS``), prismatic x, y and
  revolute phi per robot, start (0, -0.8, 0), (0, 0.8, 0), resolution 0.01, tolerance 0.
* ``pinocchio.random_ur5`` — pinocchio_env.py:1025-1276: two UR5e arms on the primitives URDF
  ``ur5e_constrained_coll_primitives.urdf`` (path or ``MRCC_URDF_UR5E``), a start pose sampled
  collision free (seeded), resolution 0.01, tolerance 0 (the security margins are set while the
  tolerance is still 0, pinocchio_env.py:99-134).
"""
from __future__ import annotations

import os

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


@register("mujoco.four_ur10")
class FourUR10Env(MrccEnvironment):
    """mujoco_env.py:1160-1204: four UR10e arms (6 joints each) around a box on a table."""

    def __init__(self, path: str = None, spheres_per_link: int = 8):
        from .mjcf import load_mjcf_arms

        path = path or os.environ.get("MRCC_MJCF_FOUR_UR10")
        if not path or not os.path.exists(path):
            raise FileNotFoundError("mujoco.four_ur10 needs the MJCF world mujoco_4_ur10_world_closer.xml "
                                    "(argument `path` or MRCC_MJCF_FOUR_UR10)")
        desc = load_mjcf_arms(path, tolerance=0.01, spheres_per_link=spheres_per_link)
        home = np.array([0, -2, 1.0, -1.0, -1.57, 1.0])
        start = NpConfiguration.from_list([home for _ in range(4)])
        super().__init__(desc, robots=["ur10_1", "ur10_2", "ur10_3", "ur10_4"], start_pos=start,
                         collision_resolution=0.01, collision_tolerance=0.01)


def _asset(path, env_var, what):
    path = path or os.environ.get(env_var)
    if not path or not os.path.exists(path):
        raise FileNotFoundError(f"needs the reference asset {what} (argument `path` or {env_var})")
    return path


@register("mujoco.swap")
class SwapEnv(MrccEnvironment):
    """mujoco_env.py:892-933: two planar cylinder robots swap places around a centre block."""

    def __init__(self, path: str = None):
        from .mjcf import load_mjcf_arms

        desc = load_mjcf_arms(_asset(path, "MRCC_MJCF_TWO_DIM", "mj_two_dim.xml"), tolerance=0.0)
        start = NpConfiguration.from_list([np.array([0.0, -1.0, 0.0]), np.array([0.0, 1.0, 0.0])])
        super().__init__(desc, robots=["a1", "a2"], start_pos=start, collision_resolution=0.01,
                         collision_tolerance=0.0)


@register("pinocchio.middle_obs")
class MiddleObsEnv(MrccEnvironment):
    """pinocchio_env.py:664-707: two planar robots on either side of a middle obstacle."""

    def __init__(self, path: str = None):
        from .urdf import load_urdf_arms

        desc = load_urdf_arms(_asset(path, "MRCC_URDF_MIDDLE_OBS", "middle_obstacle_two_agents.urdf"))
        start = NpConfiguration.from_list([np.array([0.0, -0.8, 0.0]), np.array([0.0, 0.8, 0.0])])
        super().__init__(desc, robots=["a1", "a2"], start_pos=start, collision_resolution=0.01,
                         collision_tolerance=0.0)


@register("pinocchio.random_ur5")
class DualUR5Env(MrccEnvironment):
    """pinocchio_env.py:1189-1276: two UR5e arms over a table; the start pose is a collision-free
    configuration drawn uniformly in the limits (pinocchio_env.py:1191-1196), seeded."""

    def __init__(self, path: str = None, seed: int = 0):
        from .urdf import load_dual_ur5

        desc = load_dual_ur5(_asset(path, "MRCC_URDF_UR5E", "ur5e_constrained_coll_primitives.urdf"))
        q0 = NpConfiguration.from_list([np.zeros(6), np.zeros(6)])
        super().__init__(desc, robots=["a1", "a2"], start_pos=q0, collision_resolution=0.01,
                         collision_tolerance=0.0)
        rng = np.random.default_rng(seed)
        for _ in range(64):
            Q = rng.uniform(desc.limits[0], desc.limits[1], size=(256, desc.dim))
            ok = self.batch_is_collision_free(Q)
            if ok.any():
                self.start_pos = q0.from_flat(Q[int(np.argmax(ok))])
                break
        else:
            raise RuntimeError("no collision-free start pose among 16384 samples")
