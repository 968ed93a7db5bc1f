"""Writes tests/golden/mj_four_ur10.npz: the spherized four-UR10e scene that mjcf.load_mjcf_arms
builds from the reference's mujoco world, plus independently composed MJCF body frames, so that the
GPU tests can run that scene where the reference checkout is absent.

The MJCF files are read as text (XML) from the reference checkout; nothing of the reference is
imported or executed.  Run from the repository root:
    python tests/golden/make_mjcf_fixture.py \\
        /root/reference/src/multi_robot_multi_goal_planning/assets/models/mujoco/mujoco_4_ur10_world_closer.xml
"""
import importlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
mjcf = importlib.import_module("multirobot-pathplanning-benchmark_amd.mjcf")

TOLERANCE = 0.01  # the environment's collision_tolerance (mujoco_env.py:1201)


def main(path):
    d = mjcf.load_mjcf_arms(path, tolerance=TOLERANCE)
    rng = np.random.default_rng(2024)
    qs = rng.uniform(d.limits[0], d.limits[1], size=(8, d.n_arms * d.dof))
    frames = np.stack([np.stack(mjcf.body_frames(path, q)) for q in qs])  # [8, L, 4, 4]
    np.savez(os.path.join(HERE, "mj_four_ur10.npz"), n_arms=d.n_arms, dof=d.dof, base=d.base,
             joint_fixed=d.joint_fixed, link_sphere_count=d.link_sphere_count, sphere_local=d.sphere_local,
             boxes=d.boxes, limits=d.limits, tolerance=TOLERANCE, frame_q=qs, body_frames=frames)
    print("wrote", os.path.join(HERE, "mj_four_ur10.npz"), d.n_arms, "arms x", d.dof, "joints,",
          len(d.sphere_local), "spheres,", len(d.boxes), "boxes")


if __name__ == "__main__":
    main(sys.argv[1])
