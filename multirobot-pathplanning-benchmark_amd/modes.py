"""Mode -> scene graph of manipulation problems: movable boxes held by robots or resting in the world.

The reference's manipulation environments change the scene when a task completes
(rai_base_env.py:704-839 ``set_to_mode``; pinocchio_env.py:287-372
``get_scenegraph_info_for_mode``): walking the chain of modes from the start mode, at each mode
switch the robots of the completed task are set to the next mode's entry configuration and, unless
the task is a "goto", its object ``frames[1]`` is attached to its new parent ``frames[0]`` at its
current world pose, so the pose relative to the new parent is frozen,
T_rel = T_parent(q_entry)^-1 T_world; a "remove" side effect deletes the object.

``ManipulationSpec`` holds the movable boxes and the tasks; ``scene_graph`` replays a mode's chain
into {object: (parent, T_rel)}; ``device_slots`` turns a scene graph into the engine's per-mode box
slots (slot a < n_arms: the box held by arm a, pose T_ee(q) T_rel; slot n_arms + j: object j resting
at a world pose; ``mrcc_scene_set_modes``).  The end-effector frame of a robot is the frame of its
last link.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class MovableBox:
    """A movable object: a box of half extents ``half`` starting at world pose ``pose`` (4 x 4)."""
    name: str
    half: np.ndarray
    pose: np.ndarray


@dataclass
class Task:
    """The manipulation fields of a reference Task (planning_env.py:54-100): the robots that do it,
    its type ("goto", None: no change of the scene; any other type, e.g. "pick" or "place":
    attach ``frames[1]`` to ``frames[0]``) and an optional side effect ("remove")."""
    name: str
    robots: List[str]
    type: Optional[str] = None
    frames: Optional[List[str]] = None
    side_effect: Optional[str] = None


@dataclass
class ManipulationSpec:
    objects: List[MovableBox]
    tasks: List[Task]
    root: str = "table"
    # gripper frame name -> robot name (default: "<robot>" and "<robot>_ee" name the end-effector)
    grippers: Dict[str, str] = field(default_factory=dict)


# ---------------------------------------------------------------------------------------------
def _rz(q):
    c, s = math.cos(q), math.sin(q)
    T = np.eye(4)
    T[0, 0], T[0, 1], T[1, 0], T[1, 1] = c, -s, s, c
    return T


def _from12(v):
    T = np.eye(4)
    T[:3, :3] = np.asarray(v[:9], dtype=np.float64).reshape(3, 3)
    T[:3, 3] = v[9:12]
    return T


def _from_rows12(v):
    T = np.eye(4)
    T[:3, :] = np.asarray(v, dtype=np.float64).reshape(3, 4)
    return T


def to12(T: np.ndarray) -> np.ndarray:
    """4 x 4 -> [R row-major | t] (the engine's box-transform layout)."""
    return np.concatenate([T[:3, :3].reshape(-1), T[:3, 3]])


def end_effector(desc, arm: int, q_arm: np.ndarray) -> np.ndarray:
    """World frame of the last link of `arm` for its joint values q_arm (binary64 products of 4 x 4
    transforms: DH joints Rz(q) Tz(d) Tx(a) Rx(alpha), chain joints Rz(q) C or, prismatic, Tz(q) C)."""
    T = _from12(desc.base[arm])
    jt = getattr(desc, "joint_type", None)
    for k in range(desc.dof):
        if getattr(desc, "joint_fixed", None) is not None:
            if jt is not None and jt[arm * desc.dof + k] == 1:  # prismatic: Tz(q) C
                J = np.eye(4)
                J[2, 3] = q_arm[k]
            else:
                J = _rz(q_arm[k])
            T = T @ J @ _from_rows12(desc.joint_fixed[arm, k])
        else:
            d, a, al = desc.dh[arm, k]
            ca, sa = math.cos(al), math.sin(al)
            A = np.array([[1.0, 0, 0, a], [0, ca, -sa, 0], [0, sa, ca, d], [0, 0, 0, 1]])
            T = T @ _rz(q_arm[k]) @ A
    return T


def _inv(T):
    R, p = T[:3, :3], T[:3, 3]
    out = np.eye(4)
    out[:3, :3] = R.T
    out[:3, 3] = -R.T @ p
    return out


class SceneGraphs:
    """Scene graphs of the modes of one environment (robots: names in configuration order)."""

    def __init__(self, spec: ManipulationSpec, desc, robots: Sequence[str], start_q: np.ndarray):
        self.spec = spec
        self.desc = desc
        self.robots = list(robots)
        self.start_q = np.asarray(start_q, dtype=np.float64)
        names = [o.name for o in spec.objects]
        if len(set(names)) != len(names):
            raise ValueError("movable objects need distinct names")
        self.obj_index = {n: j for j, n in enumerate(names)}

    def robot_of_frame(self, frame: str) -> Optional[int]:
        """Arm index of a gripper frame, None for the world (root or any other frame)."""
        if frame in self.spec.grippers:
            return self.robots.index(self.spec.grippers[frame])
        for i, r in enumerate(self.robots):
            if frame in (r, f"{r}_ee"):
                return i
        return None

    def _arm_q(self, q: np.ndarray, arm: int) -> np.ndarray:
        dof = self.desc.dof
        return q[arm * dof:(arm + 1) * dof]

    def scene_graph(self, mode) -> Dict[str, Tuple[Optional[int], np.ndarray]]:
        """{object: (holding arm or None, T_rel)} of a mode, replaying its chain of modes from the
        start mode (rai_base_env.py:726-812): at each switch the completed task's robots take the
        next mode's entry configuration, and a task that is not a "goto" attaches its object to
        its new parent at the object's current world pose; "remove" deletes the object."""
        chain = []
        m = mode
        while m is not None:
            chain.append(m)
            m = getattr(m, "prev_mode", None)
        chain.reverse()
        sg: Dict[str, Tuple[Optional[int], np.ndarray]] = {o.name: (None, np.array(o.pose, dtype=np.float64))
                                                            for o in self.spec.objects}
        q = self.start_q.copy()  # joint state of every robot (its last entry configuration)
        for cur, nxt in zip(chain[:-1], chain[1:]):
            changed = {t for t, u in zip(cur.task_ids, nxt.task_ids) if t != u}
            if len(changed) != 1:
                raise ValueError(f"mode switch {cur.task_ids} -> {nxt.task_ids}: exactly one task must complete")
            task = self.spec.tasks[changed.pop()]
            entry = np.asarray(nxt.entry_configuration.state() if hasattr(nxt.entry_configuration, "state")
                               else nxt.entry_configuration, dtype=np.float64)
            for r in task.robots:
                i = self.robots.index(r)
                dof = self.desc.dof
                q[i * dof:(i + 1) * dof] = entry[i * dof:(i + 1) * dof]
            if task.type is None or task.type == "goto":
                continue
            parent, obj = task.frames[0], task.frames[1]
            if obj not in sg:
                raise ValueError(f"task {task.name}: object {obj} is not in the scene")
            holder, rel = sg[obj]
            world = rel if holder is None else end_effector(self.desc, holder, self._arm_q(q, holder)) @ rel
            new = self.robot_of_frame(parent)
            sg[obj] = (None, world) if new is None else (new, _inv(end_effector(self.desc, new, self._arm_q(q, new))) @ world)
            if task.side_effect == "remove":
                del sg[obj]
        return sg

    def key(self, sg) -> tuple:
        """Hashable form of a scene graph (exact transforms)."""
        return tuple(sorted((n, h, to12(T).tobytes()) for n, (h, T) in sg.items()))

    def device_slots(self, sg):
        """(has [slots], rel [slots, 12], half [slots, 3]) with slots = n_arms + n_objects."""
        na, no = self.desc.n_arms, len(self.spec.objects)
        has = np.zeros(na + no, dtype=np.uint8)
        rel = np.zeros((na + no, 12))
        half = np.zeros((na + no, 3))
        for name, (holder, T) in sg.items():
            j = self.obj_index[name]
            h = np.asarray(self.spec.objects[j].half, dtype=np.float64)
            slot = na + j if holder is None else holder
            if holder is not None and has[slot]:
                raise ValueError(f"arm {holder} holds two objects")
            has[slot] = 1
            rel[slot] = to12(T)
            half[slot] = h
        return has, rel, half
