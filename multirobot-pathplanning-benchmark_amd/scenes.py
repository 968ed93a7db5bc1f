"""Scene descriptions and seeded workload generators.

A ``SceneDesc`` is the plain-array description that both the engine (``mrcc_scene_create_*``) and
any CPU checker consume.  Generators follow the benchmark configurations of this engine:

* C1 ``planar_disks``   — the reference's abstract environment (abstract_env.py:41-109, 255-276):
  disk robots + axis-aligned rectangles, 2 robots x 2 DoF, 20 rectangles.
* C2 ``planar_boxes``   — 4 oriented rectangular robots (x, y, theta), 100 oriented boxes.
* C3 ``quad_arms``      — 4 x 7-DoF DH arms around a 1.6 m table, link spheres, 50 boxes
  (the shape of the reference's ``vampmr.quad_panda`` problem, mr_vamp_env.py:810-876).
* C5 ``mixed_arms``     — 8 arm slots; instances use 2..8 arms and may hold boxes at the
  end-effector (modes), cf. grasp attachment in rai_base_env.py:801-805.

Random draws use ``numpy.random.default_rng(seed)`` in the order scene -> configs -> edges.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class SceneDesc:
    kind: str                                   # "disks" | "planar" | "arms"
    limits: np.ndarray                          # [2, D] joint limits
    # disks
    robot_dim: int = 0
    robot_radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    sphere_center: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    sphere_radius: np.ndarray = field(default_factory=lambda: np.zeros(0))
    rect_center: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))
    rect_size: np.ndarray = field(default_factory=lambda: np.zeros((0, 0)))  # full edge length
    # planar boxes
    robot_half: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    obstacle_pose: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    obstacle_half: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    # arms
    n_arms: int = 0
    dof: int = 0
    base: np.ndarray = field(default_factory=lambda: np.zeros((0, 12)))
    dh: np.ndarray = field(default_factory=lambda: np.zeros((0, 0, 3)))
    link_sphere_count: np.ndarray = field(default_factory=lambda: np.zeros(0, dtype=np.int32))
    sphere_local: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    boxes: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    modes: Optional[dict] = None               # n_active, has_grasp, grasp_rel, grasp_half
    # arms of any joint geometry: [n_arms, dof, 12] fixed transform C after Rz(q) of each joint
    # (rows R00 R01 R02 t0, R10 R11 R12 t1, R20 R21 R22 t2); when set, dh is unused
    joint_fixed: Optional[np.ndarray] = None
    # chain arms: [n_arms * dof] joint types, 0 revolute (Rz(q) C), 1 prismatic (Tz(q) C); None: all
    # revolute
    joint_type: Optional[np.ndarray] = None

    @property
    def dim(self) -> int:
        return int(self.limits.shape[1])

    def joint_travel(self) -> np.ndarray:
        """[n_arms * dof] travel of each prismatic joint, max(|lower|, |upper|) of its limits (0 for
        revolute joints): a configuration beyond it counts as a collision."""
        jt = np.zeros(self.dim, dtype=np.uint8) if self.joint_type is None else np.asarray(self.joint_type)
        return np.where(jt == 1, np.maximum(np.abs(self.limits[0]), np.abs(self.limits[1])), 0.0)

    @property
    def n_robots(self) -> int:
        if self.kind == "disks":
            return len(self.robot_radius)
        if self.kind == "planar":
            return len(self.robot_half)
        return self.n_arms

    def robot_slices(self):
        if self.kind == "disks":
            d = self.robot_dim
        elif self.kind == "planar":
            d = 3
        else:
            d = self.dof
        return [(i * d, (i + 1) * d) for i in range(self.n_robots)]


# ---------------------------------------------------------------------------------------------
# disks (the reference's abstract environments)

def disks_scene(robot_radius, robot_dim, limits, spheres=(), rects=()) -> SceneDesc:
    """spheres: [(center, r)], rects: [(center, full_size)] as in abstract_env.Sphere/Rectangle."""
    d = robot_dim
    sc = np.array([np.asarray(c, dtype=np.float64) for c, _ in spheres]).reshape(-1, d)
    sr = np.array([float(r) for _, r in spheres], dtype=np.float64)
    rc = np.array([np.asarray(c, dtype=np.float64) for c, _ in rects]).reshape(-1, d)
    rs = np.array([np.asarray(b, dtype=np.float64) for _, b in rects]).reshape(-1, d)
    return SceneDesc(kind="disks", limits=np.asarray(limits, dtype=np.float64), robot_dim=d,
                     robot_radius=np.asarray(robot_radius, dtype=np.float64),
                     sphere_center=sc, sphere_radius=sr, rect_center=rc, rect_size=rs)


def abstract_test_scene() -> SceneDesc:
    """abstract.test: make_middle_obstacle_n_dim_env (abstract_env.py:363-378), radii 0.1 (:390)."""
    lim = np.ones((2, 4)) * 2
    lim[0, :] = -2
    return disks_scene([0.1, 0.1], 2, lim, spheres=[(np.zeros(2), 0.2)],
                       rects=[(np.array([0, 0.4]), np.array([0.5, 0.5]))])


def center_rect_scene(n: int) -> SceneDesc:
    """abstract.center_rect_{2d,10d}: make_center_rectangle_nd (abstract_env.py:428-440)."""
    lim = np.ones((2, 2 * n)) * 2
    lim[0, :] = -2
    return disks_scene([0.1, 0.1], n, lim, rects=[(np.zeros(n), np.ones(n) * 0.5)])


def planar_disks(seed: int = 0, n_robots: int = 2, n_rects: int = 20, radius: float = 0.1,
                 rng=None) -> SceneDesc:
    """C1: 2 disks r=0.1 in [-2,2]^2, 20 rectangles (centres U[-2,2]^2, half-extents U[0.05,0.3]^2)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    centers = rng.uniform(-2, 2, size=(n_rects, 2))
    half = rng.uniform(0.05, 0.3, size=(n_rects, 2))
    lim = np.ones((2, 2 * n_robots)) * 2
    lim[0, :] = -2
    return disks_scene([radius] * n_robots, 2, lim,
                       rects=[(centers[i], 2 * half[i]) for i in range(n_rects)])


# ---------------------------------------------------------------------------------------------
# planar oriented boxes

def planar_boxes(seed: int = 0, n_robots: int = 4, n_obstacles: int = 100, rng=None) -> SceneDesc:
    """C2: robots (x, y, theta) with half-extents U[0.05,0.2]^2; obstacles with centre U[-2,2]^2,
    yaw U[-pi,pi], half-extents U[0.05,0.3]^2."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    rh = rng.uniform(0.05, 0.2, size=(n_robots, 2))
    pose = np.concatenate([rng.uniform(-2, 2, size=(n_obstacles, 2)),
                           rng.uniform(-math.pi, math.pi, size=(n_obstacles, 1))], axis=1)
    oh = rng.uniform(0.05, 0.3, size=(n_obstacles, 2))
    lo = np.tile([-2.0, -2.0, -math.pi], n_robots)
    hi = np.tile([2.0, 2.0, math.pi], n_robots)
    return SceneDesc(kind="planar", limits=np.stack([lo, hi]), robot_half=rh, obstacle_pose=pose,
                     obstacle_half=oh)


# ---------------------------------------------------------------------------------------------
# arms

ARM_SLOTS = [(0.8, 0.8), (-0.8, 0.8), (-0.8, -0.8), (0.8, -0.8),
             (0.8, 0.0), (0.0, 0.8), (-0.8, 0.0), (0.0, -0.8)]
JOINT_LIMIT = 2.9


def _arm_geometry(rng, n_arms: int, dof: int):
    """DH chains (alpha alternating -/+ pi/2, link length L_k U[0.1,0.4] along d for even joints and
    along a for odd joints) and link spheres: 4..8 spheres of radius U[0.03,0.08] evenly spaced on
    the middle 70% of the segment from the previous joint origin to the link frame origin (fixed in
    the link frame), so links two apart do not overlap at every configuration."""
    base = np.zeros((n_arms, 12))
    dh = np.zeros((n_arms, dof, 3))
    counts = []
    spheres = []
    for a in range(n_arms):
        x, y = ARM_SLOTS[a]
        yaw = math.atan2(-y, -x)
        c, s = math.cos(yaw), math.sin(yaw)
        base[a] = [c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0, x, y, 0.0]
        L = rng.uniform(0.1, 0.4, size=dof)
        for k in range(dof):
            alpha = 0.0 if k == dof - 1 else (-math.pi / 2 if k % 2 == 0 else math.pi / 2)
            d = L[k] if k % 2 == 0 else 0.0
            aa = L[k] if k % 2 == 1 else 0.0
            dh[a, k] = (d, aa, alpha)
            # previous joint origin in the link frame: (-a, -d sin(alpha), -d cos(alpha))
            prev = np.array([-aa, -d * math.sin(alpha), -d * math.cos(alpha)])
            n = int(rng.integers(4, 9))
            counts.append(n)
            r = rng.uniform(0.03, 0.08, size=n)
            for j in range(n):
                t = 0.15 + 0.7 * (j + 0.5) / n
                spheres.append([prev[0] * t, prev[1] * t, prev[2] * t, r[j]])
    return base, dh, np.array(counts, dtype=np.int32), np.array(spheres, dtype=np.float64)


def _table_boxes(rng, n_boxes: int, n_arms: int):
    """Boxes resting on the table (z = half height), half-extents U[0.03,0.12]^3, centres U on the
    table at least 0.35 m (xy) from every arm base."""
    half = rng.uniform(0.03, 0.12, size=(n_boxes, 3))
    bases = np.array(ARM_SLOTS[:n_arms])
    xy = np.zeros((n_boxes, 2))
    for i in range(n_boxes):
        while True:
            p = rng.uniform(-0.8, 0.8, size=2)
            if np.min(np.linalg.norm(bases - p, axis=1)) >= 0.35:
                break
        xy[i] = p
    centers = np.concatenate([xy, half[:, 2:3]], axis=1)
    return np.concatenate([centers, half], axis=1)


def quad_arms(seed: int = 0, n_arms: int = 4, dof: int = 7, n_boxes: int = 50, rng=None) -> SceneDesc:
    """C3: 4 x 7-DoF arms at the corners of a 1.6 m x 1.6 m table, 50 boxes on the table."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    base, dh, counts, sph = _arm_geometry(rng, n_arms, dof)
    boxes = _table_boxes(rng, n_boxes, n_arms)
    lim = np.stack([np.full(n_arms * dof, -JOINT_LIMIT), np.full(n_arms * dof, JOINT_LIMIT)])
    return SceneDesc(kind="arms", limits=lim, n_arms=n_arms, dof=dof, base=base, dh=dh,
                     link_sphere_count=counts, sphere_local=sph, boxes=boxes)


def grasp_modes(rng, n_modes: int, n_arms: int, min_arms: int = 2, p_grasp: float = 0.5,
                first_default: bool = True):
    """Modes for mixed batches: n_active U{min_arms..n_arms}; each active arm holds a box with
    half-extents U[0.02,0.06]^3 with probability p_grasp, attached at the end-effector with a
    random yaw and offset h_z + 0.02 along the end-effector z axis (frozen at grasp time)."""
    n_active = rng.integers(min_arms, n_arms + 1, size=n_modes).astype(np.int32)
    has = (rng.uniform(size=(n_modes, n_arms)) < p_grasp).astype(np.uint8)
    half = rng.uniform(0.02, 0.06, size=(n_modes, n_arms, 3))
    yaw = rng.uniform(-math.pi, math.pi, size=(n_modes, n_arms))
    rel = np.zeros((n_modes, n_arms, 12))
    rel[..., 0] = np.cos(yaw)
    rel[..., 1] = -np.sin(yaw)
    rel[..., 3] = np.sin(yaw)
    rel[..., 4] = np.cos(yaw)
    rel[..., 8] = 1.0
    rel[..., 11] = half[..., 2] + 0.02
    for m in range(n_modes):
        has[m, n_active[m]:] = 0
    if first_default:
        n_active[0] = n_arms
        has[0] = 0
    return {"n_active": n_active, "has_grasp": has, "grasp_rel": rel, "grasp_half": half}


def mixed_arms(seed: int = 0, n_modes: int = 1024, n_arms: int = 8, dof: int = 7,
               n_boxes: int = 50, rng=None) -> SceneDesc:
    """C5: 8 arm slots (4 corners + 4 edge midpoints), modes = (arm count 2..8, grasped boxes)."""
    rng = rng if rng is not None else np.random.default_rng(seed)
    sc = quad_arms(n_arms=n_arms, dof=dof, n_boxes=n_boxes, rng=rng)
    sc.modes = grasp_modes(rng, n_modes, n_arms)
    return sc


# ---------------------------------------------------------------------------------------------
# inputs

def uniform_configs(rng, desc: SceneDesc, n: int) -> np.ndarray:
    """Uniform in the joint limits (planning_env.py:1697-1708)."""
    return rng.uniform(desc.limits[0], desc.limits[1], size=(n, desc.dim))


def random_pair_edges(rng, desc: SceneDesc, n: int):
    """q1, q2 iid uniform (C1)."""
    return uniform_configs(rng, desc, n), uniform_configs(rng, desc, n)


def lognormal_edges(rng, desc: SceneDesc, n: int, resolution: float, median: float = 0.3,
                    sigma: float = 1.0, max_states: int = 500, dims=None):
    """q1 uniform; length L ~ lognormal(ln median, sigma) in the max-abs metric; direction d with
    max|d| = 1; q2 = q1 + L d; resampled until q2 is in bounds and N <= max_states."""
    lo, hi = desc.limits
    D = desc.dim
    q1 = np.zeros((n, D))
    q2 = np.zeros((n, D))
    filled = 0
    while filled < n:
        m = (n - filled) * 2 + 16
        a = rng.uniform(lo, hi, size=(m, D))
        L = rng.lognormal(math.log(median), sigma, size=m)
        d = rng.uniform(-1, 1, size=(m, D))
        if dims is not None:
            mask = np.zeros(D, dtype=bool)
            mask[dims] = True
            d[:, ~mask] = 0.0
        d /= np.max(np.abs(d), axis=1, keepdims=True)
        b = a + L[:, None] * d
        N = np.floor(np.max(np.abs(b - a), axis=1) / resolution) + 1
        ok = np.all((b >= lo) & (b <= hi), axis=1) & (N <= max_states)
        a, b = a[ok], b[ok]
        k = min(len(a), n - filled)
        q1[filled:filled + k] = a[:k]
        q2[filled:filled + k] = b[:k]
        filled += k
    return q1, q2


def mixed_instances(rng, desc: SceneDesc, n_instances: int, resolution: float, max_edges: int = 64,
                    sigma_home: float = 0.4, median: float = 0.3, sigma: float = 1.0):
    """C5 batch: instance i uses mode i (its arm count and held boxes) and is a path of U{1..64}
    edges: the first waypoint near home, each next one a lognormal step (max-abs metric over the
    instance's active joints).  Returns vertices [V, D], edge row indices idx1 / idx2 into them,
    the mode of every edge, and the instance of every edge."""
    D, dof = desc.dim, desc.dof
    n_active = desc.modes["n_active"][:n_instances]
    K = rng.integers(1, max_edges + 1, size=n_instances)
    V = np.zeros((int(np.sum(K + 1)), D))
    idx1 = np.zeros(int(np.sum(K)), dtype=np.int32)
    inst = np.repeat(np.arange(n_instances, dtype=np.int32), K)
    v0 = np.concatenate([[0], np.cumsum(K + 1)[:-1]])
    e0 = np.concatenate([[0], np.cumsum(K)[:-1]])
    for i in range(n_instances):
        da = int(n_active[i]) * dof
        w = np.zeros((K[i] + 1, D))
        w[0, :da] = rng.normal(0.0, sigma_home, da)
        L = rng.lognormal(math.log(median), sigma, size=K[i])
        d = rng.uniform(-1, 1, size=(K[i], da))
        d /= np.max(np.abs(d), axis=1, keepdims=True)
        w[1:, :da] = w[0, :da] + np.cumsum(L[:, None] * d, axis=0)
        w = np.clip(w, desc.limits[0], desc.limits[1])
        V[v0[i]:v0[i] + K[i] + 1] = w
        idx1[e0[i]:e0[i] + K[i]] = np.arange(v0[i], v0[i] + K[i], dtype=np.int32)
    return V, idx1, idx1 + 1, inst.copy(), inst
