"""Robot scenes from URDF files (the pinocchio backend's robot descriptions).

The pinocchio environments of the reference build their models with ``pin.buildModelsFromUrdf``
(pinocchio_env.py:664-670, 795-800, 882-887) or append copies of one robot URDF at base placements
(``make_dual_ur5_waypoint_env``, pinocchio_env.py:1025-1187), and count a configuration in collision
when any geometry pair of the collision model overlaps (``pin.computeCollisions`` with the security
margins set when the environment is built, pinocchio_env.py:131-134 and 474-478).  ``load_urdf_arms``
turns such a file into an arm scene of the engine (``SceneDesc`` with ``joint_fixed``):

* the link tree is read from ``<link>`` / ``<joint>`` elements; the links reached from the root
  through fixed joints are static; every movable joint (revolute, continuous, prismatic) whose
  parent is static starts a robot, a serial chain of movable joints (links joined by fixed joints
  are merged into the link before them).  URDF gives a child link's frame as
  T_child = T_parent O J(q), O the joint's origin (xyz, rpy = Rz(y) Ry(p) Rx(r)) and J the rotation
  about (revolute) or translation along (prismatic) the joint axis.  With U carrying z onto the
  axis, J = U Z(q) U^T (Z = Rz or Tz); pre_k = P_k O_k U_k (P_k: the fixed path from the previous
  robot link, or from the root, to the joint's parent) and post_k = U_k^T give the engine's chain
  F_k = F_{k-1} Z(q_k) C_k, C_k = post_k pre_{k+1} (C_last = post_last), base F_0 = placement pre_1;
* collision geometry (``<collision>``; meshes are not supported) of robot links becomes spheres in
  F_k: a cylinder of radius r and length 2h is covered by n spheres at the centres of n slabs of
  height s = 2h / n, radius sqrt(r^2 + (s/2)^2); a sphere is itself; a box is covered by the 8
  spheres around the centres of its octants; at most 8 spheres per link;
* collision geometry of static links becomes obstacle boxes (the axis-aligned box around it;
  exact for axis-aligned boxes), or is dropped (``static="drop"``);
* ``placements``: one copy of every robot of the file per base placement (pinocchio's appendModel);
  ``boxes``: extra obstacle boxes (centre, half extents) in the world;
* the contact tolerance is added as tolerance / 2 to every sphere radius and box half-extent.

The engine's pair set (links at least two apart within a robot, every link pair of two robots,
every link against every obstacle box) stands for the reference's collision pairs.  The geometry
is a sphere cover of the primitives pinocchio collides, so results are the engine's on this model.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .mjcf import MAX_SPHERES_PER_LINK, _allot, _chain12, _cylinder_spheres, _inv, _rows12, _T, _z_to
from .scenes import SceneDesc

MOVABLE = ("revolute", "continuous", "prismatic")


def _floats(s: Optional[str], n: int, default) -> np.ndarray:
    if s is None:
        return np.asarray(default, dtype=np.float64)
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    if len(v) != n:
        raise ValueError(f"expected {n} numbers, got {s!r}")
    return v


def rpy_matrix(r: float, p: float, y: float) -> np.ndarray:
    """URDF's fixed-axis roll-pitch-yaw: Rz(y) Ry(p) Rx(r)."""
    cr, sr, cp, sp, cy, sy = math.cos(r), math.sin(r), math.cos(p), math.sin(p), math.cos(y), math.sin(y)
    Rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    Ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    Rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def _origin(el: Optional[ET.Element]) -> np.ndarray:
    if el is None:
        return np.eye(4)
    xyz = _floats(el.get("xyz"), 3, [0, 0, 0])
    rpy = _floats(el.get("rpy"), 3, [0, 0, 0])
    return _T(rpy_matrix(*rpy), xyz)


class UrdfModel:
    """Links (collision geometry: (kind, pose in the link frame, size)) and joints of a URDF file.
    Sizes: box half extents [3]; cylinder (radius, half length); sphere (radius,)."""

    def __init__(self, path: str):
        root = ET.parse(path).getroot()
        self.links: Dict[str, List[Tuple[str, np.ndarray, np.ndarray]]] = {}
        for l in root.findall("link"):
            geoms = []
            for c in l.findall("collision"):
                g = c.find("geometry")
                if g is None or len(g) == 0:
                    continue
                shape = g[0]
                T = _origin(c.find("origin"))
                if shape.tag == "box":
                    geoms.append(("box", T, _floats(shape.get("size"), 3, [0, 0, 0]) / 2))
                elif shape.tag == "cylinder":
                    geoms.append(("cylinder", T, np.array([float(shape.get("radius")), float(shape.get("length")) / 2])))
                elif shape.tag == "sphere":
                    geoms.append(("sphere", T, np.array([float(shape.get("radius"))])))
                else:
                    raise ValueError(f"link {l.get('name')}: collision geometry <{shape.tag}> is not supported")
            self.links[l.get("name")] = geoms
        self.joints: List[dict] = []
        for j in root.findall("joint"):
            lim = j.find("limit")
            self.joints.append({
                "name": j.get("name"), "type": j.get("type"),
                "parent": j.find("parent").get("link"), "child": j.find("child").get("link"),
                "origin": _origin(j.find("origin")),
                "axis": _floats(j.find("axis").get("xyz") if j.find("axis") is not None else None, 3, [1, 0, 0]),
                "lower": None if lim is None or lim.get("lower") is None else float(lim.get("lower")),
                "upper": None if lim is None or lim.get("upper") is None else float(lim.get("upper")),
            })
        for j in self.joints:
            if j["type"] not in MOVABLE + ("fixed",):
                raise ValueError(f"joint {j['name']}: type {j['type']} is not supported")
        children = {j["child"] for j in self.joints}
        roots = [n for n in self.links if n not in children]
        if len(roots) != 1:
            raise ValueError(f"expected one root link, found {roots}")
        self.root = roots[0]
        self.by_parent: Dict[str, List[dict]] = {}
        for j in self.joints:
            self.by_parent.setdefault(j["parent"], []).append(j)

    def fixed_subtree(self, link: str) -> Dict[str, np.ndarray]:
        """Links reached from `link` through fixed joints, with their pose in `link`'s frame."""
        out = {link: np.eye(4)}
        todo = [link]
        while todo:
            l = todo.pop()
            for j in self.by_parent.get(l, []):
                if j["type"] == "fixed":
                    out[j["child"]] = out[l] @ j["origin"]
                    todo.append(j["child"])
        return out

    def movable_below(self, sub: Dict[str, np.ndarray]) -> List[dict]:
        """Movable joints whose parent is in `sub`, in file order."""
        return [j for j in self.joints if j["type"] in MOVABLE and j["parent"] in sub]


def _joint_motion(j: dict, q: float) -> np.ndarray:
    a = j["axis"] / np.linalg.norm(j["axis"])
    if j["type"] == "prismatic":
        return _T(None, a * q)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return _T(np.eye(3) + math.sin(q) * K + (1 - math.cos(q)) * (K @ K))


def _box_spheres(T: np.ndarray, half: np.ndarray) -> List[Tuple[np.ndarray, float]]:
    """8 spheres around the centres of a box's octants (each contains its octant)."""
    r = float(np.linalg.norm(half / 2))
    out = []
    for sx in (-1, 1):
        for sy in (-1, 1):
            for sz in (-1, 1):
                out.append(((T @ np.array([sx * half[0] / 2, sy * half[1] / 2, sz * half[2] / 2, 1.0]))[:3], r))
    return out


def _aabb(kind: str, T: np.ndarray, size: np.ndarray) -> np.ndarray:
    """Centre and half extents of the axis-aligned box around a primitive at pose T."""
    R, c = T[:3, :3], T[:3, 3]
    if kind == "box":
        return np.concatenate([c, np.abs(R) @ size])
    if kind == "sphere":
        return np.concatenate([c, [size[0]] * 3])
    a = R[:, 2]  # cylinder axis
    r, h = size
    return np.concatenate([c, h * np.abs(a) + r * np.sqrt(np.maximum(0.0, 1.0 - a * a))])


def _robot_chain(m: UrdfModel, first: dict, P0: np.ndarray, budget: int):
    """(joints, pre, post, spheres per link in F_k, limits, types, link names) of one robot."""
    chain, paths, link_geoms = [], [], []
    j, path = first, P0
    while j is not None:
        chain.append(j)
        paths.append(path)
        sub = m.fixed_subtree(j["child"])
        geoms = [(k, T_l @ T, s) for l, T_l in sub.items() for k, T, s in m.links.get(l, [])]
        link_geoms.append(geoms)
        nxt = m.movable_below(sub)
        if len(nxt) > 1:
            raise ValueError(f"link {j['child']}: branching chains are not supported")
        if nxt:
            path = sub[nxt[0]["parent"]]
            j = nxt[0]
        else:
            j = None
    pre = [paths[k] @ chain[k]["origin"] @ _T(_z_to(chain[k]["axis"])) for k in range(len(chain))]
    post = [_T(_z_to(chain[k]["axis"]).T) for k in range(len(chain))]
    dof = len(chain)
    spheres, limits, types = [], [], []
    for k, j in enumerate(chain):
        to_F = _inv(pre[k + 1]) if k + 1 < dof else np.eye(4)  # child link frame -> F_k
        geoms = link_geoms[k]
        cyl = [(T, s) for kind, T, s in geoms if kind == "cylinder"]
        sph = [(T, s) for kind, T, s in geoms if kind == "sphere"]
        box = [(T, s) for kind, T, s in geoms if kind == "box"]
        left = budget - len(sph) - 8 * len(box)
        if left < len(cyl):
            raise ValueError(f"link {j['child']}: its geometry needs more than {budget} spheres")
        link = []
        for (T, s), n in zip(cyl, _allot([2 * s[1] / max(s[0], 1e-9) for _, s in cyl], left)):
            link += _cylinder_spheres(to_F @ T, float(s[0]), float(s[1]), n)
        for T, s in sph:
            link.append(((to_F @ T)[:3, 3], float(s[0])))
        for T, s in box:
            link += _box_spheres(to_F @ T, s)
        spheres.append(link)
        if j["type"] == "continuous":
            limits.append([-math.pi, math.pi])
        else:
            if j["lower"] is None or j["upper"] is None:
                raise ValueError(f"joint {j['name']}: a {j['type']} joint needs lower and upper limits")
            limits.append([j["lower"], j["upper"]])
        types.append(1 if j["type"] == "prismatic" else 0)
    return chain, pre, post, spheres, np.asarray(limits), types


def _z_extent(kind: str, T: np.ndarray, size: np.ndarray) -> Tuple[float, float]:
    c = _aabb(kind, T, size)
    return c[2] - c[5], c[2] + c[5]


def _planar_z(m: UrdfModel, placements: Sequence[np.ndarray]) -> Optional[Tuple[float, float]]:
    """The z interval the robots' collision geometry occupies when every robot is planar (each
    revolute axis vertical and each prismatic axis horizontal in the world at q = 0: then every
    joint motion keeps z and the axes, at every configuration); None otherwise."""
    lo, hi = math.inf, -math.inf
    for P in placements:
        F = {m.root: np.asarray(P, dtype=np.float64)}
        todo = [m.root]
        while todo:
            l = todo.pop()
            for j in m.by_parent.get(l, []):
                F[j["child"]] = F[l] @ j["origin"]  # q = 0
                todo.append(j["child"])
        static = m.fixed_subtree(m.root)
        for j in m.joints:
            if j["type"] not in MOVABLE:
                continue
            ax = (F[j["parent"]] @ j["origin"])[:3, :3] @ (j["axis"] / np.linalg.norm(j["axis"]))
            if j["type"] == "prismatic" and abs(ax[2]) > 1e-12:
                return None
            if j["type"] != "prismatic" and np.linalg.norm(ax[:2]) > 1e-12:
                return None
        for l, T in F.items():
            if l in static:
                continue
            for kind, G, size in m.links.get(l, []):
                a, b = _z_extent(kind, T @ G, size)
                lo, hi = min(lo, a), max(hi, b)
    return (lo, hi) if lo <= hi else None


def load_urdf_arms(path: str, placements: Optional[Sequence[np.ndarray]] = None, boxes=(),
                   tolerance: float = 0.0, static: str = "obstacles",
                   spheres_per_link: int = MAX_SPHERES_PER_LINK) -> SceneDesc:
    """Arm scene of a URDF file (see the module doc).  Robots are listed per placement, in the file
    order of their first joints; ``desc.joint_names`` and ``desc.link_names`` name the joints and
    their child links in configuration order.

    Planar worlds (every robot planar: revolute axes vertical, prismatic axes horizontal): static
    geometry whose z interval is apart from the robots' (by more than the tolerance) can meet no
    robot at any configuration and is left out (``desc.unreachable`` names its links); a sphere
    cover of a flat cylinder reaches about its radius above and below it, and would meet such
    geometry (the table under pinocchio_env.py's planar robots, 1 cm below them)."""
    if static not in ("obstacles", "drop"):
        raise ValueError("static: 'obstacles' or 'drop'")
    if not 1 <= spheres_per_link <= MAX_SPHERES_PER_LINK:
        raise ValueError(f"spheres_per_link must be in 1..{MAX_SPHERES_PER_LINK}")
    m = UrdfModel(path)
    placements = [np.eye(4)] if placements is None else [np.asarray(P, dtype=np.float64) for P in placements]
    static_links = m.fixed_subtree(m.root)
    starts = m.movable_below(static_links)
    if not starts:
        raise ValueError("no movable joints in the URDF file")
    robots = [_robot_chain(m, j, static_links[j["parent"]], spheres_per_link) for j in starts]
    dofs = {len(r[0]) for r in robots}
    if len(dofs) != 1:
        raise ValueError(f"robots with different joint counts {sorted(dofs)}")
    dof = dofs.pop()
    ob = [np.concatenate([np.asarray(c, dtype=np.float64), np.asarray(h, dtype=np.float64) + tolerance / 2])
          for c, h in boxes]
    zr = _planar_z(m, placements)
    unreachable = []
    base, jf, counts, sph, lims, jt, jnames, lnames = [], [], [], [], [], [], [], []
    for P in placements:
        if static == "obstacles":
            for l, T_l in static_links.items():
                for kind, T, s in m.links.get(l, []):
                    if zr is not None:
                        a, b = _z_extent(kind, P @ T_l @ T, s)
                        if a > zr[1] + tolerance or b < zr[0] - tolerance:
                            unreachable.append(l)
                            continue
                    bx = _aabb(kind, P @ T_l @ T, s)
                    bx[3:] += tolerance / 2
                    ob.append(bx)
        for chain, pre, post, spheres, limits, types in robots:
            base.append(_rows12(P @ pre[0]))
            jf.append(np.stack([_chain12(post[k] @ pre[k + 1] if k + 1 < dof else post[k]) for k in range(dof)]))
            for link in spheres:
                counts.append(len(link))
                for c, r in link:
                    sph.append([c[0], c[1], c[2], r + tolerance / 2])
            lims.append(limits)
            jt += types
            jnames += [j["name"] for j in chain]
            lnames += [j["child"] for j in chain]
    n_arms = len(base)
    jt = np.asarray(jt, dtype=np.uint8)
    desc = SceneDesc(kind="arms", limits=np.concatenate(lims).T.copy(), n_arms=n_arms, dof=dof,
                     base=np.stack(base), dh=np.zeros((n_arms, dof, 3)),
                     link_sphere_count=np.asarray(counts, dtype=np.int32),
                     sphere_local=np.asarray(sph, dtype=np.float64).reshape(-1, 4),
                     boxes=np.asarray(ob, dtype=np.float64).reshape(-1, 6), joint_fixed=np.stack(jf),
                     joint_type=jt if jt.any() else None)
    desc.joint_names = jnames
    desc.link_names = lnames
    desc.unreachable = unreachable
    return desc


def link_frames(path: str, q: np.ndarray, joint_names: Sequence[str],
                placements: Optional[Sequence[np.ndarray]] = None) -> Dict[Tuple[int, str], np.ndarray]:
    """World frames of every link for the joint vector q, composed directly from the URDF
    attributes (T_child = T_parent O J(q)), one copy per placement: {(placement, link): 4 x 4}.
    `joint_names` maps q's coordinates to joints (per placement, in order).  An independent check
    of ``load_urdf_arms``."""
    m = UrdfModel(path)
    placements = [np.eye(4)] if placements is None else placements
    n_per = len(joint_names) // len(placements)
    out = {}
    for p, P in enumerate(placements):
        qv = {n: q[p * n_per + i] for i, n in enumerate(joint_names[p * n_per:(p + 1) * n_per])}
        frames = {m.root: np.asarray(P, dtype=np.float64)}
        todo = [m.root]
        while todo:
            l = todo.pop()
            for j in m.by_parent.get(l, []):
                J = np.eye(4) if j["type"] == "fixed" else _joint_motion(j, qv[j["name"]])
                frames[j["child"]] = frames[l] @ j["origin"] @ J
                todo.append(j["child"])
        for l, T in frames.items():
            out[(p, l)] = T
    return out


def geometry_world(path: str, q: np.ndarray, joint_names: Sequence[str],
                   placements: Optional[Sequence[np.ndarray]] = None) -> Dict[Tuple[int, str], list]:
    """World pose of every link's collision geometry: {(placement, link): [(kind, T, size)]}."""
    m = UrdfModel(path)
    F = link_frames(path, q, joint_names, placements)
    return {(p, l): [(k, T @ G, s) for k, G, s in m.links.get(l, [])] for (p, l), T in F.items()}


# ---------------------------------------------------------------------------------------------
def rz(theta: float) -> np.ndarray:
    return _T(np.array([[math.cos(theta), -math.sin(theta), 0], [math.sin(theta), math.cos(theta), 0], [0, 0, 1.0]]))


# make_dual_ur5_waypoint_env (pinocchio_env.py:1025-1187): bases Rz(pi/2) at (+-0.5, -0.5, 0.01),
# a table Box(4, 4, 0.01) at the origin, shoulder_lift limited to [-3.14, 0]
def placement(theta: float, t) -> np.ndarray:
    """pin.SE3(Rz(theta), t) as 4 x 4."""
    T = rz(theta)
    T[:3, 3] = t
    return T


DUAL_UR5_PLACEMENTS = [placement(math.pi / 2, [0.5, -0.5, 0.01]), placement(math.pi / 2, [-0.5, -0.5, 0.01])]
DUAL_UR5_TABLE = (np.zeros(3), np.array([2.0, 2.0, 0.005]))


def load_dual_ur5(path: str, tolerance: float = 0.0, spheres_per_link: int = MAX_SPHERES_PER_LINK) -> SceneDesc:
    """The two-UR5e world of ``pinocchio.random_ur5`` on the primitives URDF
    (``ur5e_constrained_coll_primitives.urdf``).  The geometry fixed under each robot
    (base_link_inertia) is dropped: the reference's pair filter (pinocchio_env.py:1134-1172) removes
    every pair whose first geometry is a base and every pair with the table's parent joint, so it
    tests that geometry only against the other robot's links."""
    desc = load_urdf_arms(path, placements=DUAL_UR5_PLACEMENTS, boxes=[DUAL_UR5_TABLE], tolerance=tolerance,
                          static="drop", spheres_per_link=spheres_per_link)
    lo, hi = desc.limits
    for a in range(desc.n_arms):  # pinocchio_env.py:1174-1185
        lo[a * desc.dof], hi[a * desc.dof] = -3.14, 3.14
        lo[a * desc.dof + 1], hi[a * desc.dof + 1] = -3.14, 0.0
    return desc
