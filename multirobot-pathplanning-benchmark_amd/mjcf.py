"""Robot scenes from MJCF files (the mujoco backend's robot descriptions).

The mujoco environments of the reference load MJCF worlds (mujoco_env.py; e.g.
``assets/models/mujoco/mujoco_4_ur10_world_closer.xml``, which attaches four copies of
``mujoco_ur10/ur10e.xml``) and count a configuration in collision when a contact is closer than
``collision_tolerance`` (mujoco_env.py:404-407).  ``load_mjcf_arms`` turns such a world into an
arm scene of the engine (``SceneDesc`` with ``joint_fixed``):

* every attached (or inline) robot is a serial chain of hinge and slide joints, one or more per
  body (``mujoco.swap``'s ``mj_two_dim.xml``: slide x, slide y, hinge yaw on one body).  MJCF gives
  a body's frame as B = B_parent P J_1(q_1) ... J_m(q_m), with P the body's pos / quat, J the
  rotation about a hinge's axis through its pos, or the translation along a slide's axis.  Each
  joint is one link.  Writing J_k = pre_k Z(q) post_k (pre_k carries z onto the axis, Z = Rz for a
  hinge and Tz for a slide, pre_k also carrying P for the first joint of a body), the frames
  F_k = B_k pre_{k+1} (F_last = B_last) follow F_k = F_{k-1} Z(q_k) C_k with C_k = post_k pre_{k+1}
  (C_last = post_last); the arm's base is F_0 = B_0 pre_1; slide joints are the engine's prismatic
  joints (``SceneDesc.joint_type``), and need a range (their travel);
* collision geoms (capsules, cylinders and spheres; visual geoms, contype = conaffinity = 0, are
  skipped) become spheres in F_k of the body's last joint: a capsule of radius r and half-length h
  is covered by n spheres spaced s = 2h / (n - 1) along its axis, of radius sqrt(r^2 + (s/2)^2),
  so that their union contains the capsule; a cylinder by n spheres at the centres of n slabs of
  height s = 2h / n, radius sqrt(r^2 + (s/2)^2); a link's capsules and cylinders share at most 8
  spheres;
* static box geoms and the boxes of free bodies (at their initial pose) become obstacle boxes
  (axis-aligned when their rotation is a multiple of 90 degrees about the axes; otherwise their
  axis-aligned bounding box); a plane becomes a thick box below it;
* the contact tolerance is added as tolerance / 2 to every sphere radius and box half-extent, so
  that two shapes closer than the tolerance overlap.

Joint limits come from the joints' ``range`` (with the MJCF class defaults).  The geometry is an
approximation of the meshes mujoco collides, so collision results are the engine's on this
spherized model, not mujoco's.
"""
from __future__ import annotations

import math
import os
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Tuple

import numpy as np

from .scenes import SceneDesc

MAX_SPHERES_PER_LINK = 8


# ---------------------------------------------------------------------------------------------
# rigid transforms as 4 x 4 matrices
def _quat(w, x, y, z) -> np.ndarray:
    n = math.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _axis_angle(axis, angle) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)


def _T(R=None, p=None) -> np.ndarray:
    T = np.eye(4)
    if R is not None:
        T[:3, :3] = R
    if p is not None:
        T[:3, 3] = p
    return T


def _inv(T: np.ndarray) -> np.ndarray:
    R, p = T[:3, :3], T[:3, 3]
    return _T(R.T, -R.T @ p)


def _z_to(axis) -> np.ndarray:
    """A rotation that carries the z axis onto `axis` (unit)."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    z = np.array([0.0, 0.0, 1.0])
    c = float(z @ a)
    if c > 1 - 1e-15:
        return np.eye(3)
    if c < -1 + 1e-15:
        return _axis_angle([1.0, 0.0, 0.0], math.pi)
    v = np.cross(z, a)
    return _axis_angle(v, math.atan2(np.linalg.norm(v), c))


# ---------------------------------------------------------------------------------------------
# MJCF attributes with class defaults
class _Defaults:
    """MJCF default classes: class name -> {element tag -> attributes}, inherited down the tree."""

    def __init__(self, root: ET.Element):
        self.classes: Dict[str, Dict[str, Dict[str, str]]] = {}
        top = root.find("default")
        if top is not None:
            self._walk(top, "main", {})

    def _walk(self, el: ET.Element, name: str, inherited: Dict[str, Dict[str, str]]):
        mine = {k: dict(v) for k, v in inherited.items()}
        for child in el:
            if child.tag != "default":
                mine.setdefault(child.tag, {}).update(child.attrib)
        self.classes[name] = mine
        for child in el:
            if child.tag == "default":
                self._walk(child, child.get("class", name), mine)

    def attrs(self, el: ET.Element, cls: str) -> Dict[str, str]:
        a = dict(self.classes.get(cls, self.classes.get("main", {})).get(el.tag, {}))
        a.update(el.attrib)
        return a


def _floats(s: Optional[str], n: int, default) -> np.ndarray:
    if s is None:
        return np.asarray(default, dtype=np.float64)
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    return v if n == 0 else v[:n]


def _body_pose(a: Dict[str, str], degrees: bool) -> np.ndarray:
    p = _floats(a.get("pos"), 3, [0, 0, 0])
    if "quat" in a:
        R = _quat(*_floats(a["quat"], 4, [1, 0, 0, 0]))
    elif "euler" in a:
        e = _floats(a["euler"], 3, [0, 0, 0])
        if degrees:
            e = np.radians(e)
        R = _axis_angle([1, 0, 0], e[0]) @ _axis_angle([0, 1, 0], e[1]) @ _axis_angle([0, 0, 1], e[2])
    elif "axisangle" in a:
        v = _floats(a["axisangle"], 4, [0, 0, 1, 0])
        R = _axis_angle(v[:3], math.radians(v[3]) if degrees else v[3])
    else:
        R = np.eye(3)
    return _T(R, p)


def _degrees(root: ET.Element) -> bool:
    c = root.find("compiler")
    return c is None or c.get("angle", "degree") != "radian"


# ---------------------------------------------------------------------------------------------
def _capsule_spheres(T_geom: np.ndarray, r: float, h: float, n: int) -> List[Tuple[np.ndarray, float]]:
    if n <= 1 or h <= 0:
        return [(T_geom[:3, 3], r + max(h, 0.0))]
    s = 2 * h / (n - 1)
    rr = math.sqrt(r * r + (s / 2) ** 2)
    out = []
    for i in range(n):
        z = -h + i * s
        out.append(((T_geom @ np.array([0, 0, z, 1.0]))[:3], rr))
    return out


def _link_geoms(body: ET.Element, defs: _Defaults, cls: str, deg: bool) -> List[Tuple[str, np.ndarray, np.ndarray]]:
    """Collision geoms of one body: (type, pose in the body frame, size)."""
    out = []
    for g in body.findall("geom"):
        a = defs.attrs(g, g.get("class", cls))
        if a.get("contype", "1") == "0" and a.get("conaffinity", "1") == "0":
            continue
        gtype = a.get("type", "sphere")
        if "fromto" in a:
            f = _floats(a["fromto"], 6, [0] * 6)
            c, d = (f[:3] + f[3:]) / 2, f[3:] - f[:3]
            T = _T(_z_to(d) if np.linalg.norm(d) > 0 else np.eye(3), c)
            size = np.array([_floats(a.get("size"), 1, [0])[0], np.linalg.norm(d) / 2])
        else:
            T = _body_pose(a, deg)
            size = _floats(a.get("size"), 0, [0, 0, 0])
        out.append((gtype, T, size))
    return out


def _allot(lengths: List[float], budget: int) -> List[int]:
    """Spheres per capsule: at least one each, the rest by length, at most `budget` in total."""
    n = [1] * len(lengths)
    if not lengths:
        return n
    for _ in range(budget - len(lengths)):
        i = int(np.argmax([l / (k + 1) for l, k in zip(lengths, n)]))
        if lengths[i] <= 0:
            break
        n[i] += 1
    return n


def _chain(root_body: ET.Element, defs: _Defaults, cls: str):
    """The serial chain below `root_body`: list of (body element, class, [joint attrs]) with one or
    more hinge / slide joints per body, in the order MuJoCo applies them."""
    chain = []
    body, c = root_body, cls
    while True:
        kids = [b for b in body.findall("body")]
        nxt = None
        for b in kids:
            bc = b.get("childclass", c)
            if b.find("joint") is not None or b.findall(".//joint"):
                if nxt is not None:
                    raise ValueError("only serial chains are supported")
                nxt = (b, bc)
        if nxt is None:
            return chain
        body, c = nxt
        joints = body.findall("joint")
        if not joints:
            raise ValueError(f"body {body.get('name')}: a chain body without a joint is not supported")
        jas = [defs.attrs(j, j.get("class", c)) for j in joints]
        for ja in jas:
            if ja.get("type", "hinge") not in ("hinge", "slide"):
                raise ValueError(f"body {body.get('name')}: only hinge and slide joints are supported")
        chain.append((body, c, jas))


def _cylinder_spheres(T_geom: np.ndarray, r: float, h: float, n: int) -> List[Tuple[np.ndarray, float]]:
    """n spheres covering a cylinder of radius r and half-length h (axis z of T_geom): the axis cut
    into n slabs of height s = 2h / n, one sphere of radius sqrt(r^2 + (s/2)^2) at each slab's
    centre, which contains the slab."""
    n = max(1, n)
    s = 2 * h / n
    rr = math.sqrt(r * r + (s / 2) ** 2)
    return [((T_geom @ np.array([0, 0, -h + s / 2 + i * s, 1.0]))[:3], rr) for i in range(n)]


def _robot(root_body: ET.Element, defs: _Defaults, cls: str, B0: np.ndarray, degrees: bool, budget: int):
    """Base transform, joint transforms C_k, spheres per link (in F_k), limits, link names and joint
    types (0 hinge, 1 slide) of one chain.  A body with several joints gives one link per joint; its
    geoms belong to the link of its last joint."""
    chain = _chain(root_body, defs, cls)
    if not chain:
        raise ValueError("no joints below the robot's base body")
    joints = []  # (body, class, joint attrs, body pose before the joint, last joint of its body)
    for body, c, jas in chain:
        P = _body_pose(defs.attrs(body, c), degrees)
        for i, ja in enumerate(jas):
            joints.append((body, c, ja, P if i == 0 else np.eye(4), i + 1 == len(jas)))
    pre, post, types = [], [], []
    for body, c, ja, P, _ in joints:
        axis = _floats(ja.get("axis"), 3, [0, 0, 1])
        U = _z_to(axis)
        if ja.get("type", "hinge") == "slide":  # T(axis q) = U Tz(q) U^T
            pre.append(P @ _T(U))
            post.append(_T(U.T))
            types.append(1)
        else:  # T(p) Rot(axis, q) T(-p) = T(p) U Rz(q) U^T T(-p)
            jp = _floats(ja.get("pos"), 3, [0, 0, 0])
            pre.append(P @ _T(None, jp) @ _T(U))
            post.append(_T(U.T) @ _T(None, -jp))
            types.append(0)
    dof = len(joints)
    C = [post[k] @ pre[k + 1] if k + 1 < dof else post[k] for k in range(dof)]
    base = B0 @ pre[0]
    spheres, limits = [], []
    for k, (body, c, ja, _, last) in enumerate(joints):
        link = []
        if last:
            to_F = _inv(pre[k + 1]) if k + 1 < dof else np.eye(4)  # body frame -> F_k
            geoms = _link_geoms(body, defs, c, degrees)
            caps = [(T, s) for t, T, s in geoms if t == "capsule"]
            cyls = [(T, s) for t, T, s in geoms if t == "cylinder"]
            sph = [(T, s) for t, T, s in geoms if t == "sphere"]
            other = [t for t, _, _ in geoms if t not in ("capsule", "cylinder", "sphere")]
            if other:
                raise ValueError(f"body {body.get('name')}: collision geom types {other} are not supported")
            n_alloc = _allot([2 * s[1] / max(s[0], 1e-9) for _, s in caps + cyls], budget - len(sph))
            for (T, s), n in zip(caps, n_alloc[:len(caps)]):
                link += _capsule_spheres(to_F @ T, float(s[0]), float(s[1]), n)
            for (T, s), n in zip(cyls, n_alloc[len(caps):]):
                link += _cylinder_spheres(to_F @ T, float(s[0]), float(s[1]), n)
            for T, s in sph:
                link.append(((to_F @ T)[:3, 3], float(s[0])))
            if len(link) > budget:
                raise ValueError(f"body {body.get('name')}: more than {budget} spheres")
        spheres.append(link)
        rng = ja.get("range")
        if ja.get("type", "hinge") == "slide":
            if rng is None:
                raise ValueError(f"body {body.get('name')}: a slide joint needs a range (its travel)")
            limits.append(_floats(rng, 2, [0, 0]))
        elif rng is None:
            limits.append(np.array([-math.pi, math.pi]))
        else:
            lim = _floats(rng, 2, [-math.pi, math.pi])
            limits.append(np.radians(lim) if degrees else lim)
    names = [ja.get("name", body.get("name")) for body, _, ja, _, _ in joints]
    return base, np.stack(C), spheres, np.stack(limits), names, types


def _rows12(T: np.ndarray) -> np.ndarray:
    """[R row-major | t] (the engine's base layout)."""
    return np.concatenate([T[:3, :3].reshape(-1), T[:3, 3]])


def _chain12(T: np.ndarray) -> np.ndarray:
    """Rows R00 R01 R02 t0, R10 R11 R12 t1, R20 R21 R22 t2 (the joint_fixed layout)."""
    return T[:3, :4].reshape(-1)


def _box_of(T: np.ndarray, half: np.ndarray) -> np.ndarray:
    """Centre and half-extents of the axis-aligned box of an oriented box."""
    R = T[:3, :3]
    return np.concatenate([T[:3, 3], np.abs(R) @ half])


def load_mjcf_arms(path: str, tolerance: float = 0.0, spheres_per_link: int = MAX_SPHERES_PER_LINK,
                   plane_depth: float = 1.0, plane_extent: float = 10.0) -> SceneDesc:
    """Arm scene of an MJCF world whose robots are serial hinge chains (see the module doc)."""
    if not 1 <= spheres_per_link <= MAX_SPHERES_PER_LINK:
        raise ValueError(f"spheres_per_link must be in 1..{MAX_SPHERES_PER_LINK}")
    world = ET.parse(path).getroot()
    here = os.path.dirname(os.path.abspath(path))
    wdeg = _degrees(world)
    wdefs = _Defaults(world)
    models = {}
    asset = world.find("asset")
    if asset is not None:
        for m in asset.findall("model"):
            models[m.get("name")] = os.path.join(here, m.get("file"))
    robots, boxes = [], []

    def add_box(T, half):
        boxes.append(_box_of(T, np.asarray(half, dtype=np.float64) + tolerance / 2))

    def walk(el: ET.Element, B: np.ndarray, cls: str):
        for g in el.findall("geom"):
            a = wdefs.attrs(g, g.get("class", cls))
            t = a.get("type", "sphere")
            Tg = B @ _body_pose(a, wdeg)
            if t == "box":
                add_box(Tg, _floats(a.get("size"), 3, [0, 0, 0]))
            elif t == "plane":
                # the half-space below the plane (its normal is the geom's z axis; only level planes)
                if abs(Tg[2, 2] - 1.0) > 1e-12:
                    raise ValueError("only horizontal planes are supported")
                c = Tg[:3, 3] - np.array([0, 0, plane_depth / 2])
                boxes.append(np.concatenate([c, [plane_extent, plane_extent, plane_depth / 2 + tolerance / 2]]))
        for b in el.findall("body"):
            a = wdefs.attrs(b, cls)
            Bb = B @ _body_pose(a, wdeg)
            att = b.find("attach")
            if att is not None:
                mroot = ET.parse(models[att.get("model")]).getroot()
                mdefs = _Defaults(mroot)
                rb = next(x for x in mroot.iter("body") if x.get("name") == att.get("body"))
                rcls = rb.get("childclass", "main")
                Brb = Bb @ _body_pose(mdefs.attrs(rb, rcls), _degrees(mroot))
                robots.append(_robot(rb, mdefs, rcls, Brb, _degrees(mroot), spheres_per_link))
                continue
            if b.find("freejoint") is not None or (b.find("joint") is None and not b.findall(".//joint")):
                walk(b, Bb, b.get("childclass", cls))  # static or free bodies: their geoms are obstacles
                continue
            holder = ET.Element("body")  # the chain starts at b itself
            holder.append(b)
            robots.append(_robot(holder, wdefs, cls, B, wdeg, spheres_per_link))

    walk(world.find("worldbody"), np.eye(4), "main")
    if not robots:
        raise ValueError("no robot chains in the MJCF world")
    dofs = {r[1].shape[0] for r in robots}
    if len(dofs) != 1:
        raise ValueError(f"robots with different joint counts {sorted(dofs)}")
    dof = dofs.pop()
    n_arms = len(robots)
    base = np.stack([_rows12(r[0]) for r in robots])
    joint_fixed = np.stack([np.stack([_chain12(C) for C in r[1]]) for r in robots])
    counts, sph = [], []
    for r in robots:
        for link in r[2]:
            counts.append(len(link))
            for c, rad in link:
                sph.append([c[0], c[1], c[2], rad + tolerance / 2])
    limits = np.concatenate([r[3] for r in robots]).T.copy()
    jt = np.asarray([t for r in robots for t in r[5]], dtype=np.uint8)
    desc = SceneDesc(kind="arms", limits=limits, n_arms=n_arms, dof=dof, base=base,
                     dh=np.zeros((n_arms, dof, 3)), link_sphere_count=np.asarray(counts, dtype=np.int32),
                     sphere_local=np.asarray(sph, dtype=np.float64).reshape(-1, 4),
                     boxes=np.asarray(boxes, dtype=np.float64).reshape(-1, 6), joint_fixed=joint_fixed,
                     joint_type=jt if jt.any() else None)
    desc.link_names = [n for r in robots for n in r[4]]
    return desc


def body_frames(path: str, q: np.ndarray) -> List[np.ndarray]:
    """World frames of every robot joint of an MJCF world for the joint vector q (robots in file
    order, joints in chain order), composed directly from the MJCF attributes: a body's frame is its
    parent's times its pose P, then times T(p) Rot(axis, q) T(-p) for each hinge and T(axis q) for
    each slide joint, in order; the frame after each joint is listed.  An independent check of
    ``load_mjcf_arms``."""
    world = ET.parse(path).getroot()
    here = os.path.dirname(os.path.abspath(path))
    wdeg = _degrees(world)
    wdefs = _Defaults(world)
    models = {m.get("name"): os.path.join(here, m.get("file")) for m in world.iter("model")}
    frames: List[np.ndarray] = []
    qi = [0]

    def chain_frames(rb, defs, cls, B, deg):
        for body, c, jas in _chain(rb, defs, cls):
            B = B @ _body_pose(defs.attrs(body, c), deg)
            for ja in jas:
                axis = _floats(ja.get("axis"), 3, [0, 0, 1])
                if ja.get("type", "hinge") == "slide":
                    J = _T(None, axis / np.linalg.norm(axis) * q[qi[0]])
                else:
                    jp = _floats(ja.get("pos"), 3, [0, 0, 0])
                    J = _T(None, jp) @ _T(_axis_angle(axis, q[qi[0]])) @ _T(None, -jp)
                qi[0] += 1
                B = B @ J
                frames.append(B)

    def walk(el, B, cls):
        for b in el.findall("body"):
            Bb = B @ _body_pose(wdefs.attrs(b, cls), wdeg)
            att = b.find("attach")
            if att is not None:
                mroot = ET.parse(models[att.get("model")]).getroot()
                mdefs = _Defaults(mroot)
                rb = next(x for x in mroot.iter("body") if x.get("name") == att.get("body"))
                rcls = rb.get("childclass", "main")
                chain_frames(rb, mdefs, rcls, Bb @ _body_pose(mdefs.attrs(rb, rcls), _degrees(mroot)),
                             _degrees(mroot))
                continue
            if b.find("freejoint") is not None or (b.find("joint") is None and not b.findall(".//joint")):
                walk(b, Bb, b.get("childclass", cls))
                continue
            holder = ET.Element("body")
            holder.append(b)
            chain_frames(holder, wdefs, cls, B, wdeg)

    walk(world.find("worldbody"), np.eye(4), "main")
    return frames
