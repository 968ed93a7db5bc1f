"""Independent brute-force checker of the arm scenes' composite predicate (numpy + scipy).

Test infrastructure.  Written from the definition of the predicate, not from the C oracle or the
kernels, so that a rule shared by those two (which pairs are tested, the mask, the active arms of a
mode, the held boxes) is checked by a third computation:

* kinematics as products of 4x4 homogeneous matrices (np.cos / np.sin):
  DH joints T_k = T_{k-1} Rz(q_k) Tz(d_k) Tx(a_k) Rx(alpha_k); chain joints T_k = T_{k-1} Rz(q_k) C_k
  (revolute) or T_{k-1} Tz(q_k) C_k (prismatic);
* spheres fixed in their link frames; sphere vs obstacle box (axis aligned) and sphere vs held box
  (oriented) by the distance from the centre to its projection onto the box, closed (touching
  collides); sphere vs sphere closed;
* pair set: spheres of one arm on links at least two apart; every sphere pair of two different arms;
  every sphere against every obstacle box;
* arms [0, n_active[mode]) exist in a mode;
* a held box of arm a (pose T_ee(q) T_rel, half extents h) meets the obstacle boxes, the other held
  boxes, the resting boxes and every sphere except those of its holding link, the last link of arm
  a: the attachment deactivates only the (object, new parent) pair (pinocchio_env.py:266-272; rai
  attach + setContact(-1), rai_base_env.py:801-805);
* a resting box of the mode (a movable object put down at a world pose) meets every sphere and the
  held boxes; static against static (obstacles, other resting boxes) is not a state's collision;
* robot mask: a pair counts iff every robot in it is in the mask (rai_base_env.py:560-586,
  collisions with robots outside the set are ignored);
* non-finite coordinates of the mode's arms collide, and so does a prismatic joint beyond its
  travel (chain joints T_k = T_{k-1} Tz(q_k) C_k);
* box vs box by linear programming (scipy linprog): the largest s such that both boxes, shrunk by s
  on every face, still share a point.  s >= 0: overlap (closed); no separating-axis test.

Each pair yields a signed slack (<= 0: contact).  States whose decision depends on a pair within
`eps` of contact are reported as undecided, because two correct computations with different
rounding may disagree there.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.optimize import linprog


def rot_z(q):
    c, s = math.cos(q), math.sin(q)
    T = np.eye(4)
    T[0, 0], T[0, 1], T[1, 0], T[1, 1] = c, -s, s, c
    return T


def rot_x(al):
    c, s = math.cos(al), math.sin(al)
    T = np.eye(4)
    T[1, 1], T[1, 2], T[2, 1], T[2, 2] = c, -s, s, c
    return T


def trans(x, y, z):
    T = np.eye(4)
    T[:3, 3] = (x, y, z)
    return T


def from12(v):
    """[R row-major | t] (12 numbers) -> 4x4."""
    v = np.asarray(v, dtype=np.float64)
    T = np.eye(4)
    T[:3, :3] = v[:9].reshape(3, 3)
    T[:3, 3] = v[9:12]
    return T


def from_rows12(v):
    """rows (R00 R01 R02 t0, R10 .. t1, R20 .. t2) -> 4x4 (the chain-joint layout)."""
    T = np.eye(4)
    T[:3, :] = np.asarray(v, dtype=np.float64).reshape(3, 4)
    return T


def link_frames(desc, q):
    """[n_arms * dof, 4, 4] world frames of every link."""
    na, dof = desc.n_arms, desc.dof
    out = np.zeros((na * dof, 4, 4))
    for a in range(na):
        T = from12(desc.base[a])
        for k in range(dof):
            l = a * dof + k
            if getattr(desc, "joint_fixed", None) is not None:
                jt = getattr(desc, "joint_type", None)
                J = trans(0, 0, q[l]) if jt is not None and jt[l] == 1 else rot_z(q[l])  # prismatic / revolute
                T = T @ J @ from_rows12(desc.joint_fixed[a, k])
            else:
                d, aa, al = desc.dh[a, k]
                T = T @ rot_z(q[l]) @ trans(0, 0, d) @ trans(aa, 0, 0) @ rot_x(al)
            out[l] = T
    return out


def sphere_table(desc):
    """(link of each sphere, arm of each sphere) in the listing order of sphere_local."""
    link = np.repeat(np.arange(desc.n_arms * desc.dof), desc.link_sphere_count)
    return link, link // desc.dof


def _box_rows(T, h):
    """A x <= b of the box {x : |R^T (x - t)| <= h} as rows (u_k . x <= u_k . t + h_k, -u_k . x <= ...)."""
    R, t = T[:3, :3], T[:3, 3]
    A = np.concatenate([R.T, -R.T])
    b = np.concatenate([R.T @ t + h, -(R.T @ t) + h])
    return A, b


def box_box_slack(T1, h1, T2, h2, eps):
    """-s*, s* the largest shrink of both boxes (per face) at which they still share a point.
    A cheap bounding-sphere test settles clearly separated pairs without the LP."""
    gap = np.linalg.norm(T1[:3, 3] - T2[:3, 3]) - np.linalg.norm(h1) - np.linalg.norm(h2)
    if gap > 4 * eps:
        return gap
    A1, b1 = _box_rows(T1, h1)
    A2, b2 = _box_rows(T2, h2)
    A = np.concatenate([A1, A2])
    # u . x <= u . t + h - s  <=>  u . x + s <= u . t + h
    A_ub = np.concatenate([A, np.ones((12, 1))], axis=1)
    b_ub = np.concatenate([b1, b2])
    r = linprog(c=[0, 0, 0, -1], A_ub=A_ub, b_ub=b_ub,
                bounds=[(None, None)] * 3 + [(-100.0, None)], method="highs")
    assert r.status == 0, r.message
    return -r.x[3]


def sph_box_slack(c, r, T, h):
    """squared distance from c to the box T (half h) minus r^2"""
    R, t = T[:3, :3], T[:3, 3]
    l = R.T @ (c - t)
    e = np.clip(l, -h, h) - l
    return float(e @ e - r * r)


CATS = ("sphere_box", "self", "inter_arm", "held_box", "held_sphere_other", "held_sphere_own",
        "held_held", "rest_sphere", "rest_held")


def pair_slacks(desc, q, mode=0, eps=1e-9, mutant=None):
    """Every pair of the state's pair set: list of (category, robots (frozenset), slack).
    Also returns the excluded held-box vs holding-link pairs under category "held_sphere_holder".
    `mutant` plants a mistake, for tests that show the comparison would catch it: "gap1" (self
    pairs of adjacent links counted), "all_active" (every arm present in every mode)."""
    q = np.asarray(q, dtype=np.float64)
    dof, na = desc.dof, desc.n_arms
    if desc.modes is not None:
        n_act = na if mutant == "all_active" else int(desc.modes["n_active"][mode])
        has = np.asarray(desc.modes["has_grasp"][mode], dtype=bool)
        rel = desc.modes["grasp_rel"][mode]
        half = desc.modes["grasp_half"][mode]
    else:
        n_act, has = na, np.zeros(na, dtype=bool)
        rel = half = None
    n_slots = len(has)
    F = link_frames(desc, q)
    link, arm = sphere_table(desc)
    loc = desc.sphere_local
    C = np.einsum("sij,sj->si", F[link][:, :3, :3], loc[:, :3]) + F[link][:, :3, 3]
    r = loc[:, 3]
    live = arm < n_act
    out = []
    idx = np.nonzero(live)[0]
    boxes = desc.boxes
    for i in idx:
        for b in range(len(boxes)):
            lo, hi = boxes[b, :3] - boxes[b, 3:], boxes[b, :3] + boxes[b, 3:]
            e = np.clip(C[i], lo, hi) - C[i]
            out.append(("sphere_box", frozenset([arm[i]]), float(e @ e - r[i] ** 2)))
    for x in range(len(idx)):
        i = idx[x]
        for y in range(x + 1, len(idx)):
            j = idx[y]
            if arm[i] == arm[j]:
                if abs(int(link[i]) - int(link[j])) < (1 if mutant == "gap1" else 2):
                    continue
                cat = "self"
            else:
                cat = "inter_arm"
            d = C[i] - C[j]
            out.append((cat, frozenset([arm[i], arm[j]]), float(d @ d - (r[i] + r[j]) ** 2)))
    G = {}
    for a in range(n_act):
        if has[a]:
            G[a] = F[a * dof + dof - 1] @ from12(rel[a])
    for j in range(na, n_slots):  # resting boxes: world poses
        if has[j]:
            hj = np.asarray(half[j], dtype=np.float64)
            Gj = from12(rel[j])
            for i in idx:
                out.append(("rest_sphere", frozenset([arm[i]]), sph_box_slack(C[i], r[i], Gj, hj)))
            for a, Ga in G.items():
                out.append(("rest_held", frozenset([a]),
                            box_box_slack(Ga, np.asarray(half[a], dtype=np.float64), Gj, hj, eps)))
    for a, Ga in G.items():
        ha = np.asarray(half[a], dtype=np.float64)
        for b in range(len(boxes)):
            out.append(("held_box", frozenset([a]),
                        box_box_slack(trans(*boxes[b, :3]), boxes[b, 3:], Ga, ha, eps)))
        for i in idx:
            if arm[i] != a:
                cat = "held_sphere_other"
            elif link[i] == a * dof + dof - 1:
                cat = "held_sphere_holder"  # excluded from the pair set
            else:
                cat = "held_sphere_own"
            out.append((cat, frozenset([a, arm[i]]), sph_box_slack(C[i], r[i], Ga, ha)))
        for b2, Gb in G.items():
            if b2 > a:
                out.append(("held_held", frozenset([a, b2]),
                            box_box_slack(Ga, ha, Gb, np.asarray(half[b2], dtype=np.float64), eps)))
    return out


def decide(pairs, mask=0xFFFFFFFF, eps=1e-9):
    """True: collision, False: free, None: some counted pair within eps of contact decides."""
    counted = [s for c, rob, s in pairs
               if c != "held_sphere_holder" and all((mask >> int(x)) & 1 for x in rob)]
    if any(s < -eps for s in counted):
        return True
    if all(s > eps for s in counted):
        return False
    return None


def out_of_range(desc, q, mode=0):
    """A coordinate of the mode's arms is non-finite, or a prismatic joint is beyond its travel."""
    n_act = desc.n_arms if desc.modes is None else int(desc.modes["n_active"][mode])
    qa = np.asarray(q, dtype=np.float64)[:n_act * desc.dof]
    if not np.all(np.isfinite(qa)):
        return True
    jt = getattr(desc, "joint_type", None)
    if jt is None:
        return False
    tr = desc.joint_travel()[:n_act * desc.dof]
    return bool(np.any((np.asarray(jt)[:n_act * desc.dof] == 1) & (np.abs(qa) > tr)))


def collides(desc, q, mode=0, mask=0xFFFFFFFF, eps=1e-9):
    if out_of_range(desc, q, mode):
        return True
    return decide(pair_slacks(desc, q, mode, eps), mask, eps)


def deciding_category(pairs, mask=0xFFFFFFFF, eps=1e-9):
    """The category of the only colliding counted pairs when exactly one category collides (the
    state would be free without it), else None."""
    cats = set()
    for c, rob, s in pairs:
        if c != "held_sphere_holder" and all((mask >> int(x)) & 1 for x in rob) and s < -eps:
            cats.add(c)
    return next(iter(cats)) if len(cats) == 1 else None


# ---------------------------------------------------------------------------------------------
def _rand_rot(rng):
    u = rng.normal(size=4)
    w, x, y, z = u / np.linalg.norm(u)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def tiny_arm_scene(rng, scenes, n_arms=3, dof=4, n_boxes=3, n_modes=6, chain=False, spread=0.35,
                   held_half=(0.02, 0.09), rel_offset=0.15, sphere_r=(0.02, 0.07), n_rest=0, prism=0.0):
    """A small arm scene with modes: arm bases on a circle of radius `spread` (close enough for
    inter-arm contacts), 1..3 spheres per link, a few boxes, and modes with fewer active arms and
    held boxes of various offsets.  Chain scenes: each joint is prismatic with probability `prism`
    (limits +-0.3)."""
    base = np.zeros((n_arms, 12))
    for a in range(n_arms):
        ang = 2 * math.pi * a / n_arms
        x, y = spread * math.cos(ang), spread * math.sin(ang)
        yaw = rng.uniform(-math.pi, math.pi)
        c, s = math.cos(yaw), math.sin(yaw)
        base[a] = [c, -s, 0, s, c, 0, 0, 0, 1, x, y, rng.uniform(0.0, 0.1)]
    dh = np.zeros((n_arms, dof, 3))
    jf = np.zeros((n_arms, dof, 12))
    counts, sph = [], []
    for a in range(n_arms):
        for k in range(dof):
            d, aa = rng.uniform(0.0, 0.2), rng.uniform(0.05, 0.2)
            al = rng.choice([0.0, math.pi / 2, -math.pi / 2, rng.uniform(-math.pi, math.pi)])
            dh[a, k] = (d, aa, al)
            if chain:
                T = np.eye(4)
                T[:3, :3] = _rand_rot(rng)
                T[:3, 3] = rng.uniform(-0.15, 0.15, 3)
                jf[a, k] = T[:3, :].reshape(-1)
                prev = -T[:3, :3].T @ T[:3, 3]
            else:
                prev = np.array([-aa, -d * math.sin(al), -d * math.cos(al)])
            n = int(rng.integers(1, 4))
            counts.append(n)
            for j in range(n):
                t = rng.uniform(0.1, 0.9)
                sph.append([*(prev * t + rng.normal(0, 0.01, 3)), rng.uniform(*sphere_r)])
    boxes = np.concatenate([rng.uniform(-0.4, 0.4, (n_boxes, 2)), rng.uniform(0.0, 0.3, (n_boxes, 1)),
                            rng.uniform(0.03, 0.1, (n_boxes, 3))], axis=1)
    D = n_arms * dof
    desc = scenes.SceneDesc(kind="arms", limits=np.stack([np.full(D, -math.pi), np.full(D, math.pi)]),
                            n_arms=n_arms, dof=dof, base=base, dh=dh,
                            link_sphere_count=np.array(counts, dtype=np.int32),
                            sphere_local=np.array(sph, dtype=np.float64), boxes=boxes,
                            joint_fixed=jf if chain else None)
    if chain and prism > 0:
        desc.joint_type = (rng.uniform(size=D) < prism).astype(np.uint8)
        desc.limits[:, desc.joint_type == 1] = [[-0.3], [0.3]]
    n_active = rng.integers(1, n_arms + 1, size=n_modes).astype(np.int32)
    n_active[0] = n_arms
    ns = n_arms + n_rest
    has = (rng.uniform(size=(n_modes, ns)) < 0.8).astype(np.uint8)
    has[0] = 0
    for m in range(n_modes):
        has[m, n_active[m]:n_arms] = 0
    rel = np.zeros((n_modes, ns, 12))
    half = rng.uniform(held_half[0], held_half[1], size=(n_modes, ns, 3))
    for m in range(n_modes):
        for a in range(ns):
            rel[m, a, :9] = _rand_rot(rng).reshape(-1)
            rel[m, a, 9:] = rng.uniform(-rel_offset, rel_offset, 3)
            if a >= n_arms:  # resting boxes among the arms
                rel[m, a, 9:] = [*rng.uniform(-0.4, 0.4, 2), rng.uniform(0.0, 0.4)]
    desc.modes = {"n_active": n_active, "has_grasp": has, "grasp_rel": rel, "grasp_half": half, "n_rest": n_rest}
    return desc
