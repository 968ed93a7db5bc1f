"""Pins the CPU oracle to facts it does not restate itself.

* C1 semantics: a numpy restatement that uses the reference's own numpy calls
  (oracle/abstract_np.py) must agree bit for bit, including near contact and in 10-D.
* Touching cases fixed by the reference's operators: disk-disk touching is free (strict <,
  abstract_env.py:264-268), disk-rectangle touching collides (<=, :84).
* sincos: against libm.
* DH forward kinematics: against an independent numpy product of 4x4 homogeneous matrices.
* 2D / 3D separating-axis tests: against polygon intersection / LP feasibility.
* Sphere-box: against the textbook per-axis distance form.
* Path schedules: the reference's call counts (tests/test.py:93-122) through the oracle's
  restatement of the schedules (oracle/path_schedule.py) and the package's default, with a mocked
  is_collision_free; the default's boolean against the schedules on random paths.
* Clearance gradient: against central finite differences of the FP64 cost.
"""
import math
from unittest import mock

import numpy as np
import pytest

from helpers import O, pkg, scenes
from oracle import path_schedule as PS
from oracle.abstract_np import AbstractNp
from oracle_env import OracleEnv


# ---------------------------------------------------------------------------------------------
# C1 vs the numpy restatement
@pytest.mark.parametrize("seed", [0, 1])
def test_c1_configs_match_numpy(seed):
    rng = np.random.default_rng(seed)
    desc = scenes.planar_disks(rng=rng)
    o, ref = O.Scene(desc), AbstractNp(desc)
    Q = scenes.uniform_configs(rng, desc, 3000)
    # near contact: robot 2 at distance ~0.2 of robot 1
    phi = rng.uniform(-math.pi, math.pi, 1500)
    R = 0.2 + rng.choice([-1e-15, 0.0, 1e-15, 2e-16, -2e-16], 1500)
    Q[:1500, 2] = Q[:1500, 0] + R * np.cos(phi)
    Q[:1500, 3] = Q[:1500, 1] + R * np.sin(phi)
    got = o.check_configs(Q)
    exp = np.array([ref.is_collision_free(q) for q in Q])
    assert np.array_equal(got, exp)


def test_c1_edges_match_numpy():
    rng = np.random.default_rng(2)
    desc = scenes.planar_disks(rng=rng)
    o, ref = O.Scene(desc), AbstractNp(desc)
    q1, q2 = scenes.random_pair_edges(rng, desc, 300)
    q2 = q1 + (q2 - q1) * 0.2
    for a, b in zip(q1, q2):
        assert o.edge_free(a, b, 0.01) == ref.is_edge_collision_free(a, b, 0.01)
        N = max(2, int(np.max(np.abs(a - b)) / 0.01) + 1)
        s = int(rng.integers(0, N + 1))
        assert o.edge_free(a, b, 0.01, N=N, N_start=s, N_max=s + 10, include_endpoints=True) == \
            ref.is_edge_collision_free(a, b, 0.01, include_endpoints=True, N_start=s, N_max=s + 10, N=N)


def test_center_rect_10d_matches_numpy():
    """10-D robots: norm through np.dot (sequential fma chain), np.sum pairwise over 10 terms."""
    desc = scenes.center_rect_scene(10)
    o, ref = O.Scene(desc), AbstractNp(desc)
    rng = np.random.default_rng(3)
    Q = rng.uniform(-0.45, 0.45, size=(4000, 20)) * rng.uniform(0.2, 1.0, size=(4000, 1))
    got = o.check_configs(Q)
    exp = np.array([ref.is_collision_free(q) for q in Q])
    assert 0.05 < exp.mean() < 0.95
    assert np.array_equal(got, exp)


def test_touching_semantics():
    lim = np.array([[-2.0] * 4, [2.0] * 4])
    desc = scenes.disks_scene([0.25, 0.25], 2, lim, rects=[(np.array([0.0, 0.0]), np.array([1.0, 1.0]))])
    o = O.Scene(desc)
    # disks touching each other (exact distance 0.5 = 0.25 + 0.25): free (strict <)
    assert o.config_free(np.array([1.0, 1.5, 1.5, 1.5]))
    assert not o.config_free(np.array([1.0, 1.5, 1.4999999999999998, 1.5]))
    # disk touching the rectangle face at x = 0.5 (exact): collision (<=)
    assert not o.config_free(np.array([0.75, 0.0, -1.75, 1.75]))
    assert o.config_free(np.array([0.7500000000000001, 0.0, -1.75, 1.75]))


# ---------------------------------------------------------------------------------------------
def test_sincos_vs_libm():
    rng = np.random.default_rng(4)
    xs = np.concatenate([rng.uniform(-10, 10, 50000), np.linspace(-7, 7, 2001)])
    err = 0.0
    for x in xs:
        s, c = O.sincos(x)
        err = max(err, abs(s - math.sin(x)), abs(c - math.cos(x)))
    assert err <= 4.5e-16


def _dh_numpy(desc, q):
    """[links, 4, 4] from np.cos / np.sin and matrix products Rz Tz Tx Rx."""
    out = []
    for a in range(desc.n_arms):
        T = np.eye(4)
        T[:3, :3] = desc.base[a, :9].reshape(3, 3)
        T[:3, 3] = desc.base[a, 9:]
        for k in range(desc.dof):
            d, aa, al = desc.dh[a, k]
            th = q[a * desc.dof + k]
            Rz = np.array([[np.cos(th), -np.sin(th), 0, 0], [np.sin(th), np.cos(th), 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
            Tz = np.eye(4)
            Tz[2, 3] = d
            Tx = np.eye(4)
            Tx[0, 3] = aa
            Rx = np.array([[1, 0, 0, 0], [0, np.cos(al), -np.sin(al), 0], [0, np.sin(al), np.cos(al), 0], [0, 0, 0, 1]])
            T = T @ Rz @ Tz @ Tx @ Rx
            out.append(T.copy())
    return np.array(out)


def test_fk_vs_homogeneous_product():
    rng = np.random.default_rng(5)
    desc = scenes.quad_arms(rng=rng)
    o = O.Scene(desc)
    for q in scenes.uniform_configs(rng, desc, 50):
        T = o.fk_matrices(q)
        ref = _dh_numpy(desc, q)
        assert np.max(np.abs(T - ref)) < 1e-13
        R = T[:, :3, :3]
        assert np.max(np.abs(R @ np.transpose(R, (0, 2, 1)) - np.eye(3))) < 1e-13


def test_fk_zero_configuration_closed_form():
    """All joint angles 0 and all alpha 0: R = base rotation, p = base + R (sum a, 0, sum d)."""
    desc = scenes.quad_arms(rng=np.random.default_rng(6))
    desc.dh[:, :, 2] = 0.0
    o = O.Scene(desc)
    T = o.fk_matrices(np.zeros(desc.dim))
    for a in range(desc.n_arms):
        Rb = desc.base[a, :9].reshape(3, 3)
        for k in range(desc.dof):
            sa = desc.dh[a, : k + 1, 1].sum()
            sd = desc.dh[a, : k + 1, 0].sum()
            exp = desc.base[a, 9:] + Rb @ np.array([sa, 0.0, sd])
            assert np.allclose(T[a * desc.dof + k, :3, 3], exp, atol=1e-14)


def _poly(b):
    x, y, ux, uy, vx, vy, h0, h1 = b
    c = np.array([x, y])
    u = np.array([ux, uy]) * h0
    v = np.array([vx, vy]) * h1
    return [c + u + v, c - u + v, c - u - v, c + u - v]


def _seg_intersect(p1, p2, p3, p4):
    def orient(a, b, c):
        return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])

    d1, d2 = orient(p3, p4, p1), orient(p3, p4, p2)
    d3, d4 = orient(p1, p2, p3), orient(p1, p2, p4)
    return (d1 * d2 < 0) and (d3 * d4 < 0)


def _inside(p, poly):
    def cross(u, v):
        return u[0] * v[1] - u[1] * v[0]

    s = [cross(poly[(i + 1) % 4] - poly[i], p - poly[i]) for i in range(4)]
    return all(v > 0 for v in s) or all(v < 0 for v in s)


def test_obb2_vs_polygon_intersection():
    rng = np.random.default_rng(7)
    checked = 0
    for _ in range(4000):
        def box():
            th = rng.uniform(-math.pi, math.pi)
            s, c = O.sincos(th)
            return np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), c, s, -s, c, rng.uniform(0.05, 0.6),
                             rng.uniform(0.05, 0.6)])
        A, B = box(), box()
        PA, PB = _poly(A), _poly(B)
        inter = any(_inside(p, PB) for p in PA) or any(_inside(p, PA) for p in PB) or any(
            _seg_intersect(PA[i], PA[(i + 1) % 4], PB[j], PB[(j + 1) % 4]) for i in range(4) for j in range(4))
        # skip near-touching pairs: grow / shrink A by 1e-7 must not change the brute-force answer
        Ag, As = A.copy(), A.copy()
        Ag[6:] += 1e-7
        As[6:] -= 1e-7
        PG, PS = _poly(Ag), _poly(As)
        ig = any(_inside(p, PB) for p in PG) or any(_inside(p, PG) for p in PB) or any(
            _seg_intersect(PG[i], PG[(i + 1) % 4], PB[j], PB[(j + 1) % 4]) for i in range(4) for j in range(4))
        is_ = any(_inside(p, PB) for p in PS) or any(_inside(p, PS) for p in PB) or any(
            _seg_intersect(PS[i], PS[(i + 1) % 4], PB[j], PB[(j + 1) % 4]) for i in range(4) for j in range(4))
        if ig != is_:
            continue
        checked += 1
        assert O.obb2_overlap(A, B) == inter
    assert checked > 3500


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def test_obb3_vs_lp_feasibility():
    from scipy.optimize import linprog

    rng = np.random.default_rng(8)
    checked = 0
    for _ in range(600):
        RA, RB = _rot(rng), _rot(rng)
        if rng.uniform() < 0.3:
            RA = np.eye(3)  # AABB obstacle vs held box
        tA, tB = rng.uniform(-0.4, 0.4, 3), rng.uniform(-0.4, 0.4, 3)
        a, b = rng.uniform(0.02, 0.3, 3), rng.uniform(0.02, 0.3, 3)
        # max s: |R^T (x - t)| + s <= half for both boxes; overlap iff s* >= 0
        rows, rhs = [], []
        for R, t, h in ((RA, tA, a), (RB, tB, b)):
            for k in range(3):
                for sgn in (1, -1):
                    rows.append(np.concatenate([sgn * R[:, k], [1.0]]))
                    rhs.append(h[k] + sgn * R[:, k] @ t)
        res = linprog(c=[0, 0, 0, -1], A_ub=np.array(rows), b_ub=np.array(rhs), bounds=[(None, None)] * 3 + [(None, 1)])
        s = -res.fun
        if abs(s) < 1e-7:
            continue
        checked += 1
        A12 = np.concatenate([RA.reshape(-1), tA])
        B12 = np.concatenate([RB.reshape(-1), tB])
        assert O.obb3_overlap(A12, a, B12, b) == (s > 0)
    assert checked > 550


def test_sphere_box_textbook_form():
    """Arm scene of one joint (a = d = 0) whose single sphere sits at the base translation."""
    rng = np.random.default_rng(9)
    for _ in range(300):
        c = rng.uniform(-1, 1, 3)
        r = rng.uniform(0.01, 0.4)
        box = np.concatenate([rng.uniform(-0.5, 0.5, 3), rng.uniform(0.05, 0.5, 3)])
        base = np.array([[1.0, 0, 0, 0, 1, 0, 0, 0, 1, *c]])
        desc = scenes.SceneDesc(kind="arms", limits=np.array([[-3.0], [3.0]]), n_arms=1, dof=1, base=base,
                                dh=np.zeros((1, 1, 3)), link_sphere_count=np.array([1], dtype=np.int32),
                                sphere_local=np.array([[0, 0, 0, r]]), boxes=box[None, :])
        lo, hi = box[:3] - box[3:], box[:3] + box[3:]
        d2 = sum(max(lo[k] - c[k], 0.0, c[k] - hi[k]) ** 2 for k in range(3))
        if abs(d2 - r * r) < 1e-12:
            continue
        assert O.Scene(desc).config_free(np.array([rng.uniform(-3, 3)])) == (d2 > r * r)


def test_sphere_obb_vs_projection():
    rng = np.random.default_rng(10)
    for _ in range(500):
        R = _rot(rng)
        t = rng.uniform(-0.5, 0.5, 3)
        h = rng.uniform(0.02, 0.3, 3)
        c = rng.uniform(-0.8, 0.8, 3)
        r = rng.uniform(0.01, 0.3)
        local = R.T @ (c - t)
        d2 = np.sum((np.clip(local, -h, h) - local) ** 2)
        if abs(d2 - r * r) < 1e-10:
            continue
        assert O.sph_obb(c, r, np.concatenate([R.reshape(-1), t]), h) == (d2 <= r * r)


# ---------------------------------------------------------------------------------------------
def test_path_collision_checking_call_counts():
    """tests/test.py:93-122 with the registry key of abstract_env.py:381: the oracle's restatement
    of the reference's schedules (oracle/path_schedule.py) and the package's default
    is_path_collision_free make the same number of is_collision_free calls."""
    env = OracleEnv(scenes.abstract_test_scene())
    q1 = env.start_pos.from_flat(np.array([-1, 0, 1, 1.0]))
    q2 = env.start_pos.from_flat(np.array([-1, 1, 1, 0.0]))
    q3 = env.start_pos.from_flat(np.array([-1, 2, 1, -1.0]))
    s = [pkg.State(q, env.start_mode) for q in (q1, q2, q3)]
    for check in (lambda **kw: PS.sequential_path_collision_free(env, s, **kw),
                  lambda **kw: env.is_path_collision_free(s, **kw)):
        assert check(resolution=0.5)
        for kw, calls in (({"resolution": 0.5, "check_edges_in_order": True}, 5), ({"resolution": 0.5}, 5),
                          ({"resolution": 0.1, "check_edges_in_order": True}, 21), ({"resolution": 0.1}, 21)):
            with mock.patch.object(env, "is_collision_free", return_value=True) as m:
                check(**kw)
                assert m.call_count == calls, kw


def test_default_path_check_matches_schedules():
    """The package's default is_path_collision_free gives the reference schedules' boolean on random
    paths of the abstract test scene (CPU, both through the oracle)."""
    env = OracleEnv(scenes.center_rect_scene(2))
    rng = np.random.default_rng(31)
    for trial in range(150):
        K = int(rng.integers(2, 6))
        path = [pkg.State(env.start_pos.from_flat(rng.uniform(-2, 2, 4)), env.start_mode) for _ in range(K)]
        for in_order in (False, True):
            for se in (False, True):
                assert env.is_path_collision_free(path, check_edges_in_order=in_order, check_start_and_end=se) == \
                    PS.sequential_path_collision_free(env, path, check_edges_in_order=in_order,
                                                      check_start_and_end=se), (trial, in_order, se)


def test_windows_partition_the_edge():
    """Windows of 10 BFS positions (the path schedule) visit every state exactly once."""
    for N in range(1, 200):
        seen = []
        for s in range(0, N + 1, 10):
            seen += list(pkg.generate_binary_search_indices(N)[s:min(N, s + 10)])
        assert sorted(seen) == list(range(N))
        assert sum(O.edge_num_states(N, s, s + 10, True) for s in range(0, N + 1, 10)) == N


# ---------------------------------------------------------------------------------------------
def test_clearance_gradient_vs_finite_differences():
    rng = np.random.default_rng(11)
    desc = scenes.quad_arms(rng=rng)
    o = O.Scene(desc)
    h = 1e-6
    n_checked = 0
    for _ in range(6):
        q = np.clip(rng.normal(0, 0.6, desc.dim), -2.9, 2.9)
        if o.clearance_cost(q, 0.02) <= 0:
            continue
        g = o.clearance_grad(q, 0.02)
        fd = np.array([(o.clearance_cost(q + h * e, 0.02) - o.clearance_cost(q - h * e, 0.02)) / (2 * h)
                       for e in np.eye(desc.dim)])
        assert np.linalg.norm(g - fd) <= 1e-5 * max(np.linalg.norm(fd), 1e-8)
        n_checked += 1
    assert n_checked >= 3


def test_clearance_single_pair_closed_form():
    """One sphere above a box: cost = (m + r - d)^2 with d the gap to the top face."""
    base = np.array([[1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0.0, 0.0, 0.31]])
    desc = scenes.SceneDesc(kind="arms", limits=np.array([[-3.0], [3.0]]), n_arms=1, dof=1, base=base,
                            dh=np.zeros((1, 1, 3)), link_sphere_count=np.array([1], dtype=np.int32),
                            sphere_local=np.array([[0, 0, 0, 0.05]]),
                            boxes=np.array([[0.0, 0.0, 0.1, 0.2, 0.2, 0.2]]))
    o = O.Scene(desc)
    d = 0.31 - 0.3
    assert abs(o.clearance_cost(np.array([0.4]), 0.02) - (0.02 + 0.05 - d) ** 2) < 1e-15
