"""GPU parity of the engine against the CPU oracle (bit-exact booleans, bit-exact FP64 kinematics).

Every case prints its name before running so a failing GPU step names itself.
"""
import ctypes
import math

import numpy as np
import pytest

from helpers import O, backend, envs, gpu_configs, gpu_edges, near_home, pkg, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return backend()


# ---------------------------------------------------------------------------------------------
# binary64 building blocks
def test_sincos_bit_exact(B):
    _say("sincos bit-exact vs oracle")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-10, 10, 20000), rng.uniform(-3.2, 3.2, 20000),
                        np.array([0.0, -0.0, math.pi / 4, math.pi / 2, math.pi, 2.9, -2.9, 1e-300])])
    xd = torch.from_numpy(x).cuda()
    s = torch.empty_like(xd)
    c = torch.empty_like(xd)
    pkg._lib.check(pkg.lib().mrcc_sincos(ctypes.c_void_p(xd.data_ptr()), len(x), ctypes.c_void_p(s.data_ptr()),
                                         ctypes.c_void_p(c.data_ptr()), None))
    torch.cuda.synchronize()
    ref = np.array([O.sincos(v) for v in x])
    assert np.array_equal(s.cpu().numpy(), ref[:, 0])
    assert np.array_equal(c.cpu().numpy(), ref[:, 1])


def test_arm_fk_bit_exact(B):
    _say("arm FK bit-exact vs oracle (C3 scene, 2000 configs)")
    rng = np.random.default_rng(1)
    desc = scenes.quad_arms(rng=rng)
    sc = B.Scene(desc)
    o = O.Scene(desc)
    Q = scenes.uniform_configs(rng, desc, 2000)
    qd = torch.from_numpy(Q).cuda()
    out = torch.empty((2000, desc.dim, 12), dtype=torch.float64, device="cuda")
    pkg._lib.check(pkg.lib().mrcc_arm_fk(sc.handle, ctypes.c_void_p(qd.data_ptr()), 2000, desc.dim, 1,
                                         ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for r in range(0, 2000, 97):
        assert np.array_equal(got[r], o.arm_fk(Q[r]).reshape(desc.dim, 12)), r


# ---------------------------------------------------------------------------------------------
# C1: the reference's abstract environment
@pytest.fixture(scope="module")
def c1(B):
    rng = np.random.default_rng(2)
    desc = scenes.planar_disks(rng=rng)
    return desc, B.Scene(desc), O.Scene(desc), rng


def test_c1_configs_full_size(c1):
    _say("C1 10,000 configs vs oracle")
    desc, sc, o, rng = c1
    Q = scenes.uniform_configs(rng, desc, 10000)
    assert np.array_equal(gpu_configs(sc, Q), o.check_configs(Q))


@pytest.mark.parametrize("form", ["persistent", "graph", "single_launch", "per_call"])
def test_c1_edges_full_size(c1, form, B):
    _say(f"C1 10,000 edges res 0.01 vs oracle ({form})")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(3)
    q1, q2 = scenes.random_pair_edges(rng, desc, 10000)
    ref = o.check_edges(q1, q2, 0.01)
    if form == "per_call":
        got = gpu_edges(sc, q1, q2, 0.01, per_call=True)
    elif form == "single_launch":
        got = gpu_edges(sc, q1, q2, 0.01, tune=B.make_tune(persistent=0, sorted=0, use_graph=0))
    elif form == "graph":
        got = gpu_edges(sc, q1, q2, 0.01, tune=B.make_tune(use_graph=1))
    else:
        got = gpu_edges(sc, q1, q2, 0.01)
    assert np.array_equal(got, ref)


def test_c1_edges_free_endpoints(c1):
    _say("C1 edges between free configurations (reference harness variant)")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(4)
    Q = scenes.uniform_configs(rng, desc, 40000)
    F = Q[o.check_configs(Q)]
    q1, q2 = F[: len(F) // 2], F[len(F) // 2: 2 * (len(F) // 2)]
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01), o.check_edges(q1, q2, 0.01))


@pytest.mark.parametrize("incl", [False, True])
def test_c1_windows(c1, incl):
    _say(f"C1 windowed edges (N_start / N_max / explicit N), include_endpoints={incl}")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(5 + incl)
    n = 6000
    q1, q2 = scenes.random_pair_edges(rng, desc, n)
    q2 = q1 + (q2 - q1) * rng.uniform(0, 1, size=(n, 1)) ** 2
    N = np.maximum(2, (np.max(np.abs(q2 - q1), axis=1) / 0.01).astype(np.int64) + 1)
    N = np.where(rng.uniform(size=n) < 0.3, N + rng.integers(0, 5, n), N).astype(np.int32)
    S = (rng.uniform(size=n) * (N + 1)).astype(np.int32)
    S = np.where(rng.uniform(size=n) < 0.3, 0, S)
    M = np.where(rng.uniform(size=n) < 0.2, -1, S + rng.integers(0, 40, n)).astype(np.int32)
    ref = o.check_edges(q1, q2, 0.01, N=N, N_start=S, N_max=M, include_endpoints=incl)
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01, N=N, N_start=S, N_max=M, incl=incl), ref)


def test_c1_pass_count_limits(c1, B):
    _say("C1 edges at the queue's pass-count limit (16384 passes of 8 states) and beyond it")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(20)
    Q = scenes.uniform_configs(rng, desc, 4000)
    F = Q[o.check_configs(Q)]
    # short free segments walked with huge explicit N: the states stay close to free endpoints
    q1 = F[:12]
    q2 = q1 + rng.uniform(-1e-3, 1e-3, q1.shape)
    N = np.array([131074, 131082, 200000, 131073, 5, 2, 3, 131074, 70000, 150001, 131075, 64],
                 dtype=np.int32)  # 131074: exactly 16384 passes of 8 states; 131082: 16385
    ref = o.check_edges(q1, q2, 0.01, N=N)
    for tune in (B.make_tune(slots_per_pass=8, use_graph=0), B.make_tune(slots_per_pass=8, use_graph=1)):
        assert np.array_equal(gpu_edges(sc, q1, q2, 0.01, N=N, tune=tune), ref)
    # and colliding walks through an obstacle, found from a long edge's slots
    q3 = np.repeat(desc.rect_center[:1], 3, axis=0)
    q3 = np.concatenate([q3, q3], axis=1)[:, :4]
    q4 = q3 + 0.001
    N3 = np.array([131074, 200000, 40], dtype=np.int32)
    assert np.array_equal(gpu_edges(sc, q3, q4, 0.01, N=N3), o.check_edges(q3, q4, 0.01, N=N3))


def test_c1_near_contact(c1):
    _say("C1 near-contact states (FP64 recheck path)")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(7)
    n = 20000
    Q = scenes.uniform_configs(rng, desc, n)
    phi = rng.uniform(-math.pi, math.pi, n)
    R = 0.2 + rng.choice([-1e-7, -1e-12, -1e-15, 0.0, 1e-15, 1e-12, 1e-7], n)
    Q[:, 2] = Q[:, 0] + R * np.cos(phi)
    Q[:, 3] = Q[:, 1] + R * np.sin(phi)
    # robots touching rectangle faces
    k = n // 2
    rc, hs = desc.rect_center, desc.rect_size / 2
    j = rng.integers(0, len(rc), k)
    Q[:k, 0] = rc[j, 0] + hs[j, 0] + 0.1 + rng.choice([-1e-13, 0.0, 1e-13], k)
    Q[:k, 1] = rc[j, 1] + rng.uniform(-1, 1, k) * hs[j, 1]
    assert np.array_equal(gpu_configs(sc, Q), o.check_configs(Q))


def test_abstract_sphere_near_contact(B):
    """abstract.test's sphere obstacle (r 0.2 at the origin, abstract_env.py:363-378) and two disks
    of r 0.1: disk centres within 1e-15 .. 1e-7 of 0.3 from the origin (the binary64 recheck of
    the disk-sphere pair), disk pairs near 0.2 apart, masks and non-finite values."""
    _say("abstract.test disk-sphere and disk-disk near contact (binary64 rechecks), masks, NaN")
    desc = scenes.abstract_test_scene()
    sc, o = B.Scene(desc), O.Scene(desc)
    rng = np.random.default_rng(19)
    n = 20000
    Q = rng.uniform(-2, 2, (n, 4))
    eps = rng.choice([-1e-7, -1e-12, -1e-15, 0.0, 1e-15, 1e-12, 1e-7], n)
    phi = rng.uniform(-math.pi, math.pi, n)
    k = n // 2
    Q[:k, 0] = (0.3 + eps[:k]) * np.cos(phi[:k])  # robot 0 touching the sphere
    Q[:k, 1] = (0.3 + eps[:k]) * np.sin(phi[:k])
    Q[k:, 2] = Q[k:, 0] + (0.2 + eps[k:]) * np.cos(phi[k:])  # robot 1 touching robot 0
    Q[k:, 3] = Q[k:, 1] + (0.2 + eps[k:]) * np.sin(phi[k:])
    ref = o.check_configs(Q)
    assert 0.2 < ref.mean() < 0.8
    assert np.array_equal(gpu_configs(sc, Q), ref)
    for mask in (1, 2):
        assert np.array_equal(gpu_configs(sc, Q, mask=mask), o.check_configs(Q, mask=mask))
    Qn = Q[:8].copy()
    Qn[0, 0] = np.nan
    Qn[1, 3] = np.inf
    Qn[2, 1] = -np.inf
    assert np.array_equal(gpu_configs(sc, Qn), o.check_configs(Qn))
    q2 = np.clip(Q[:3000] + rng.normal(0, 0.02, (3000, 4)), -2, 2)
    assert np.array_equal(gpu_edges(sc, Q[:3000], q2, 0.01), o.check_edges(Q[:3000], q2, 0.01))
    assert np.array_equal(gpu_edges(sc, Q[:300], q2[:300], 0.01, per_call=True), o.check_edges(Q[:300], q2[:300], 0.01))


def test_c1_non_finite(c1):
    _say("C1 NaN / inf coordinates follow numpy semantics")
    desc, sc, o, _ = c1
    Q = np.array([[np.nan, 0, 1, 1], [np.inf, 0, 1, 1], [0.5, 0.5, np.nan, np.nan], [-np.inf, np.inf, 0, 0]])
    assert np.array_equal(gpu_configs(sc, Q), o.check_configs(Q))


def test_c1_masks(c1):
    _say("C1 per-robot masks")
    desc, sc, o, rng = c1
    Q = scenes.uniform_configs(rng, desc, 5000)
    for mask in (1, 2, 3):
        assert np.array_equal(gpu_configs(sc, Q, mask=mask), o.check_configs(Q, mask=mask)), mask


# ---------------------------------------------------------------------------------------------
# abstract environments through the BaseProblem API (per-call host path)
def test_abstract_test_kat(B):
    _say("abstract.test KAT edge + per-call calls")
    E = envs()
    env = pkg.get_env_by_name("abstract.test")
    q1 = env.start_pos.from_flat(np.array([-1, 0, 1, 1.0]))
    q2 = env.start_pos.from_flat(np.array([-1, 1, 1, 0.0]))
    q3 = env.start_pos.from_flat(np.array([-1, 2, 1, -1.0]))
    assert env.is_edge_collision_free(q1, q2, env.start_mode)
    s = [pkg.State(q, env.start_mode) for q in (q1, q2, q3)]
    assert env.is_path_collision_free(s, resolution=0.5)
    o = O.Scene(env.desc)
    rng = np.random.default_rng(8)
    for _ in range(200):
        a, b = rng.uniform(-2, 2, 4), rng.uniform(-2, 2, 4)
        res = float(rng.choice([0.01, 0.05, 0.3]))
        qa, qb = env.start_pos.from_flat(a), env.start_pos.from_flat(b)
        assert env.is_collision_free(qa, None) == o.config_free(a)
        assert env.is_edge_collision_free(qa, qb, None, resolution=res) == o.edge_free(a, b, res)
        N = max(2, int(np.max(np.abs(a - b)) / res) + 1)
        ns = int(rng.integers(0, N + 1))
        nm = int(ns + rng.integers(0, 12))
        assert env.is_edge_collision_free(qa, qb, None, resolution=res, N_start=ns, N_max=nm, N=N) == \
            o.edge_free(a, b, res, N=N, N_start=ns, N_max=nm)
        assert env.is_collision_free_for_robot("a1", a, None) == o.config_free(a, mask=1)


def test_abstract_batch_forms(B):
    _say("abstract.test batch forms: batch_is_collision_free, _batch_is_collision_free")
    env = pkg.get_env_by_name("abstract.test")
    o = O.Scene(env.desc)
    rng = np.random.default_rng(18)
    Q = rng.uniform(-2, 2, size=(500, 4))
    ref = o.check_configs(Q)
    assert np.array_equal(env.batch_is_collision_free(Q), ref)
    free = [env.start_pos.from_flat(q) for q in Q[ref][:40]]
    hit = [env.start_pos.from_flat(q) for q in Q[~ref][:1]]
    assert env._batch_is_collision_free(free, None) is True
    assert env._batch_is_collision_free(free + hit, None) is False
    assert env._batch_is_collision_free([], None) is True


def test_path_schedules_match_sequential(B):
    _say("batched is_path_collision_free == reference sequential schedules")
    env = pkg.get_env_by_name("abstract.center_rect_2d")
    base_cls = pkg.BaseProblem
    rng = np.random.default_rng(9)
    for trial in range(60):
        K = int(rng.integers(2, 7))
        P = [env.start_pos.from_flat(rng.uniform(-2, 2, 4)) for _ in range(K)]
        path = [pkg.State(q, env.start_mode) for q in P]
        for in_order in (False, True):
            for se in (False, True):
                got = env.is_path_collision_free(path, check_edges_in_order=in_order, check_start_and_end=se)
                ref = base_cls.is_path_collision_free(env, path, check_edges_in_order=in_order,
                                                      check_start_and_end=se)
                assert got == ref, (trial, in_order, se)


# ---------------------------------------------------------------------------------------------
# C2: planar oriented boxes
@pytest.fixture(scope="module")
def c2(B):
    rng = np.random.default_rng(10)
    desc = scenes.planar_boxes(rng=rng)
    return desc, B.Scene(desc), O.Scene(desc), rng


def test_c2_configs(c2):
    _say("C2 configs (uniform + sparse scene variants) vs oracle")
    desc, sc, o, rng = c2
    Q = scenes.uniform_configs(rng, desc, 100000)
    assert np.array_equal(gpu_configs(sc, Q), o.check_configs(Q))


def test_c2_edges(c2):
    _say("C2 lognormal edges vs oracle")
    desc, sc, o, rng = c2
    q1, q2 = scenes.lognormal_edges(rng, desc, 20000, 0.01)
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01), o.check_edges(q1, q2, 0.01))


def test_c2_sparse_scene_and_masks(B):
    _say("C2 sparse scene (5 obstacles): free edges, masks")
    rng = np.random.default_rng(11)
    desc = scenes.planar_boxes(rng=rng, n_obstacles=5)
    sc, o = B.Scene(desc), O.Scene(desc)
    Q = scenes.uniform_configs(rng, desc, 50000)
    for mask in (0xFFFFFFFF, 1, 0b0110):
        assert np.array_equal(gpu_configs(sc, Q, mask=mask), o.check_configs(Q, mask=mask))
    q1, q2 = scenes.lognormal_edges(rng, desc, 10000, 0.01)
    ref = o.check_edges(q1, q2, 0.01)
    assert ref.mean() > 0.05
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01), ref)


def test_c2_environment_per_call(B):
    _say("C2 per-call host path (MrccEnvironment): configs, edges, windows, masks")
    rng = np.random.default_rng(21)
    desc = scenes.planar_boxes(rng=rng, n_obstacles=5)
    env, o = B.MrccEnvironment(desc, collision_resolution=0.01), O.Scene(desc)
    Q = scenes.uniform_configs(rng, desc, 200)
    for q in Q:
        assert env.is_collision_free(q, None) == o.config_free(q)
        assert env.is_collision_free_for_robot(["a2", "a3"], q, None) == o.config_free(q, mask=0b0110)
    q1, q2 = scenes.lognormal_edges(rng, desc, 200, 0.01)
    n_free = 0
    for i in range(len(q1)):
        got = env.is_edge_collision_free(q1[i], q2[i], None)
        assert got == o.edge_free(q1[i], q2[i], 0.01), i
        n_free += got
        if i % 3 == 0:
            N = pkg.edge_num_states(q1[i], q2[i], 0.01)
            ns = int(rng.integers(0, N + 1))
            nm = ns + int(rng.integers(0, 12))
            assert env.is_edge_collision_free(q1[i], q2[i], None, N_start=ns, N_max=nm) == \
                o.edge_free(q1[i], q2[i], 0.01, N=N, N_start=ns, N_max=nm), i
    assert 10 <= n_free <= len(q1) - 10


def _lane_scene(kind, rng):
    if kind == "disk_many_spheres":  # 2 x 3000 binary32 sphere floats + thresholds: beyond 32 KiB
        c = rng.uniform(-2, 2, (3000, 2))
        lim = np.array([[-2.0] * 4, [2.0] * 4])
        return scenes.disks_scene([0.02, 0.02], 2, lim, spheres=[(c[i], 0.005) for i in range(3000)])
    if kind == "disk2_1500":  # 6,000 floats of sphere tables: staged, through the tail loop
        c = rng.uniform(-2, 2, (1500, 2))
        lim = np.array([[-2.0] * 4, [2.0] * 4])
        return scenes.disks_scene([0.03, 0.03], 2, lim, spheres=[(c[i], 0.01) for i in range(1500)])
    if kind == "disk3d_300":  # generic evaluator, 1,800 floats of sphere tables at 256 threads
        lim = np.array([[-2.0] * 9, [2.0] * 9])
        sph = [(rng.uniform(-2, 2, 3), rng.uniform(0.02, 0.08)) for _ in range(300)]
        return scenes.disks_scene([0.1, 0.15, 0.2], 3, lim, spheres=sph)
    if kind == "disk3d":  # 3 robots in 3-D with spheres and boxes (the generic disk evaluator)
        lim = np.array([[-2.0] * 9, [2.0] * 9])
        sph = [(rng.uniform(-2, 2, 3), rng.uniform(0.1, 0.4)) for _ in range(12)]
        rec = [(rng.uniform(-2, 2, 3), rng.uniform(0.2, 0.8, 3)) for _ in range(10)]
        return scenes.disks_scene([0.1, 0.15, 0.2], 3, lim, spheres=sph, rects=rec)
    if kind == "disk_c1":
        return scenes.planar_disks(rng=rng)
    nr, no = {"planar4_100": (4, 100), "planar4_300": (4, 300), "planar3_100": (3, 100),
              "planar4_700": (4, 700), "planar3_300": (3, 300), "planar4_1100": (4, 1100)}[kind]
    return scenes.planar_boxes(rng=rng, n_robots=nr, n_obstacles=no)


@pytest.mark.parametrize("kind", ["planar4_100", "planar4_300", "planar3_100", "planar4_700", "planar3_300",
                                  "planar4_1100", "disk_c1", "disk3d", "disk2_1500", "disk3d_300",
                                  "disk_many_spheres"])
def test_lane_per_call_scene_sizes(B, kind):
    """The per-call lane kernel stages the binary32 tables in LDS up to 8192 floats (planar boxes up
    to 1019 obstacles) and reads global memory beyond; four loads per thread are issued first and
    longer tables finish in a tail loop (planar4_700 at 1024 threads, planar3_300 at 256); planar
    edges split a state's obstacles over up to 64 lanes when the edge has few states.  Edge
    lengths from 1 to ~300 states, masks, against the oracle."""
    rng = np.random.default_rng(sum(map(ord, kind)))
    desc = _lane_scene(kind, rng)
    sc, o = B.Scene(desc), O.Scene(desc)
    lo, hi = desc.limits
    n = 400
    q1 = rng.uniform(lo, hi, (n, desc.dim))
    step = np.exp(rng.uniform(np.log(0.003), np.log(3.0), (n, 1)))
    d = rng.normal(size=(n, desc.dim))
    d /= np.max(np.abs(d), axis=1, keepdims=True)
    q2 = np.clip(q1 + step * d, lo, hi)
    ref = o.check_edges(q1, q2, 0.01)
    _say(f"{kind}: per-call edges vs oracle ({ref.mean():.3f} free)")
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01, per_call=True), ref)
    m = 0b0101
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01, mask=m, per_call=True), o.check_edges(q1, q2, 0.01, mask=m))
    # the one-configuration host call runs the same kernel on the row itself (no interpolation):
    # free and colliding configurations, -0.0 coordinates, non-finite values, masks
    Q = np.concatenate([q1[:150], 0.5 * (q1[:150] + q2[:150])])
    Q[::7, 0] = -0.0
    Q[3, 1] = np.nan
    Q[5, -1] = np.inf
    res = ctypes.c_int32(-1)
    for mask in (0xFFFFFFFF, m):
        got = []
        for q in Q:
            qc = np.ascontiguousarray(q)
            pkg._lib.check(pkg.lib().mrcc_check_config_host(sc.handle, qc.ctypes.data, 0, mask, ctypes.byref(res)))
            got.append(bool(res.value))
        assert np.array_equal(np.array(got), o.check_configs(Q, mask=mask)), mask


# ---------------------------------------------------------------------------------------------
# C3: 4 x 7-DoF arms
@pytest.fixture(scope="module")
def c3(B):
    rng = np.random.default_rng(12)
    desc = scenes.quad_arms(rng=rng)
    return desc, B.Scene(desc), O.Scene(desc), rng


@pytest.mark.parametrize("bp", [True, False], ids=["broadphase", "bruteforce"])
def test_c3_configs(c3, bp):
    _say(f"C3 configs: uniform + near-home vs oracle (broadphase={bp})")
    desc, sc, o, rng = c3
    sc.set_broadphase(bp)
    Q = np.concatenate([scenes.uniform_configs(rng, desc, 4000), near_home(rng, desc, 8000)])
    ref = o.check_configs(Q)
    assert 0.005 < ref.mean() < 0.95  # both outcomes represented (>= 60 free configurations)
    assert np.array_equal(gpu_configs(sc, Q), ref)
    sc.set_broadphase(True)


@pytest.mark.parametrize("bp", [True, False], ids=["broadphase", "bruteforce"])
def test_c3_edges(c3, B, bp):
    _say(f"C3 edges near home + lognormal uniform vs oracle (persistent, single launch, per call; broadphase={bp})")
    desc, sc, o, rng = c3
    sc.set_broadphase(bp)
    a = near_home(rng, desc, 3000)
    b = np.clip(a + rng.normal(0, 0.3, a.shape), desc.limits[0], desc.limits[1])
    u1, u2 = scenes.lognormal_edges(rng, desc, 1000, 0.05)
    q1, q2 = np.concatenate([a, u1]), np.concatenate([b, u2])
    ref = o.check_edges(q1, q2, 0.05)
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.05), ref)
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.05, tune=B.make_tune(persistent=0, sorted=0, use_graph=0)), ref)
    assert np.array_equal(gpu_edges(sc, q1[:300], q2[:300], 0.05, per_call=True), ref[:300])
    sc.set_broadphase(True)


def test_c3_per_call_state_split(c3, B):
    """The per-call kernel splits a state's narrow phase over up to 4 waves when the edge has fewer
    states than waves, and waves stop once one finds a collision: short edges (1..10 states), free
    and colliding, against the oracle; with held boxes (modes) and on chain joints."""
    _say("C3 per call: state split over waves (1..10 states per edge), modes, chain joints vs oracle")
    desc, sc, o, rng = c3
    a = near_home(rng, desc, 2400, sigma=0.5)
    d = rng.normal(0, 1, a.shape)
    d /= np.max(np.abs(d), axis=1, keepdims=True)
    L = 0.05 * rng.integers(1, 12, size=(len(a), 1)) + rng.uniform(0, 0.049, (len(a), 1))  # N - 2 = 0..10
    b = np.clip(a + L * d, desc.limits[0], desc.limits[1])
    ref = o.check_edges(a, b, 0.05)
    assert 0.1 < ref.mean() < 0.9
    assert np.array_equal(gpu_edges(sc, a, b, 0.05, per_call=True), ref)
    # held boxes and fewer arms (the GRASP instantiation)
    md = scenes.mixed_arms(rng=np.random.default_rng(23), n_arms=4, n_modes=16)
    sm, om = B.Scene(md), O.Scene(md)
    a2 = near_home(rng, md, 1200, sigma=0.5)
    b2 = np.clip(a2 + rng.normal(0, 0.08, a2.shape), md.limits[0], md.limits[1])
    modes = rng.integers(0, 16, len(a2)).astype(np.int32)
    ref2 = om.check_edges(a2, b2, 0.05, modes=modes)
    assert 0.05 < ref2.mean() < 0.95
    assert np.array_equal(gpu_edges(sm, a2, b2, 0.05, modes=modes, per_call=True), ref2)
    # chain joints (prismatic among them): the brute-force scene family
    import arm_bruteforce as BF
    cd = BF.tiny_arm_scene(np.random.default_rng(33), scenes, chain=True, prism=0.4, n_boxes=2, spread=0.5)
    sc3, o3 = B.Scene(cd), O.Scene(cd)
    a3 = rng.uniform(cd.limits[0], cd.limits[1], size=(1200, cd.dim)) * 0.5
    b3 = np.clip(a3 + rng.normal(0, 0.1, a3.shape), cd.limits[0], cd.limits[1])
    m3 = rng.integers(0, len(cd.modes["n_active"]), len(a3)).astype(np.int32)
    ref3 = o3.check_edges(a3, b3, 0.05, modes=m3)
    assert 0.05 < ref3.mean() < 0.95
    assert np.array_equal(gpu_edges(sc3, a3, b3, 0.05, modes=m3, per_call=True), ref3)


@pytest.mark.parametrize("spt,g,w", [(1, 1, 1), (16, 1, 12), (64, 4, 16), (7, 2, 3)])
def test_c3_queue_shapes(c3, B, spt, g, w):
    _say(f"C3 persistent queue shapes: slots/ticket {spt}, slots/pass {g}, waves/CU {w}; stats invariants")
    desc, sc, o, rng = c3
    a = near_home(rng, desc, 1500, sigma=0.7)
    b = np.clip(a + rng.normal(0, 0.4, a.shape), desc.limits[0], desc.limits[1])
    # plus uniform lognormal edges: almost all collide at their first state, so the later passes
    # hold chunks without live states and waves take several tickets at once
    u1, u2 = scenes.lognormal_edges(rng, desc, 1000, 0.05)
    a, b = np.concatenate([a, u1]), np.concatenate([b, u2])
    ref = o.check_edges(a, b, 0.05)
    t = B.make_tune(slots_per_pass=g, slots_per_ticket=spt, waves_per_cu=w, use_graph=0, collect_stats=1)
    assert np.array_equal(gpu_edges(sc, a, b, 0.05, tune=t), ref)


def test_c3_windows(c3):
    _say("C3 windowed edges (N_start / N_max / explicit N) through the arm queue")
    desc, sc, o, rng = c3
    n = 1500
    a = near_home(rng, desc, n, sigma=0.6)
    b = np.clip(a + rng.normal(0, 0.3, a.shape), desc.limits[0], desc.limits[1])
    N = np.maximum(2, (np.max(np.abs(b - a), axis=1) / 0.05).astype(np.int64) + 1)
    N = np.where(rng.uniform(size=n) < 0.3, N + rng.integers(0, 5, n), N).astype(np.int32)
    S = (rng.uniform(size=n) * (N + 1)).astype(np.int32)
    S = np.where(rng.uniform(size=n) < 0.3, 0, S)
    M = np.where(rng.uniform(size=n) < 0.2, -1, S + rng.integers(0, 10, n)).astype(np.int32)
    for incl in (False, True):
        ref = o.check_edges(a, b, 0.05, N=N, N_start=S, N_max=M, include_endpoints=incl)
        assert np.array_equal(gpu_edges(sc, a, b, 0.05, N=N, N_start=S, N_max=M, incl=incl), ref), incl


def test_plan_timing_and_stats(c3, B):
    _say("plan timing events (time_kernels) and stats invariants; arm slots_per_ticket limit")
    desc, sc, o, rng = c3
    a = near_home(rng, desc, 2000, sigma=0.7)
    b = np.clip(a + rng.normal(0, 0.4, a.shape), desc.limits[0], desc.limits[1])
    ref = o.check_edges(a, b, 0.05)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.empty(len(a), dtype=torch.uint8, device="cuda")
    batch = sc.edge_batch(ta, tb, out, 0.05)
    G = 2
    plan = B.EdgePlan(sc, batch, B.make_tune(slots_per_pass=G, use_graph=0, time_kernels=1, collect_stats=1))
    plan.run()
    t = plan.timing()
    assert 0.0 < t["kernel_ms"] <= t["run_ms"] and 0.0 <= t["build_ms"] <= t["run_ms"]
    assert np.array_equal(out.cpu().numpy().astype(bool), ref)
    st = plan.stats()
    N = np.maximum(2, (np.max(np.abs(b - a), axis=1) / 0.05).astype(np.int64) + 1)
    # an edge of M = N - 2 states takes ceil(M / G) passes of G slots each
    assert st["slots"] == int(np.sum(G * ((N - 2 + G - 1) // G)))
    assert st["states"] + st["dead_skips"] <= st["slots"]
    # every state of a free edge is evaluated; a colliding edge needs at least one
    assert st["states"] >= int(np.sum(N[ref] - 2)) + int(np.sum(~ref))
    gplan = B.EdgePlan(sc, batch, B.make_tune(use_graph=1))
    with pytest.raises(pkg.MrccError):
        gplan.timing()
    with pytest.raises(pkg.MrccError):
        B.EdgePlan(sc, batch, B.make_tune(slots_per_ticket=65))


def test_c3_non_finite_edges(c3, B):
    _say("C3 edges with NaN / inf away from coordinate 0 (N rule skips them; states collide)")
    desc, sc, o, rng = c3
    Q = near_home(rng, desc, 4000, sigma=0.4)
    a = Q[o.check_configs(Q)]
    b = np.clip(a + rng.normal(0, 0.05, a.shape), desc.limits[0], desc.limits[1])
    keep = o.check_edges(a, b, 0.05)
    a, b = a[keep][:60], b[keep][:60]  # free edges: the non-finite states decide the outcome
    assert len(a) >= 12
    a[::3, 5] = np.nan
    b[1::3, 17] = np.nan
    ref = o.check_edges(a, b, 0.05)
    assert np.array_equal(gpu_edges(sc, a, b, 0.05), ref)
    # a NaN in coordinate 0, or an infinite length, makes N unbounded: the batch is rejected, as
    # int(nan) / int(inf) raise in the reference's N rule (abstract_env.py:321-323)
    for k, v in ((0, np.nan), (9, np.inf)):
        a2 = a.copy()
        a2[2, k] = v
        with pytest.raises(pkg.MrccError):
            gpu_edges(sc, a2, b, 0.05)


def test_c3_environment_per_call(c3, B):
    _say("C3 per-call host path (MrccEnvironment): configs, edges, windows, masks, non-finite lengths")
    desc, sc, o, rng = c3
    env = B.MrccEnvironment(desc, collision_resolution=0.05)
    Q = np.concatenate([near_home(rng, desc, 150, sigma=0.4), scenes.uniform_configs(rng, desc, 50)])
    for q in Q:
        assert env.is_collision_free(q, None) == o.config_free(q)
    assert env.is_collision_free_for_robot(["a1", "a3"], Q[0], None) == o.config_free(Q[0], mask=0b0101)
    # edges from free starts (short steps: about half of them free) and from random starts
    s = near_home(rng, desc, 3000, sigma=0.3)
    a = np.concatenate([s[o.check_configs(s)][:60], near_home(rng, desc, 140, sigma=0.4)])
    b = np.clip(a + rng.normal(0, 0.1, a.shape), desc.limits[0], desc.limits[1])
    n_free = 0
    for i in range(len(a)):
        got = env.is_edge_collision_free(a[i], b[i], None)
        assert got == o.edge_free(a[i], b[i], 0.05), i
        n_free += got
        if i % 4 == 0:  # windows of the binary order (the path schedule), endpoints included
            N = pkg.edge_num_states(a[i], b[i], 0.05)
            ns = int(rng.integers(0, N + 1))
            nm = ns + int(rng.integers(0, 12))
            assert env.is_edge_collision_free(a[i], b[i], None, N_start=ns, N_max=nm, include_endpoints=True) == \
                o.edge_free(a[i], b[i], 0.05, N=N, N_start=ns, N_max=nm, include_endpoints=True), i
    assert 10 <= n_free <= len(a) - 10  # both outcomes occur
    # the reference's N rule: a NaN away from coordinate 0 is skipped by the distance (the states
    # holding it collide); in coordinate 0 int(nan) raises ValueError, an infinite length
    # OverflowError (abstract_env.py:321-323)
    a2 = a[0].copy()
    a2[5] = np.nan
    assert env.is_edge_collision_free(a2, b[0], None) == o.edge_free(a2, b[0], 0.05)
    a2 = a[0].copy()
    a2[0] = np.nan
    with pytest.raises(ValueError):
        env.is_edge_collision_free(a2, b[0], None)
    a2 = a[0].copy()
    a2[9] = np.inf
    with pytest.raises(OverflowError):
        env.is_edge_collision_free(a2, b[0], None)
    # the staging block is reused call after call: a colliding and a free call alternate
    free_i = next(i for i in range(len(a)) if o.edge_free(a[i], b[i], 0.05))
    hit_i = next(i for i in range(len(a)) if not o.edge_free(a[i], b[i], 0.05))
    for _ in range(20):
        assert env.is_edge_collision_free(a[free_i], b[free_i], None)
        assert not env.is_edge_collision_free(a[hit_i], b[hit_i], None)


def test_c3_environment_long_edges(c3, B):
    """The one-edge host call spreads an arm edge of more than two rounds of states over several
    workgroups (verdict preset to free, a flag kernel after the check); shorter edges and windows
    keep the one-workgroup kernel.  Edges of 3 to ~150 states from free starts, windows on both
    sides of the switch, a robot mask, and calls of both kinds alternating on the staging block."""
    _say("C3 per-call host path, long edges (several workgroups) vs oracle")
    desc, sc, o, rng = c3
    env = B.MrccEnvironment(desc, collision_resolution=0.05)
    s = near_home(rng, desc, 4000, sigma=0.3)
    a = np.concatenate([s[o.check_configs(s)][:60], near_home(rng, desc, 60, sigma=0.45)])
    d = rng.normal(size=a.shape)
    d[:60] = 0.0  # from the free starts: one joint moves (long edges that stay free occur)
    d[np.arange(60), rng.integers(0, desc.dim, 60)] = rng.choice([-1.0, 1.0], 60)
    d /= np.max(np.abs(d), axis=1, keepdims=True)
    states = np.exp(rng.uniform(np.log(2.0), np.log(150.0), (len(a), 1)))  # 3 .. 150 states
    b = np.clip(a + states * 0.05 * d, desc.limits[0], desc.limits[1])
    got = np.array([env.is_edge_collision_free(a[i], b[i], None) for i in range(len(a))])
    ref = o.check_edges(a, b, 0.05)
    assert np.array_equal(got, ref)
    assert 10 <= ref.sum() <= len(a) - 10
    N = np.array([pkg.edge_num_states(a[i], b[i], 0.05) for i in range(len(a))])
    assert (N > 20).sum() > 20 and (N < 8).sum() > 5
    for i in range(0, len(a), 3):
        ns = int(rng.integers(0, N[i] + 1))
        nm = ns + int(rng.choice([3, 8, 9, 40, 200]))
        for incl in (False, True):
            assert env.is_edge_collision_free(a[i], b[i], None, N_start=ns, N_max=nm, include_endpoints=incl) == \
                o.edge_free(a[i], b[i], 0.05, N=int(N[i]), N_start=ns, N_max=nm, include_endpoints=incl), (i, ns, nm)
    res = ctypes.c_int32(-1)
    for i in range(0, len(a), 5):  # robot mask through the C entry point
        ai, bi = np.ascontiguousarray(a[i]), np.ascontiguousarray(b[i])
        pkg._lib.check(pkg.lib().mrcc_check_edge_host(sc.handle, ai.ctypes.data, bi.ctypes.data, ctypes.c_double(0.05),
                                                      0, 0, -1, 0, 0, 0b0101, ctypes.byref(res)))
        assert bool(res.value) == o.edge_free(a[i], b[i], 0.05, mask=0b0101), i
    long_free = [i for i in range(len(a)) if ref[i] and N[i] > 20]
    long_hit = [i for i in range(len(a)) if not ref[i] and N[i] > 20]
    short = [i for i in range(len(a)) if N[i] < 8]
    assert long_free and long_hit and short
    for _ in range(10):
        for i in (long_free[0], long_hit[0], short[0]):
            assert env.is_edge_collision_free(a[i], b[i], None) == ref[i]


def test_c3_masks(c3):
    _say("C3 per-arm masks")
    desc, sc, o, rng = c3
    Q = near_home(rng, desc, 3000, sigma=0.8)
    for mask in (1, 0b0101, 0b1110):
        assert np.array_equal(gpu_configs(sc, Q, mask=mask), o.check_configs(Q, mask=mask)), mask


@pytest.mark.parametrize("bp", [True, False], ids=["broadphase", "bruteforce"])
def test_c3_near_contact(c3, bp):
    _say(f"C3 near-contact: configs bisected onto the contact boundary (broadphase={bp})")
    desc, sc, o, rng = c3
    sc.set_broadphase(bp)
    A = near_home(rng, desc, 400)
    Bq = near_home(rng, desc, 400, sigma=1.2)
    fa, fb = o.check_configs(A), o.check_configs(Bq)
    pairs = [(a, b) for a, b, x, y in zip(A, Bq, fa, fb) if x != y][:120]
    pts = []
    for a, b in pairs:
        lo, hi = 0.0, 1.0  # free at lo side iff o.config_free(a)
        fa_ = o.config_free(a)
        for _ in range(48):
            m = 0.5 * (lo + hi)
            if o.config_free(a + m * (b - a)) == fa_:
                lo = m
            else:
                hi = m
        for t in (lo, hi):
            pts.append(a + t * (b - a))
    P = np.array(pts)
    assert np.array_equal(gpu_configs(sc, P), o.check_configs(P))
    sc.set_broadphase(True)


# ---------------------------------------------------------------------------------------------
# C5: mixed arm counts and held boxes
@pytest.mark.parametrize("bp", [True, False], ids=["broadphase", "bruteforce"])
def test_c5_modes(B, bp):
    _say(f"C5 mixed modes: configs and edges with mode ids vs oracle (broadphase={bp})")
    rng = np.random.default_rng(13)
    desc = scenes.mixed_arms(rng=rng, n_modes=256)
    sc, o = B.Scene(desc), O.Scene(desc)
    sc.set_broadphase(bp)
    n = 6000
    Q = near_home(rng, desc, n, sigma=0.5)
    modes = rng.integers(0, 256, n).astype(np.int32)
    ref = o.check_configs(Q, modes=modes)
    assert 0.02 < ref.mean() < 0.98
    assert np.array_equal(gpu_configs(sc, Q, modes=modes), ref)
    for mask in (0b1, 0b1010_0101):
        assert np.array_equal(gpu_configs(sc, Q[:2000], modes=modes[:2000], mask=mask),
                              o.check_configs(Q[:2000], modes=modes[:2000], mask=mask))
    q2 = np.clip(Q + rng.normal(0, 0.2, Q.shape), desc.limits[0], desc.limits[1])
    refe = o.check_edges(Q[:3000], q2[:3000], 0.05, modes=modes[:3000])
    assert np.array_equal(gpu_edges(sc, Q[:3000], q2[:3000], 0.05, modes=modes[:3000]), refe)
    # more modes than the LDS table holds: the per-mode arm counts are read from global memory
    desc2 = scenes.mixed_arms(rng=np.random.default_rng(19), n_modes=3000)
    sc2, o2 = B.Scene(desc2), O.Scene(desc2)
    sc2.set_broadphase(bp)
    m2 = rng.integers(0, 3000, 3000).astype(np.int32)
    Q2 = near_home(rng, desc2, 3000, sigma=0.5)
    assert np.array_equal(gpu_configs(sc2, Q2, modes=m2), o2.check_configs(Q2, modes=m2))


# ---------------------------------------------------------------------------------------------
# plans, graphs, errors
def test_graph_replay_stable(c1, B):
    _say("graph-captured plan replayed 4x with eager work in between")
    desc, sc, o, _ = c1
    rng = np.random.default_rng(14)
    q1, q2 = scenes.random_pair_edges(rng, desc, 5000)
    ref = o.check_edges(q1, q2, 0.01)
    a, b = torch.from_numpy(q1).cuda(), torch.from_numpy(q2).cuda()
    out = torch.empty(5000, dtype=torch.uint8, device="cuda")
    plan = B.EdgePlan(sc, sc.edge_batch(a, b, out, 0.01), B.make_tune(use_graph=1, collect_stats=1))
    for k in range(4):
        out.fill_(9)
        plan.run()
        torch.cuda.synchronize()
        _ = (a * 2).sum().item()
        assert np.array_equal(out.cpu().numpy().astype(bool), ref), k
    st = plan.stats()
    assert st["states"] > 0


def test_empty_and_errors(c1, B):
    _say("empty batches, N_start > N error, injected failure")
    desc, sc, o, _ = c1
    assert gpu_edges(sc, np.zeros((0, 4)), np.zeros((0, 4)), 0.01).shape == (0,)
    q1 = np.zeros((2, 4))
    q2 = np.ones((2, 4))
    with pytest.raises(pkg.MrccError):
        gpu_edges(sc, q1, q2, 0.01, N_start=np.array([5, 1000], dtype=np.int32))
    L = pkg.lib()
    L.mrcc_debug_inject_failure(-5, 1)
    with pytest.raises(pkg.MrccError):
        gpu_edges(sc, q1, q2, 0.01)
    assert gpu_edges(sc, q1, q2, 0.01).shape == (2,)


# ---------------------------------------------------------------------------------------------
# differentiable clearance (C4)
@pytest.mark.parametrize("bp", [True, False], ids=["culled", "full_pair_set"])
def test_clearance_cost_and_grad(B, bp):
    _say(f"clearance cost vs FP64 oracle; gradient vs FP64 central FD (link cull {bp})")
    rng = np.random.default_rng(15)
    desc = scenes.quad_arms(rng=rng)
    sc, o = B.Scene(desc), O.Scene(desc)
    sc.set_broadphase(bp)
    Q = near_home(rng, desc, 64, sigma=0.6)
    qd = torch.from_numpy(Q.astype(np.float32)).cuda()
    cost = torch.empty(64, dtype=torch.float32, device="cuda")
    L = pkg.lib()
    pkg._lib.check(L.mrcc_clearance_forward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 64, desc.dim, ctypes.c_float(0.02),
                                            ctypes.c_void_p(cost.data_ptr()), None))
    g_out = torch.ones(64, dtype=torch.float32, device="cuda")
    grad = torch.empty_like(qd)
    pkg._lib.check(L.mrcc_clearance_backward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 64, desc.dim, ctypes.c_float(0.02),
                                             ctypes.c_void_p(g_out.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                             None))
    torch.cuda.synchronize()
    c = cost.cpu().numpy()
    g = grad.cpu().numpy()
    Q32 = Q.astype(np.float32).astype(np.float64)
    ref_c = np.array([o.clearance_cost(q, 0.02) for q in Q32])
    assert np.all(np.abs(c - ref_c) <= 1e-4 * np.maximum(ref_c, 1e-3))
    h = 1e-6
    errs = []
    for r in range(16):
        q = Q32[r]
        if ref_c[r] <= 0:
            continue
        fd = np.zeros(desc.dim)
        for k in range(desc.dim):
            e = np.zeros(desc.dim)
            e[k] = h
            fd[k] = (o.clearance_cost(q + e, 0.02) - o.clearance_cost(q - e, 0.02)) / (2 * h)
        an = o.clearance_grad(q, 0.02)
        assert np.linalg.norm(an - fd) <= 1e-5 * max(np.linalg.norm(fd), 1e-6)
        errs.append(np.linalg.norm(g[r] - fd) / max(np.linalg.norm(fd), 1e-6))
    assert errs and max(errs) <= 1e-4, errs
