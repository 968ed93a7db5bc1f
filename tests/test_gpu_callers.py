"""Batched call-site helpers (callers.py) against the reference's per-call loops, restated here
from rrtstar_base.py:1339-1446, prm_graph.py:663-684 and planner_eitstar.py:630-665 with the
environment's per-call ``is_edge_collision_free``."""
import numpy as np
import pytest

from helpers import envs, near_home, pkg

pytestmark = pytest.mark.gpu


def _env():
    envs()
    return pkg.get_env_by_name("mrcc.quad_arms")


def _cfgs(env, rows):
    return [env.start_pos.from_flat(np.ascontiguousarray(r, dtype=np.float64)) for r in rows]


def _cluster(env, rng, k, spread):
    centre = near_home(rng, env.desc, 1)[0]
    rows = np.clip(centre + rng.normal(0.0, spread, size=(k, env.desc.dim)), env.limits[0], env.limits[1])
    return centre, rows


def test_find_parent_matches_loop():
    callers = pkg.callers
    env = _env()
    rng = np.random.default_rng(0)
    decided_by_check = 0
    for trial in range(60):
        centre, rows = _cluster(env, rng, 24, 0.35)
        q_new = _cfgs(env, [centre])[0]
        near = _cfgs(env, rows)
        ids = rng.permutation(1000)[:24]
        nearest = ids[rng.integers(24)]
        batch_cost = env.batch_config_cost(q_new, rows)
        n_near_costs = rng.uniform(0, 3, 24)
        got = callers.find_parent_batched(env, env.start_mode, near, q_new, ids, nearest, batch_cost, n_near_costs)
        # the reference loop (rrtstar_base.py:1363-1381)
        idx = int(np.where(ids == nearest)[0][0])
        c = n_near_costs + batch_cost
        c_min, c_par, chosen = c[idx], batch_cost[idx], idx
        mask = c < c_min
        if np.any(mask):
            for i in np.where(mask)[0][np.argsort(c[mask])]:
                if env.is_edge_collision_free(near[i], q_new, env.start_mode):
                    c_min, c_par, chosen = c[i], batch_cost[i], int(i)
                    decided_by_check += chosen != idx
                    break
        assert got == (chosen, float(c_min), float(c_par)), trial
    assert decided_by_check > 5  # the check changed the choice in a fair share of the trials


def test_rewire_matches_loop():
    callers = pkg.callers
    env = _env()
    rng = np.random.default_rng(1)
    n_free = n_coll = 0
    for trial in range(40):
        centre, rows = _cluster(env, rng, 32, 0.35)
        q_new = _cfgs(env, [centre])[0]
        near = _cfgs(env, rows)
        eligible = rng.random(32) < 0.7
        got = callers.rewire_batched(env, env.start_mode, q_new, near, eligible)
        for i in range(32):
            exp = bool(eligible[i]) and env.is_edge_collision_free(q_new, near[i], env.start_mode)
            assert got[i] == exp, (trial, i)
            n_free += exp
            n_coll += bool(eligible[i]) and not exp
    assert n_free > 50 and n_coll > 50


def test_prevalidate_matches_loop():
    """The loop (prm_graph.py:663-684) reaches the queued edges in order; the first direction of
    a pair it reaches is checked and cached for both directions."""
    callers = pkg.callers
    env = _env()
    rng = np.random.default_rng(2)
    centre, rows = _cluster(env, rng, 60, 0.25)
    qs = dict(enumerate(_cfgs(env, rows)))
    modes = {i: env.start_mode for i in qs}
    edges = [(int(a), int(b)) for a, b in rng.integers(0, 60, size=(300, 2)) if a != b]
    wl, bl = {}, {}
    assert callers.prevalidate_edges(env, edges, qs, modes, None, wl, bl) > 0
    lw, lb = {}, {}  # the loop's caches
    n_free = n_coll = 0
    for a, b in edges:
        if a in lw.get(b, ()) or b in lb.get(a, ()):
            continue
        free = env.is_edge_collision_free(qs[a], qs[b], env.start_mode)
        cache = lw if free else lb
        cache.setdefault(a, set()).add(b)
        cache.setdefault(b, set()).add(a)
        n_free += free
        n_coll += not free
    assert wl == lw and bl == lb
    assert n_free >= 5 and n_coll >= 5
    assert callers.prevalidate_edges(env, edges, qs, modes, None, wl, bl) == 0  # all cached


def test_sparse_checks_match_loop():
    callers = pkg.callers
    env = _env()
    rng = np.random.default_rng(4)
    centre, rows = _cluster(env, rng, 200, 0.6)
    starts, ends = _cfgs(env, rows[:100]), _cfgs(env, rows[100:])
    Ns = [max(2, int(np.max(np.abs(a.state() - b.state())) / 0.05) + 1) for a, b in zip(starts, ends)]
    modes = [env.start_mode] * 100
    for N_start, N_max in ((0, 10), (10, 20), (20, 30), (0, 1000)):
        got = callers.sparse_checks_batched(env, starts, ends, modes, N_start, N_max, Ns)
        for i in range(100):
            if N_start > Ns[i]:
                assert got[i]
                continue
            exp = env.is_edge_collision_free(starts[i], ends[i], env.start_mode, N_start=N_start, N_max=N_max, N=Ns[i])
            assert got[i] == exp, (N_start, i)


# ---------------------------------------------------------------------------------------------
# mrcc_check_edges_host (short lists on host memory) against the oracle
def _host_edges(sc, q1, q2, res, N=None, N_start=None, N_max=None, incl=False, modes=None, mask=0xFFFFFFFF):
    L = pkg.lib()
    n = len(q1)
    q1 = np.ascontiguousarray(q1, dtype=np.float64)
    q2 = np.ascontiguousarray(q2, dtype=np.float64)
    iv = lambda v: None if v is None else np.ascontiguousarray(v, dtype=np.int32)
    arrs = [iv(N), iv(N_start), iv(N_max), iv(modes)]
    ptr = lambda v: None if v is None else v.ctypes.data
    out = np.full(n, 7, dtype=np.uint8)
    rc = L.mrcc_check_edges_host(sc.handle, q1.ctypes.data, q2.ctypes.data, n, float(res),
                                 ptr(arrs[0]), ptr(arrs[1]), ptr(arrs[2]), int(incl), ptr(arrs[3]), mask,
                                 out.ctypes.data)
    return rc, out


@pytest.mark.parametrize("n", [1, 7, 30, 500, 3000])
def test_host_list_c3_matches_oracle(n):
    from helpers import O, backend, scenes
    rng = np.random.default_rng(n)
    desc = scenes.quad_arms(rng=rng)
    sc = backend().Scene(desc)
    o = O.Scene(desc)
    base = near_home(rng, desc, n)
    q1 = base
    q2 = np.clip(base + rng.normal(0, 0.4, base.shape), desc.limits[0], desc.limits[1])
    print(f"[case] c3 host list n={n}", flush=True)
    rc, got = _host_edges(sc, q1, q2, 0.05)
    assert rc == 0 and set(np.unique(got)) <= {0, 1}
    exp = o.check_edges(q1, q2, 0.05)
    assert np.array_equal(got.astype(bool), exp)
    # windows of the binary order with explicit N, endpoints, and a robot mask
    N = np.maximum(2, (np.max(np.abs(q1 - q2), axis=1) / 0.05).astype(np.int64) + 1)
    S = np.minimum(N, rng.integers(0, 12, n))
    M = np.where(rng.random(n) < 0.3, -1, S + rng.integers(0, 15, n))
    for incl in (False, True):
        rc, got = _host_edges(sc, q1, q2, 0.05, N=N, N_start=S, N_max=M, incl=incl, mask=0b0101)
        exp = o.check_edges(q1, q2, 0.05, N=N, N_start=S, N_max=M, include_endpoints=incl, mask=0b0101)
        assert rc == 0 and np.array_equal(got.astype(bool), exp), incl


def test_host_list_c1_and_c5_match_oracle():
    from helpers import O, backend, scenes
    rng = np.random.default_rng(9)
    desc = scenes.planar_disks(rng=rng)
    sc, o = backend().Scene(desc), O.Scene(desc)
    q1, q2 = scenes.random_pair_edges(rng, desc, 400)
    rc, got = _host_edges(sc, q1, q2, 0.01)
    assert rc == 0 and np.array_equal(got.astype(bool), o.check_edges(q1, q2, 0.01))
    desc = scenes.mixed_arms(rng=rng, n_modes=64)
    sc, o = backend().Scene(desc), O.Scene(desc)
    V, i1, i2, modes, inst = scenes.mixed_instances(rng, desc, 64, 0.05)
    k = min(len(i1), 1500)
    rc, got = _host_edges(sc, V[i1[:k]], V[i2[:k]], 0.05, modes=modes[:k])
    exp = o.check_edges(V[i1[:k]], V[i2[:k]], 0.05, modes=modes[:k])
    assert rc == 0 and np.array_equal(got.astype(bool), exp)


def test_host_list_errors():
    from helpers import backend, scenes
    rng = np.random.default_rng(10)
    desc = scenes.quad_arms(rng=rng)
    sc = backend().Scene(desc)
    q = near_home(rng, desc, 4)
    rc, _ = _host_edges(sc, q, q + 0.1, 0.05, N_start=[0, 0, 50, 0])  # N_start > N
    assert rc == -1
    rc, _ = _host_edges(sc, q, q + 0.1, 0.05, N_start=[0, -1, 0, 0])
    assert rc == -1
    rc, _ = _host_edges(sc, q, q + 0.1, 0.0)  # no resolution and no N
    assert rc == -1
    rc, got = _host_edges(sc, q[:0], q[:0], 0.05)
    assert rc == 0
    # N entries <= 0 take the resolution's N
    rc, a = _host_edges(sc, q, q + 0.3, 0.05, N=[0, 0, 0, 0])
    rc2, b = _host_edges(sc, q, q + 0.3, 0.05)
    assert rc == 0 and rc2 == 0 and np.array_equal(a, b)
