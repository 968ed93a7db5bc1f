"""Device parity of arm scenes with chain joints (T <- T Rz(q) C): the four-UR10e scene imported
from the reference's mujoco world (committed fixture) and DH arms written as chain joints.
Bit-exact against the oracle: forward kinematics, configurations, edges (batched and short host
lists), masks, clearance cost on the chain scene against the oracle's binary64 cost."""
import ctypes

import numpy as np
import pytest

from helpers import O, gpu_configs, gpu_edges, pkg, scenes
from test_mjcf import dh_as_chain, fixture_desc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import importlib
    return importlib.import_module("multirobot-pathplanning-benchmark_amd.backend")


@pytest.fixture(scope="module")
def ur10(B):
    desc, _ = fixture_desc()
    return desc, B.Scene(desc), O.Scene(desc)


def test_chain_fk_bit_exact(B, ur10):
    import torch
    desc, sc, o = ur10
    rng = np.random.default_rng(1)
    Q = rng.uniform(desc.limits[0], desc.limits[1], size=(1000, desc.dim))
    qd = torch.from_numpy(Q).cuda()
    out = torch.empty((1000, desc.dim, 12), dtype=torch.float64, device="cuda")
    pkg._lib.check(pkg.lib().mrcc_arm_fk(sc.handle, ctypes.c_void_p(qd.data_ptr()), 1000, desc.dim, 1,
                                         ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for r in range(0, 1000, 37):
        assert np.array_equal(got[r], o.arm_fk(Q[r]).reshape(desc.dim, 12)), r


def test_ur10_configs(ur10):
    desc, sc, o = ur10
    rng = np.random.default_rng(2)
    home = np.tile([0, -2, 1.0, -1.0, -1.57, 1.0], 4)
    Q = np.concatenate([rng.uniform(desc.limits[0], desc.limits[1], size=(20000, desc.dim)),
                        home + rng.normal(0, 0.3, size=(20000, desc.dim))])
    got = gpu_configs(sc, Q)
    exp = o.check_configs(Q)
    assert np.array_equal(got, exp)
    assert 0.05 < exp.mean() < 0.95
    for mask in (0b0001, 0b0110):
        assert np.array_equal(gpu_configs(sc, Q[:5000], mask=mask), o.check_configs(Q[:5000], mask=mask))


def test_ur10_edges(ur10):
    desc, sc, o = ur10
    rng = np.random.default_rng(3)
    home = np.tile([0, -2, 1.0, -1.0, -1.57, 1.0], 4)
    q1 = home + rng.normal(0, 0.3, size=(6000, desc.dim))
    q2 = q1 + rng.normal(0, 0.3, size=(6000, desc.dim))
    got = gpu_edges(sc, q1, q2, 0.01)
    exp = o.check_edges(q1, q2, 0.01)
    assert np.array_equal(got, exp)
    assert 0.05 < exp.mean() < 0.95
    # the environment's short-list path (mrcc_check_edges_host) on the same edges
    import importlib
    backend = importlib.import_module("multirobot-pathplanning-benchmark_amd.backend")
    env = backend.MrccEnvironment(desc, collision_resolution=0.01)
    env_free = env.batch_is_edge_collision_free(q1[:300], q2[:300])
    assert np.array_equal(env_free, exp[:300])


def test_dh_as_chain_on_device(B):
    rng = np.random.default_rng(4)
    desc = scenes.quad_arms(rng=rng)
    chain = dh_as_chain(desc)
    a, b = B.Scene(desc), B.Scene(chain)
    Q = scenes.uniform_configs(rng, desc, 20000)
    assert np.array_equal(gpu_configs(a, Q), gpu_configs(b, Q))
    q1, q2 = scenes.lognormal_edges(rng, desc, 20000, 0.05)
    ea, eb = gpu_edges(a, q1, q2, 0.05), gpu_edges(b, q1, q2, 0.05)
    assert np.array_equal(ea, eb)
    assert np.array_equal(eb, O.Scene(chain).check_edges(q1, q2, 0.05))


def test_ur10_clearance_matches_oracle(ur10):
    """Clearance cost (binary32) against the oracle's binary64 cost, and its joint gradient against
    binary64 central differences: the joint axis of a chain joint is z of the frame before it."""
    import torch
    desc, sc, o = ur10
    rng = np.random.default_rng(5)
    home = np.tile([0, -2, 1.0, -1.0, -1.57, 1.0], 4)
    Q = (home + rng.normal(0, 0.5, size=(64, desc.dim))).astype(np.float32)
    L = pkg.lib()
    qd = torch.from_numpy(Q).cuda()
    cost = torch.empty(64, dtype=torch.float32, device="cuda")
    pkg._lib.check(L.mrcc_clearance_forward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 64, desc.dim, ctypes.c_float(0.02),
                                            ctypes.c_void_p(cost.data_ptr()), None))
    g_out = torch.ones(64, dtype=torch.float32, device="cuda")
    grad = torch.empty_like(qd)
    pkg._lib.check(L.mrcc_clearance_backward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 64, desc.dim, ctypes.c_float(0.02),
                                             ctypes.c_void_p(g_out.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
                                             None))
    torch.cuda.synchronize()
    c, g = cost.cpu().numpy(), grad.cpu().numpy()
    Q64 = Q.astype(np.float64)
    ref = np.array([o.clearance_cost(q, 0.02) for q in Q64])
    assert np.all(np.abs(c - ref) <= 1e-4 * np.maximum(ref, 1e-3))
    assert (ref > 0).sum() > 10
    h, checked = 1e-6, 0
    for r in range(64):
        if ref[r] <= 1e-3 or checked >= 6:
            continue
        fd = np.zeros(desc.dim)
        for k in range(desc.dim):
            e = np.zeros(desc.dim)
            e[k] = h
            fd[k] = (o.clearance_cost(Q64[r] + e, 0.02) - o.clearance_cost(Q64[r] - e, 0.02)) / (2 * h)
        assert np.linalg.norm(g[r] - fd) <= 1e-3 * np.linalg.norm(fd) + 1e-6, r
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("bp", [True, False])
def test_ur10_modes_and_held_boxes(B, bp):
    """Chain joints with modes: fewer active arms and boxes held at the end-effector (the grasp
    instantiation of the chain kernels), configurations and edges against the oracle."""
    desc, _ = fixture_desc()
    rng = np.random.default_rng(6)
    desc.modes = scenes.grasp_modes(rng, 64, desc.n_arms)
    sc, o = B.Scene(desc), O.Scene(desc)
    sc.set_broadphase(bp)
    home = np.tile([0, -2, 1.0, -1.0, -1.57, 1.0], desc.n_arms)
    Q = np.clip(home + rng.normal(0, 0.4, size=(6000, desc.dim)), desc.limits[0], desc.limits[1])
    modes = rng.integers(0, 64, 6000).astype(np.int32)
    ref = o.check_configs(Q, modes=modes)
    assert 0.05 < ref.mean() < 0.95
    assert np.array_equal(gpu_configs(sc, Q, modes=modes), ref)
    q2 = np.clip(Q + rng.normal(0, 0.15, Q.shape), desc.limits[0], desc.limits[1])
    refe = o.check_edges(Q[:3000], q2[:3000], 0.01, modes=modes[:3000])
    assert np.array_equal(gpu_edges(sc, Q[:3000], q2[:3000], 0.01, modes=modes[:3000]), refe)
