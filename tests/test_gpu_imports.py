"""Device parity of imported scenes with prismatic (slide) joints and cylinder covers: the committed
fixtures of mujoco.swap (mj_two_dim.xml), pinocchio.middle_obs (middle_obstacle_two_agents.urdf)
and pinocchio.random_ur5 (ur5e_constrained_coll_primitives.urdf), and the brute-force scenes with
prismatic chain joints.  Bit-exact against the oracle: forward kinematics, configurations (both
broad-phase settings), prismatic coordinates beyond their travel, edges batched and per call;
clearance cost against the oracle's binary64 cost and its gradient (prismatic joints: d c / d q = z)
against central differences."""
import ctypes

import numpy as np
import pytest

from helpers import O, backend, gpu_configs, gpu_edges, pkg
from test_importers import WORLDS, fixture

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def B():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return backend()


def _say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.mark.parametrize("name", [w for w, _ in WORLDS])
def test_imported_fk_bit_exact(B, name):
    _say(f"{name}: forward kinematics vs oracle")
    desc, _ = fixture(name)
    sc, o = B.Scene(desc), O.Scene(desc)
    rng = np.random.default_rng(1)
    Q = rng.uniform(desc.limits[0], desc.limits[1], size=(500, desc.dim))
    qd = torch.from_numpy(Q).cuda()
    out = torch.empty((500, desc.dim, 12), dtype=torch.float64, device="cuda")
    pkg._lib.check(pkg.lib().mrcc_arm_fk(sc.handle, ctypes.c_void_p(qd.data_ptr()), 500, desc.dim, 1,
                                         ctypes.c_void_p(out.data_ptr()), None))
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for r in range(0, 500, 23):
        assert np.array_equal(got[r], o.arm_fk(Q[r]).reshape(desc.dim, 12)), r


@pytest.mark.parametrize("name", [w for w, _ in WORLDS])
@pytest.mark.parametrize("bp", [1, 0])
def test_imported_configs(B, name, bp):
    _say(f"{name}: configurations vs oracle, broadphase {bp}")
    desc, _ = fixture(name)
    sc, o = B.Scene(desc), O.Scene(desc)
    sc.set_broadphase(bp)
    rng = np.random.default_rng(2)
    lo, hi = desc.limits
    Q = rng.uniform(lo, hi, size=(30000, desc.dim))
    # a few coordinates beyond the limits: prismatic ones count as collisions
    Q[:3000] = rng.uniform(lo - 0.3 * (hi - lo), hi + 0.3 * (hi - lo), size=(3000, desc.dim))
    exp = o.check_configs(Q)
    assert np.array_equal(gpu_configs(sc, Q), exp)
    assert 0.02 < exp[3000:].mean() < 0.98
    if desc.joint_type is not None:
        tr = desc.joint_travel()
        beyond = np.any((desc.joint_type == 1) & (np.abs(Q) > tr), axis=1)
        assert beyond.sum() > 100 and not exp[beyond].any()


@pytest.mark.parametrize("name", [w for w, _ in WORLDS])
def test_imported_edges(B, name):
    _say(f"{name}: edges batched and per call vs oracle")
    desc, _ = fixture(name)
    sc, o = B.Scene(desc), O.Scene(desc)
    rng = np.random.default_rng(3)
    lo, hi = desc.limits
    q1 = rng.uniform(lo, hi, size=(4000, desc.dim))
    q2 = np.clip(q1 + rng.normal(0, 0.15, size=q1.shape) * (hi - lo) / 2, lo, hi)
    exp = o.check_edges(q1, q2, 0.01)
    assert np.array_equal(gpu_edges(sc, q1, q2, 0.01), exp)
    assert 0.02 < exp.mean() < 0.98
    assert np.array_equal(gpu_edges(sc, q1[:300], q2[:300], 0.01, per_call=True), exp[:300])
    env = B.MrccEnvironment(desc, collision_resolution=0.01)
    assert np.array_equal(env.batch_is_edge_collision_free(q1[:200], q2[:200]), exp[:200])


@pytest.mark.parametrize("name", ["pin_middle_obs.npz", "pin_dual_ur5.npz"])
def test_imported_clearance(B, name):
    _say(f"{name}: clearance cost and gradient vs oracle / central differences")
    desc, _ = fixture(name)
    sc, o = B.Scene(desc), O.Scene(desc)
    rng = np.random.default_rng(5)
    Q = rng.uniform(desc.limits[0], desc.limits[1], size=(256, desc.dim)).astype(np.float32)
    L = pkg.lib()
    qd = torch.from_numpy(Q).cuda()
    cost = torch.empty(256, dtype=torch.float32, device="cuda")
    pkg._lib.check(L.mrcc_clearance_forward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 256, desc.dim,
                                            ctypes.c_float(0.02), ctypes.c_void_p(cost.data_ptr()), None))
    g_out = torch.ones(256, dtype=torch.float32, device="cuda")
    grad = torch.empty_like(qd)
    pkg._lib.check(L.mrcc_clearance_backward(sc.handle, ctypes.c_void_p(qd.data_ptr()), 256, desc.dim,
                                             ctypes.c_float(0.02), ctypes.c_void_p(g_out.data_ptr()),
                                             ctypes.c_void_p(grad.data_ptr()), None))
    torch.cuda.synchronize()
    c, g = cost.cpu().numpy(), grad.cpu().numpy()
    Q64 = Q.astype(np.float64)
    ref = np.array([o.clearance_cost(q, 0.02) for q in Q64])
    assert np.all(np.abs(c - ref) <= 1e-4 * np.maximum(ref, 1e-3))
    assert (ref > 1e-3).sum() > 10
    h, checked, pr_moved = 1e-6, 0, 0
    pr = np.zeros(desc.dim, dtype=bool) if desc.joint_type is None else desc.joint_type == 1
    for r in range(256):
        if ref[r] <= 1e-3 or checked >= 8:
            continue
        fd = np.zeros(desc.dim)
        for k in range(desc.dim):
            e = np.zeros(desc.dim)
            e[k] = h
            fd[k] = (o.clearance_cost(Q64[r] + e, 0.02) - o.clearance_cost(Q64[r] - e, 0.02)) / (2 * h)
        assert np.linalg.norm(g[r] - fd) <= 1e-3 * np.linalg.norm(fd) + 1e-6, (r, g[r], fd)
        pr_moved += np.any(np.abs(fd[pr]) > 1e-3)
        checked += 1
    assert checked >= 4
    if pr.any():
        assert pr_moved >= 2  # the prismatic joints' gradient terms are exercised
