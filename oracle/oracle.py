"""ctypes binding of the CPU oracle (oracle/_build/libmrcc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker or the timed CPU baseline.  The product never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libmrcc_oracle.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force: bool = False) -> str:
    """Compiles the oracle with oracle/Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        L.or_sincos.argtypes = [ctypes.c_double, _dp, _dp]
        L.or_bfs_indices.argtypes = [ctypes.c_int, _ip]
        L.or_edge_N.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_double]
        L.or_edge_num_states.argtypes = [ctypes.c_int] * 4
        L.or_edge_num_states.restype = ctypes.c_int64
        L.or_scene_disks.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, _dp,
                                     ctypes.c_int, _dp, _dp]
        L.or_scene_disks.restype = ctypes.c_void_p
        L.or_scene_planar.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp]
        L.or_scene_planar.restype = ctypes.c_void_p
        L.or_scene_arms.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _dp, ctypes.c_int, _dp]
        L.or_scene_arms.restype = ctypes.c_void_p
        L.or_scene_arm_chains.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _ip, _dp, ctypes.c_int, _dp]
        L.or_scene_arm_chains.restype = ctypes.c_void_p
        L.or_scene_set_modes.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, ctypes.c_int, _u8p, _dp, _dp]
        L.or_scene_set_joint_types.argtypes = [ctypes.c_void_p, _u8p, _dp]
        L.or_scene_free.argtypes = [ctypes.c_void_p]
        L.or_scene_dim.argtypes = [ctypes.c_void_p]
        L.or_config_free.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_uint32]
        L.or_edge_free.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_uint32, _i64p]
        L.or_check_configs.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64, ctypes.c_int64, _ip,
                                       ctypes.c_uint32, _u8p, ctypes.c_int]
        L.or_check_edges.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_double, _ip, _ip, _ip, ctypes.c_int, _ip,
                                     ctypes.c_uint32, _u8p, ctypes.c_int, _i64p]
        L.or_arm_fk.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.or_clearance_cost.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double]
        L.or_clearance_cost.restype = ctypes.c_double
        L.or_clearance_grad.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
        L.or_clearance_cost_batch.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int64,
                                              ctypes.c_double, _dp, ctypes.c_int]
        L.or_config_cost.argtypes = [_dp, ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_double, _dp, ctypes.c_int]
        L.or_test_obb2.argtypes = [_dp, _dp]
        L.or_test_obb3.argtypes = [_dp, _dp, _dp, _dp]
        L.or_test_sph_obb.argtypes = [_dp, ctypes.c_double, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def sincos(x: float):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def obb2_overlap(A8, B8) -> bool:
    a, pa = _d(A8)
    b, pb = _d(B8)
    return bool(lib().or_test_obb2(pa, pb))


def obb3_overlap(A12, a3, B12, b3) -> bool:
    arrs = [_d(x) for x in (A12, a3, B12, b3)]
    return bool(lib().or_test_obb3(*[p for _, p in arrs]))


def sph_obb(c3, r, G12, h3) -> bool:
    c, pc = _d(c3)
    g, pg = _d(G12)
    h, ph = _d(h3)
    return bool(lib().or_test_sph_obb(pc, float(r), pg, ph))


def bfs_indices(N: int):
    out = np.zeros(N, dtype=np.int32)
    rc = lib().or_bfs_indices(N, out.ctypes.data_as(_ip))
    assert rc == 0
    return tuple(int(v) for v in out)


def edge_N(q1, q2, res):
    q1, p1 = _d(q1)
    q2, p2 = _d(q2)
    return lib().or_edge_N(p1, p2, q1.size, float(res))


def edge_num_states(N, N_start=0, N_max=None, include_endpoints=False):
    return int(lib().or_edge_num_states(N, N_start, -1 if N_max is None else N_max,
                                        int(include_endpoints)))


class Scene:
    """An oracle scene built from the same description the engine takes (scenes.SceneDesc)."""

    def __init__(self, desc):
        L = lib()
        self._keep = []
        k = desc.kind
        if k == "disks":
            args = [_d(desc.robot_radius), _d(desc.sphere_center), _d(desc.sphere_radius),
                    _d(desc.rect_center), _d(desc.rect_size)]
            self._keep += [a for a, _ in args]
            self.h = L.or_scene_disks(len(desc.robot_radius), desc.robot_dim, args[0][1],
                                      len(desc.sphere_radius), args[1][1], args[2][1],
                                      len(desc.rect_center), args[3][1], args[4][1])
        elif k == "planar":
            args = [_d(desc.robot_half), _d(desc.obstacle_pose), _d(desc.obstacle_half)]
            self._keep += [a for a, _ in args]
            self.h = L.or_scene_planar(len(desc.robot_half), args[0][1], len(desc.obstacle_pose),
                                       args[1][1], args[2][1])
        elif k == "arms":
            b, pb = _d(desc.base)
            dh, pdh = _d(desc.dh)
            cnt, pcnt = _i(desc.link_sphere_count)
            sl, psl = _d(desc.sphere_local)
            bx, pbx = _d(desc.boxes)
            self._keep += [b, dh, cnt, sl, bx]
            if getattr(desc, "joint_fixed", None) is not None:
                jf, pjf = _d(desc.joint_fixed)
                self._keep.append(jf)
                self.h = L.or_scene_arm_chains(desc.n_arms, desc.dof, pb, pjf, pcnt, psl, len(desc.boxes), pbx)
            else:
                self.h = L.or_scene_arms(desc.n_arms, desc.dof, pb, pdh, pcnt, psl, len(desc.boxes), pbx)
            if self.h and getattr(desc, "joint_type", None) is not None:
                jt = np.ascontiguousarray(desc.joint_type, dtype=np.uint8)
                tr, ptr = _d(desc.joint_travel())
                self._keep += [jt, tr]
                if L.or_scene_set_joint_types(self.h, jt.ctypes.data_as(_u8p), ptr) != 0:
                    raise ValueError("oracle rejected the joint types")
            if desc.modes is not None:
                m = desc.modes
                na, pna = _i(m["n_active"])
                g = np.ascontiguousarray(m["has_grasp"], dtype=np.uint8)
                rel, prel = _d(m["grasp_rel"])
                half, phalf = _d(m["grasp_half"])
                self._keep += [na, g, rel, half]
                rc = L.or_scene_set_modes(self.h, len(na), pna, int(m.get("n_rest", 0)), g.ctypes.data_as(_u8p),
                                          prel, phalf)
                assert rc == 0
        else:
            raise ValueError(k)
        if not self.h:
            raise ValueError("oracle rejected the scene description")
        self.D = L.or_scene_dim(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_scene_free(self.h)
            self.h = None

    def config_free(self, q, mode=0, mask=0xFFFFFFFF) -> bool:
        q, p = _d(q)
        r = lib().or_config_free(self.h, p, mode, mask)
        if r < 0:
            raise ValueError("oracle error")
        return bool(r)

    def edge_free(self, q1, q2, res, N=None, N_start=0, N_max=None, include_endpoints=False,
                  mode=0, mask=0xFFFFFFFF, count=False):
        q1, p1 = _d(q1)
        q2, p2 = _d(q2)
        c = ctypes.c_int64(0)
        r = lib().or_edge_free(self.h, p1, p2, float(res), 0 if N is None else N, N_start,
                               -1 if N_max is None else N_max, int(include_endpoints), mode, mask,
                               ctypes.byref(c))
        if r < 0:
            raise AssertionError("N_start > N")
        return (bool(r), c.value) if count else bool(r)

    def check_configs(self, q, modes=None, mask=0xFFFFFFFF, threads=0):
        q, p = _d(q)
        n = q.shape[0]
        out = np.zeros(n, dtype=np.uint8)
        m, pm = _i(modes)
        rc = lib().or_check_configs(self.h, p, n, q.shape[1], pm, mask, out.ctypes.data_as(_u8p),
                                    threads)
        if rc != 0:
            raise ValueError("oracle error")
        return out.astype(bool)

    def check_edges(self, q1, q2, res, N=None, N_start=None, N_max=None, include_endpoints=False,
                    modes=None, mask=0xFFFFFFFF, threads=0, count=False):
        q1, p1 = _d(q1)
        q2, p2 = _d(q2)
        n = q1.shape[0]
        out = np.zeros(n, dtype=np.uint8)
        Na, pN = _i(N)
        Sa, pS = _i(N_start)
        Ma, pM = _i(N_max)
        m, pm = _i(modes)
        c = ctypes.c_int64(0)
        rc = lib().or_check_edges(self.h, p1, p2, n, q1.shape[1], float(res), pN, pS, pM,
                                  int(include_endpoints), pm, mask, out.ctypes.data_as(_u8p),
                                  threads, ctypes.byref(c))
        if rc != 0:
            raise AssertionError("oracle error (N_start > N?)")
        return (out.astype(bool), c.value) if count else out.astype(bool)

    def arm_fk(self, q):
        q, p = _d(q)
        out = np.zeros(self.D * 12, dtype=np.float64)
        assert lib().or_arm_fk(self.h, p, out.ctypes.data_as(_dp)) == 0
        return out.reshape(-1, 3, 4)

    def fk_matrices(self, q):
        """[links, 4, 4] homogeneous transforms (rows of R | t)."""
        X = self.arm_fk(q).reshape(-1, 12)
        T = np.zeros((X.shape[0], 4, 4))
        T[:, :3, :3] = X[:, :9].reshape(-1, 3, 3)
        T[:, :3, 3] = X[:, 9:]
        T[:, 3, 3] = 1.0
        return T

    def clearance_cost(self, q, margin):
        q, p = _d(q)
        return lib().or_clearance_cost(self.h, p, float(margin))

    def clearance_grad(self, q, margin):
        q, p = _d(q)
        g = np.zeros(q.size, dtype=np.float64)
        assert lib().or_clearance_grad(self.h, p, float(margin), g.ctypes.data_as(_dp)) == 0
        return g

    def clearance_cost_batch(self, q, margin, threads=0):
        q, p = _d(q)
        out = np.zeros(q.shape[0], dtype=np.float64)
        assert lib().or_clearance_cost_batch(self.h, p, q.shape[0], float(margin),
                                             out.ctypes.data_as(_dp), threads) == 0
        return out


COST_METRICS = {"max": 0, "euclidean": 1, "absmax": 2}
COST_REDUCTIONS = {"max": 0, "sum": 1, "max_only": 2}


def config_cost(p, q, slices, metric="max", reduction="max", w=0.01, threads=1):
    """or_config_cost over rows: p (n, D) or (D,) broadcast, q (n, D) float64; slices [(s, e)].
    metric: "max" (np.max(np.abs)), "euclidean", "absmax" (compute_abs_max_reduction);
    reduction: "max" (max + w * sum), "sum", "max_only" (compute_max_reduction)."""
    q = np.ascontiguousarray(np.atleast_2d(np.asarray(q, dtype=np.float64)))
    p = np.ascontiguousarray(np.asarray(p, dtype=np.float64))
    n, D = q.shape
    ps = 0 if p.ndim == 1 else D
    if p.ndim != 1:
        assert p.shape == q.shape
    sl = np.ascontiguousarray(np.asarray(slices, dtype=np.int32).reshape(-1))
    out = np.empty(n, dtype=np.float64)
    (_, pp), (_, pq), (_, psl) = _d(p), _d(q), _i(sl)
    rc = lib().or_config_cost(pp, ps, pq, D, n, D, psl, len(sl) // 2, COST_METRICS[metric],
                              COST_REDUCTIONS[reduction], float(w),
                              out.ctypes.data_as(_dp), int(threads))
    if rc != 0:
        raise ValueError("or_config_cost: bad argument")
    return out
