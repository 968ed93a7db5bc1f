"""GPU collision backend: a ``BaseProblem`` whose collision checks run in libmrcc (gfx950).

Per-call methods keep the reference signatures (planning_env.py:1724-1881) so planners call them
unchanged; each call is one kernel launch on inputs staged through pinned memory.  Batch methods
(``batch_is_collision_free``, ``batch_is_edge_collision_free``) and ``is_path_collision_free`` issue
one launch for many configurations / edges.  PyTorch provides device buffers and the stream.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import EdgeBatch, Tune, check, dptr, iptr, lib
from .configuration import Configuration, NpConfiguration, batch_config_cost, config_cost, edge_num_states
from .planning_env import BaseProblem, Mode, State, generate_binary_search_indices
from .scenes import SceneDesc

ALL_ROBOTS = 0xFFFFFFFF
HOST_BATCH_MAX = 4096  # edge lists up to this length go through mrcc_check_edges_host


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("libmrcc batch calls need a HIP device (torch.cuda.is_available() is False)")
    return torch


def _stream_handle():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Scene:
    """Owner of one ``mrcc_scene`` handle built from a ``SceneDesc``."""

    def __init__(self, desc: SceneDesc, device: int = 0):
        L = lib()
        check(L.mrcc_init(device))
        self.desc = desc
        self.device = device
        h = ctypes.c_void_p()
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        if desc.kind == "disks":
            rr, sc, sr = f64(desc.robot_radius), f64(desc.sphere_center), f64(desc.sphere_radius)
            rc, rs = f64(desc.rect_center), f64(desc.rect_size)
            keep = [rr, sc, sr, rc, rs]
            check(L.mrcc_scene_create_disks(len(rr), desc.robot_dim, dptr(rr), len(sr), dptr(sc), dptr(sr),
                                            len(rc), dptr(rc), dptr(rs), ctypes.byref(h)))
        elif desc.kind == "planar":
            rh, op, oh = f64(desc.robot_half), f64(desc.obstacle_pose), f64(desc.obstacle_half)
            keep = [rh, op, oh]
            check(L.mrcc_scene_create_planar_boxes(len(rh), dptr(rh), len(op), dptr(op), dptr(oh),
                                                   ctypes.byref(h)))
        elif desc.kind == "arms":
            b, dh, sl, bx = f64(desc.base), f64(desc.dh), f64(desc.sphere_local), f64(desc.boxes)
            cnt = np.ascontiguousarray(desc.link_sphere_count, dtype=np.int32)
            keep = [b, dh, sl, bx, cnt]
            if desc.joint_fixed is not None:
                jf = f64(desc.joint_fixed)
                keep.append(jf)
                check(L.mrcc_scene_create_arm_chains(desc.n_arms, desc.dof, dptr(b), dptr(jf), iptr(cnt), dptr(sl),
                                                     len(bx), dptr(bx), ctypes.byref(h)))
            else:
                check(L.mrcc_scene_create_arms(desc.n_arms, desc.dof, dptr(b), dptr(dh), iptr(cnt), dptr(sl),
                                               len(bx), dptr(bx), ctypes.byref(h)))
        else:
            raise ValueError(f"unknown scene kind {desc.kind}")
        del keep
        self.handle = h
        if desc.kind == "arms" and getattr(desc, "joint_type", None) is not None:
            jt = np.ascontiguousarray(desc.joint_type, dtype=np.uint8)
            tr = np.ascontiguousarray(desc.joint_travel(), dtype=np.float64)
            if jt.shape != (desc.n_arms * desc.dof,):
                raise ValueError(f"joint_type: [{desc.n_arms * desc.dof}] (one per joint)")
            check(L.mrcc_scene_set_joint_types(self.handle, jt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                               dptr(tr)))
        if desc.kind == "arms" and desc.modes is not None:
            self.set_modes(desc.modes)
        self.broadphase = True
        info = np.zeros(4, dtype=np.int32)
        check(L.mrcc_scene_info(self.handle, iptr(info)))
        self.kind, self.D, self.n_robots, self.n_modes = (int(v) for v in info)

    def set_broadphase(self, enabled: bool) -> None:
        """Arm scenes: link bounding-sphere culling on (default) or brute-force pair set."""
        check(lib().mrcc_scene_set_broadphase(self.handle, int(bool(enabled))))
        self.broadphase = bool(enabled)

    def set_modes(self, modes: dict) -> None:
        na = np.ascontiguousarray(modes["n_active"], dtype=np.int32)
        g = np.ascontiguousarray(modes["has_grasp"], dtype=np.uint8)
        rel = np.ascontiguousarray(modes["grasp_rel"], dtype=np.float64)
        half = np.ascontiguousarray(modes["grasp_half"], dtype=np.float64)
        n_rest = int(modes.get("n_rest", 0))
        ns = self.desc.n_arms + n_rest
        if g.shape != (len(na), ns) or rel.shape != (len(na), ns, 12) or half.shape != (len(na), ns, 3):
            raise ValueError(f"modes: has_grasp [m, {ns}], grasp_rel [m, {ns}, 12], grasp_half [m, {ns}, 3] "
                             f"(n_arms + n_rest box slots per mode)")
        check(lib().mrcc_scene_set_modes(self.handle, len(na), iptr(na), n_rest,
                                         g.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), dptr(rel), dptr(half)))
        self.n_modes = len(na)

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().mrcc_scene_destroy(h)
            except Exception:
                pass
            self.handle = None

    # -------------------------------------------------------------------------------------------
    # device-tensor entry points (inputs already resident)
    def _check_rows(self, q, name="q"):
        """A float64 HIP tensor [n, D] of this scene (the kernels read D coordinates per row)."""
        if q.dtype != _torch().float64 or not q.is_cuda or q.dim() != 2 or q.shape[1] != self.D:
            raise ValueError(f"{name}: need a float64 HIP tensor [n, {self.D}], got {q.dtype} "
                             f"{tuple(q.shape)} on {q.device}")

    def check_configs_device(self, q, out, modes=None, robot_mask: int = ALL_ROBOTS, stream=None):
        """q: float64 device tensor [n, D] (rows contiguous) or any 2-D strided view."""
        self._check_rows(q)
        n = q.shape[0]
        if out.shape[0] != n or (modes is not None and modes.shape[0] != n):
            raise ValueError("out / modes need one entry per configuration")
        check(lib().mrcc_check_configs(self.handle, ctypes.c_void_p(q.data_ptr()), n, q.stride(0), q.stride(1),
                                       ctypes.c_void_p(modes.data_ptr()) if modes is not None else None,
                                       robot_mask, ctypes.c_void_p(out.data_ptr()),
                                       stream if stream is not None else _stream_handle()))

    def edge_batch(self, q1, q2, out, resolution, N=None, N_start=None, N_max=None,
                   include_endpoints=False, modes=None, robot_mask=ALL_ROBOTS, idx1=None, idx2=None):
        """Builds the C batch descriptor over device tensors (kept alive by the caller)."""
        self._check_rows(q1, "q1")
        self._check_rows(q2, "q2")
        if q1.stride() != q2.stride():
            raise ValueError("q1 and q2 must share the layout")
        n = out.shape[0]
        if idx1 is None and q1.shape[0] < n or idx2 is None and q2.shape[0] < n:
            raise ValueError("q1 / q2 need a row per edge (or idx1 / idx2)")
        for t, nm in ((N, "N"), (N_start, "N_start"), (N_max, "N_max"), (modes, "modes"), (idx1, "idx1"),
                      (idx2, "idx2")):
            if t is not None and (t.dtype != _torch().int32 or t.shape[0] != n):
                raise ValueError(f"{nm}: need int32 [n_edges]")
        p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        b = EdgeBatch()
        b.q1, b.q2 = p(q1), p(q2)
        b.row_stride, b.dim_stride = q1.stride(0), q1.stride(1)
        b.idx1, b.idx2 = p(idx1), p(idx2)
        b.n_edges = out.shape[0]
        b.resolution = float(resolution)
        b.N, b.N_start, b.N_max = p(N), p(N_start), p(N_max)
        b.include_endpoints = int(bool(include_endpoints))
        b.mode_ids = p(modes)
        b.robot_mask = robot_mask
        b.out = p(out)
        return b

    def check_edges_device(self, batch: EdgeBatch, tune: Optional[Tune] = None, stream=None):
        check(lib().mrcc_check_edges(self.handle, ctypes.byref(batch), ctypes.byref(tune) if tune else None,
                                     stream if stream is not None else _stream_handle()))

    def check_edges_noalloc_device(self, batch: EdgeBatch, n_invalid=None, tune: Optional[Tune] = None,
                                   stream=None):
        """mrcc_check_edges_noalloc: the batched check on the library's per-GPU workspace (no device
        allocation, asynchronous); n_invalid: optional int64 device tensor [1]."""
        check(lib().mrcc_check_edges_noalloc(self.handle, ctypes.byref(batch), ctypes.byref(tune) if tune else None,
                                             ctypes.c_void_p(n_invalid.data_ptr()) if n_invalid is not None else None,
                                             stream if stream is not None else _stream_handle()))

    def check_edges_per_call_device(self, batch: EdgeBatch, stream=None):
        check(lib().mrcc_check_edges_per_call(self.handle, ctypes.byref(batch),
                                              stream if stream is not None else _stream_handle()))


class EdgePlan:
    """A planned batched edge check bound to device tensors (graph-captured by default)."""

    def __init__(self, scene: Scene, batch: EdgeBatch, tune: Optional[Tune] = None, keep=()):
        self.scene = scene
        self.batch = batch
        self._keep = keep
        self.handle = ctypes.c_void_p()
        check(lib().mrcc_edge_plan_create(scene.handle, ctypes.byref(batch),
                                          ctypes.byref(tune) if tune is not None else None,
                                          _stream_handle(), ctypes.byref(self.handle)))

    def run(self, stream=None):
        check(lib().mrcc_edge_plan_run(self.handle, stream if stream is not None else _stream_handle()))

    def stats(self):
        s = (ctypes.c_int64 * 8)()
        check(lib().mrcc_edge_plan_stats(self.handle, s))
        return {"states": s[0], "rechecks": s[1], "dead_skips": s[2], "slots": s[3],
                "link_box_tests": s[4], "sphere_box_tests": s[5], "link_pair_tests": s[6],
                "sphere_pair_tests": s[7]}

    def timing(self):
        """Milliseconds of the last run (plans made with time_kernels=1 and use_graph=0):
        queue build, main kernel, whole run; from HIP events on the run's stream."""
        t = (ctypes.c_double * 3)()
        check(lib().mrcc_edge_plan_timing(self.handle, t))
        return {"build_ms": t[0], "kernel_ms": t[1], "run_ms": t[2]}

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().mrcc_edge_plan_destroy(h)
            except Exception:
                pass
            self.handle = None


def make_tune(slots_per_pass=0, slots_per_ticket=0, waves_per_cu=0, sorted=1, persistent=1,
              use_graph=1, collect_stats=0, time_kernels=0) -> Tune:
    t = Tune()
    t.slots_per_pass, t.slots_per_ticket, t.waves_per_cu = slots_per_pass, slots_per_ticket, waves_per_cu
    t.sorted, t.persistent, t.use_graph, t.collect_stats = sorted, persistent, use_graph, collect_stats
    t.time_kernels = time_kernels
    return t


class MrccEnvironment(BaseProblem):
    """``BaseProblem`` backend on the GPU engine.

    ``tolerance`` arguments are accepted and unused, as in AbstractEnvironment
    (abstract_env.py:316-317): the predicates are exact booleans.
    """

    def __init__(self, desc: SceneDesc, robots: Optional[List[str]] = None, start_pos=None,
                 collision_resolution: float = 0.01, collision_tolerance: float = 0.01,
                 device: int = 0, manipulation=None):
        """``manipulation``: a ``modes.ManipulationSpec`` (movable boxes and tasks) for problems whose
        modes change the scene; the device modes are then made from each mode's scene graph."""
        self.desc = desc
        if manipulation is not None:
            if desc.kind != "arms" or desc.modes is not None:
                raise ValueError("manipulation problems need an arm scene without preset modes")
        self.scene = Scene(desc, device)
        self.limits = desc.limits
        n = desc.n_robots
        self.robots = robots if robots is not None else [f"a{i + 1}" for i in range(n)]
        sl = desc.robot_slices()
        self.robot_idx = {r: list(range(s, e)) for r, (s, e) in zip(self.robots, sl)}
        self.robot_dims = {r: e - s for r, (s, e) in zip(self.robots, sl)}
        if start_pos is None:
            start_pos = NpConfiguration(np.zeros(desc.dim), sl)
        self.start_pos = start_pos
        self.start_mode = Mode([0] * n, start_pos)
        self.collision_resolution = collision_resolution
        self.collision_tolerance = collision_tolerance
        self._slices = sl
        self._current_mode = self.start_mode
        self._graphs = None
        # the per-call entry points' result (one per environment: calls on it are synchronous)
        self._res = ctypes.c_int32()
        self._res_p = ctypes.pointer(self._res)
        if manipulation is not None:
            from .modes import SceneGraphs

            self._graphs = SceneGraphs(manipulation, desc, self.robots, start_pos.state())
            self._sg_ids: Dict[tuple, int] = {}
            self._dev_modes: List[tuple] = []
            self._mode_of(self.start_mode)
        BaseProblem.__init__(self)

    # ---- modes ----------------------------------------------------------------------------
    def _mode_of(self, mode) -> int:
        """Device mode of a manipulation problem's mode: its scene graph (replayed from the start
        mode, modes.SceneGraphs) mapped to a device mode, made on first use."""
        cached = getattr(mode, "_mrcc_device_mode", None)
        if cached is not None and cached[0] is self:
            return cached[1]
        sg = self._graphs.scene_graph(mode)
        key = self._graphs.key(sg)
        mid = self._sg_ids.get(key)
        if mid is None:
            mid = len(self._dev_modes)
            self._dev_modes.append(self._graphs.device_slots(sg))
            n = len(self._dev_modes)
            self.scene.set_modes({"n_active": np.full(n, self.desc.n_arms, dtype=np.int32),
                                  "has_grasp": np.stack([m[0] for m in self._dev_modes]),
                                  "grasp_rel": np.stack([m[1] for m in self._dev_modes]),
                                  "grasp_half": np.stack([m[2] for m in self._dev_modes]),
                                  "n_rest": len(self._graphs.spec.objects)})
            self._sg_ids[key] = mid
        try:
            mode._mrcc_device_mode = (self, mid)
        except AttributeError:
            pass
        return mid

    def _rows(self, qs) -> np.ndarray:
        """Configurations (an [n, D] array or a sequence of configurations) as contiguous float64
        rows; np.array over the rows (np.stack costs several times more for short lists)."""
        if isinstance(qs, np.ndarray):
            return np.ascontiguousarray(qs, dtype=np.float64)
        if len(qs) == 0:
            return np.zeros((0, self.scene.D))
        return np.array([self._flat(q) for q in qs], dtype=np.float64)

    def _mode_ids(self, modes) -> np.ndarray:
        """int32 mode_id of each entry, once per distinct object (a candidate list repeats one mode)."""
        n = len(modes)
        if n and all(m is modes[0] for m in modes):
            return np.full(n, self.mode_id(modes[0]), dtype=np.int32)
        seen = {}
        out = np.empty(n, dtype=np.int32)
        for k, m in enumerate(modes):
            v = seen.get(id(m))
            if v is None:
                v = seen[id(m)] = self.mode_id(m)
            out[k] = v
        return out

    def mode_id(self, mode) -> int:
        """Device mode id of a mode: an int as is; in a manipulation problem, the device mode of
        the Mode's scene graph; otherwise a Mode by its ``mrcc_mode_id``.  A Mode without one is
        mode 0 only in a scene with a single mode; otherwise ValueError (a wrong scene graph would
        give wrong answers silently)."""
        if mode is None:
            return 0
        if isinstance(mode, (int, np.integer)):
            m = int(mode)
        elif self._graphs is not None:
            return self._mode_of(mode)
        else:
            m = getattr(mode, "mrcc_mode_id", None)
            if m is None:
                if self.scene.n_modes > 1:
                    raise ValueError("mode has no device mode id (mrcc_mode_id); register it with "
                                     "register_mode")
                return 0
            m = int(m)
        if not 0 <= m < self.scene.n_modes:
            raise ValueError(f"mode id {m} outside [0, {self.scene.n_modes})")
        return m

    def set_to_mode(self, mode: Mode) -> None:
        """rai_base_env.py:704-839: the scene of `mode` (its device mode is made if new)."""
        if self._graphs is not None:
            self._mode_of(mode)
        self._current_mode = mode

    def get_scenegraph_info_for_mode(self, mode: Mode, is_start_mode: bool = False) -> Dict:
        """The mode's scene graph as rai reports it (rai_base_env.py:678-702): movable object ->
        (parent frame, pose relative to it rounded to 3 decimals, as bytes); {} without movable
        objects."""
        if self._graphs is None:
            return {}
        self._mode_of(mode)
        out = {}
        for name, (holder, T) in self._graphs.scene_graph(mode).items():
            rel = np.round(T, 3)
            rel = np.where(abs(rel) < 1e-6, 0.0, rel)
            out[name] = (self._graphs.spec.root if holder is None else self.robots[holder], rel.tobytes())
        return out

    def robot_mask(self, r) -> int:
        if isinstance(r, str):
            r = [r]
        m = 0
        for name in r:
            m |= 1 << self.robots.index(name)
        return m

    # ---- per call --------------------------------------------------------------------------
    def _flat(self, q) -> np.ndarray:
        v = np.ascontiguousarray(q.state() if isinstance(q, Configuration) else q, dtype=np.float64)
        if v.shape != (self.scene.D,):
            raise ValueError(f"configuration of shape {v.shape}, the scene has D = {self.scene.D}")
        return v

    def is_collision_free(self, q: Optional[Configuration], mode: Optional[Mode]) -> bool:
        """abstract_env.py:255-276 contract; ValueError when q is None (:256-257)."""
        if q is None:
            raise ValueError
        v = self._flat(q)
        check(lib().mrcc_check_config_host(self.scene.handle, v.ctypes.data, self.mode_id(mode), ALL_ROBOTS,
                                           self._res_p))
        return bool(self._res.value)

    def is_collision_free_for_robot(self, r, q, m=None, collision_tolerance=None, set_mode=True) -> bool:
        """Collisions involving robots r and no robot outside r (rai_base_env.py:515-615)."""
        v = self._flat(q)
        check(lib().mrcc_check_config_host(self.scene.handle, v.ctypes.data, self.mode_id(m), self.robot_mask(r),
                                           self._res_p))
        return bool(self._res.value)

    def is_edge_collision_free(self, q1, q2, mode, resolution=None, tolerance=None,
                               include_endpoints=False, N_start=0, N_max=None, N=None) -> bool:
        """abstract_env.py:301-354: N states at `resolution`, BFS order, window [N_start, N_max).

        N, when not given, is the reference's (:321-323, the distance by compute_abs_max_reduction's
        rule); the library computes it in the common case (N_start = 0).  Here it is computed only
        for the check N_start > N (the reference asserts) and, after a failed call, to raise the
        reference's exception for a distance that is not a number or infinite."""
        if resolution is None:
            resolution = self.collision_resolution
        a, b = self._flat(q1), self._flat(q2)
        if N is None and N_start != 0:
            N = edge_num_states(a, b, resolution)
        if N is not None and N_start > N:
            assert False
        if N_start < 0:
            raise ValueError("N_start < 0 is not supported")
        rc = lib().mrcc_check_edge_host(self.scene.handle, a.ctypes.data, b.ctypes.data, float(resolution),
                                        0 if N is None else int(N), int(N_start),
                                        -1 if N_max is None else int(N_max), int(bool(include_endpoints)),
                                        self.mode_id(mode), ALL_ROBOTS, self._res_p)
        if rc != _lib.MRCC_OK:
            if N is None:
                edge_num_states(a, b, resolution)  # ValueError / OverflowError as int() in the reference
            check(rc)
        return bool(self._res.value)

    # ---- batched ---------------------------------------------------------------------------
    def batch_is_collision_free(self, qs, modes=None, robots=None) -> np.ndarray:
        torch = _torch()
        Q = self._rows(qs)
        if Q.ndim != 2 or Q.shape[1] != self.scene.D:
            raise ValueError(f"configurations: need [n, {self.scene.D}], got {Q.shape}")
        n = Q.shape[0]
        if modes is not None and len(modes) != n:
            raise ValueError("modes need one entry per configuration")
        if n == 0:
            return np.zeros(0, dtype=bool)
        qd = torch.from_numpy(Q).cuda()
        out = torch.empty(n, dtype=torch.uint8, device="cuda")
        md = None
        if modes is not None:
            md = torch.from_numpy(np.asarray(self._mode_ids(modes), dtype=np.int32)).cuda()
        mask = ALL_ROBOTS if robots is None else self.robot_mask(robots)
        self.scene.check_configs_device(qd, out, md, mask)
        return out.cpu().numpy().astype(bool)

    def _batch_is_collision_free(self, qs, mode=None) -> bool:
        """True iff every configuration of `qs` is collision free (abstract_env.py:278-299).

        The reference's batched form (unused by its planners) evaluates its distances with
        array reductions, whose rounding can differ from the per-configuration predicates by an
        ulp at exact contact; this form applies the per-configuration predicates of
        ``is_collision_free`` to every configuration, so it agrees with a loop over that method.
        """
        if len(qs) == 0:
            return True
        if mode is None:
            modes = None
        elif isinstance(mode, (list, tuple)):
            modes = mode
        else:
            modes = [mode] * len(qs)
        return bool(np.all(self.batch_is_collision_free(qs, modes)))

    def batch_is_edge_collision_free(self, q1s, q2s, modes=None, resolution=None, include_endpoints=False,
                                     N_start=None, N_max=None, N=None, robots=None,
                                     tune: Optional[Tune] = None) -> np.ndarray:
        torch = _torch()
        if resolution is None:
            resolution = self.collision_resolution
        A, B = self._rows(q1s), self._rows(q2s)
        if A.ndim != 2 or A.shape != B.shape or A.shape[1] != self.scene.D:
            raise ValueError(f"q1s, q2s: need two [n, {self.scene.D}] arrays, got {A.shape} and {B.shape}")
        n = A.shape[0]
        if modes is not None and len(modes) != n:
            raise ValueError("modes need one entry per edge")
        if n == 0:
            return np.zeros(0, dtype=bool)
        iv = lambda v: None if v is None else torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).cuda()
        if N_start is not None:
            Nv = np.asarray(N) if N is not None else np.maximum(
                2, (np.max(np.abs(A - B), axis=1) / resolution).astype(np.int64) + 1)
            if np.any(np.asarray(N_start) > Nv):
                assert False, "N_start > N"
            if np.any(np.asarray(N_start) < 0):
                raise ValueError("N_start < 0 is not supported")
        mask = ALL_ROBOTS if robots is None else self.robot_mask(robots)
        if n <= HOST_BATCH_MAX and tune is None:
            # a planner's candidate list: one staging copy each way and one launch, no plan
            hv = lambda v: None if v is None else np.ascontiguousarray(np.broadcast_to(v, (n,)), dtype=np.int32)
            Nh, Sh, Mh = hv(N), hv(N_start), hv(N_max)
            Mh = None if Mh is None else np.where(Mh < 0, -1, Mh).astype(np.int32)
            mh = None if modes is None else np.asarray(self._mode_ids(modes), dtype=np.int32)
            res = np.empty(n, dtype=np.uint8)
            # plain addresses (c_void_p): ctypes.cast, which data_as uses, costs microseconds a call
            p = lambda v: None if v is None else v.ctypes.data
            check(lib().mrcc_check_edges_host(self.scene.handle, A.ctypes.data, B.ctypes.data, n, float(resolution),
                                              p(Nh), p(Sh), p(Mh), int(bool(include_endpoints)), p(mh), mask,
                                              res.ctypes.data))
            return res.astype(bool)
        a, b = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
        out = torch.empty(n, dtype=torch.uint8, device="cuda")
        Nt, St, Mt = iv(N), iv(N_start), iv(N_max)  # N_max entries < 0 mean None
        md = iv(None if modes is None else self._mode_ids(modes))
        batch = self.scene.edge_batch(a, b, out, resolution, Nt, St, Mt, include_endpoints, md, mask)
        self.scene.check_edges_device(batch, tune)
        return out.cpu().numpy().astype(bool)

    def is_path_collision_free(self, path: List[State], binary_order: bool = True,
                               resolution: Optional[float] = None, tolerance: Optional[float] = None,
                               check_edges_in_order: bool = False, check_start_and_end: bool = True) -> bool:
        """Same boolean as planning_env.py:1765-1881, from one batched launch of the vertices and
        one of the edges: both schedules check every interior state of every edge; the vertex set
        differs (the in-order schedule always checks the last vertex)."""
        if resolution is None:
            resolution = self.collision_resolution
        K = len(path)
        if K < 2:
            raise IndexError("a path needs at least two states")
        verts = [i for i in range(K - 1) if not (i == 0 and not check_start_and_end)]
        if check_edges_in_order or check_start_and_end:
            verts.append(K - 1)
        V = np.stack([self._flat(s.q) for s in path])
        vm = [path[i].mode for i in verts]
        ok_v = self.batch_is_collision_free(V[verts], modes=vm)
        if not ok_v.all():
            return False
        em = [path[i].mode for i in range(K - 1)]
        ok_e = self.batch_is_edge_collision_free(V[:-1], V[1:], modes=em, resolution=resolution)
        return bool(ok_e.all())

    # ---- misc ------------------------------------------------------------------------------
    def config_cost(self, start: Configuration, end: Configuration) -> float:
        """abstract_env.py:238-239: config_cost(start, end, cost_metric, cost_reduction)."""
        return config_cost(start, end, self.cost_metric, self.cost_reduction)

    def batch_config_cost(self, starts, ends, tmp_agent_slice=None):
        """abstract_env.py:241-253 (one device launch, mrcc_config_cost)."""
        return batch_config_cost(starts, ends, self.cost_metric, self.cost_reduction,
                                 tmp_agent_slice=tmp_agent_slice)
