"""Differentiable clearance cost of arm scenes and the gradient path-repair step.

``clearance_cost(scene, q, margin)`` is a ``torch.autograd.Function`` over libmrcc's forward
(``mrcc_clearance_forward``) and hand-written backward (``mrcc_clearance_backward``, FK recomputed):
  cost(q) = sum max(0, m + r_i + r_j - |c_i - c_j|)^2 + sum max(0, m + r_i - dist(c_i, box))^2
over the pair set of the boolean check, in binary32.

``repair_paths`` moves the interior waypoints of paths down this cost (endpoints fixed) and
re-validates every repaired path with the exact boolean checker: a path is accepted only if all
its waypoints and edges are collision free, so accepted paths keep the reference's semantics
(is_path_collision_free, planning_env.py:1765-1881).  It is meant to sit beside the shortcutting
step of the planners (shortcutting.py:87-244).
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib


def _torch():
    import torch

    return torch


class _ClearanceFn:
    """Holds the autograd Function class (torch imported lazily)."""

    cls = None

    @classmethod
    def get(cls):
        if cls.cls is None:
            torch = _torch()

            class ClearanceCost(torch.autograd.Function):
                @staticmethod
                def forward(ctx, q, scene, margin):
                    q = q.contiguous()
                    n = q.shape[0]
                    cost = torch.empty(n, dtype=torch.float32, device=q.device)
                    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
                    check(lib().mrcc_clearance_forward(scene.handle, ctypes.c_void_p(q.data_ptr()), n, q.shape[1],
                                                       ctypes.c_float(margin), ctypes.c_void_p(cost.data_ptr()),
                                                       st))
                    ctx.save_for_backward(q)
                    ctx.scene = scene
                    ctx.margin = margin
                    return cost

                @staticmethod
                def backward(ctx, grad_cost):
                    (q,) = ctx.saved_tensors
                    g = grad_cost.contiguous().to(torch.float32)
                    grad_q = torch.empty_like(q)
                    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
                    check(lib().mrcc_clearance_backward(ctx.scene.handle, ctypes.c_void_p(q.data_ptr()), q.shape[0],
                                                        q.shape[1], ctypes.c_float(ctx.margin),
                                                        ctypes.c_void_p(g.data_ptr()),
                                                        ctypes.c_void_p(grad_q.data_ptr()), st))
                    return grad_q, None, None

            cls.cls = ClearanceCost
        return cls.cls


def clearance_cost(scene, q, margin: float = 0.02):
    """q: float32 device tensor [..., D] -> cost [...] (differentiable w.r.t. q)."""
    torch = _torch()
    if q.dtype != torch.float32 or not q.is_cuda:
        raise TypeError("clearance_cost needs a float32 HIP tensor")
    if q.shape[-1] != scene.D:
        raise ValueError(f"clearance_cost: rows of {q.shape[-1]} coordinates, the scene has D = {scene.D}")
    shape = q.shape[:-1]
    flat = q.reshape(-1, q.shape[-1])
    return _ClearanceFn.get().apply(flat, scene, float(margin)).reshape(shape)


def validate_paths(env, paths: np.ndarray, modes=None) -> np.ndarray:
    """Exact boolean check of [P, W, D] paths: every waypoint and every edge collision free."""
    P, W, D = paths.shape
    verts = paths.reshape(-1, D)
    vm = None if modes is None else np.repeat(np.asarray(modes), W)
    ok_v = env.batch_is_collision_free(verts, modes=vm).reshape(P, W)
    q1 = paths[:, :-1].reshape(-1, D)
    q2 = paths[:, 1:].reshape(-1, D)
    em = None if modes is None else np.repeat(np.asarray(modes), W - 1)
    ok_e = env.batch_is_edge_collision_free(q1, q2, modes=em).reshape(P, W - 1)
    return ok_v.all(axis=1) & ok_e.all(axis=1)


def repair_paths(env, paths: np.ndarray, steps: int = 60, lr: float = 0.05, margin: float = 0.02,
                 smooth: float = 0.1, modes=None):
    """Gradient descent of clearance cost + smoothness on interior waypoints; returns
    (paths_out, accepted): a path is replaced by its repaired version only when the exact checker
    accepts the repaired path; otherwise the input path is kept."""
    torch = _torch()
    P, W, D = paths.shape
    lo = torch.tensor(env.limits[0], dtype=torch.float32, device="cuda")
    hi = torch.tensor(env.limits[1], dtype=torch.float32, device="cuda")
    q = torch.tensor(paths, dtype=torch.float32, device="cuda")
    fixed = q[:, [0, W - 1]].clone()
    q.requires_grad_(True)
    opt = torch.optim.Adam([q], lr=lr)
    for _ in range(steps):
        opt.zero_grad()
        c = clearance_cost(env.scene, q, margin).sum()
        s = ((q[:, 1:] - q[:, :-1]) ** 2).sum()
        (c + smooth * s).backward()
        q.grad[:, 0] = 0
        q.grad[:, W - 1] = 0
        opt.step()
        with torch.no_grad():
            q.copy_(torch.maximum(torch.minimum(q, hi), lo))
            q[:, 0] = fixed[:, 0]
            q[:, W - 1] = fixed[:, 1]
    rep = q.detach().double().cpu().numpy()
    rep[:, 0] = paths[:, 0]
    rep[:, -1] = paths[:, -1]
    ok = validate_paths(env, rep, modes)
    out = np.where(ok[:, None, None], rep, paths)
    return out, ok
