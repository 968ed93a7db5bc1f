"""Path post-processing beside the planners' shortcutting (planners/shortcutting.py:87-244):
gradient repair and shortening of a composite path on the differentiable clearance cost, every
result re-validated by the exact boolean checker.

* ``interpolate_path`` / ``path_cost``: the reference's helpers (problems/util.py:51-101): a path
  discretised at a resolution in a configuration metric, and the summed batch cost of consecutive
  states;
* ``gradient_shortcut(env, path)``: interpolates the path, then moves its interior states down
  clearance cost + w * sum of squared steps (binary32, the hand-written backward of
  ``clearance_cost``), keeping the end states and both states of every mode switch fixed.  The
  moved path replaces the input only when ``env.is_path_collision_free`` accepts it (batched, the
  reference's boolean) and, for an input that was already collision free, only when its
  ``path_cost`` is lower, so an accepted path keeps the reference's semantics.  It is meant to run
  before or after ``robot_mode_shortcut``.
"""
from __future__ import annotations

from typing import Callable, List, Optional, Tuple

import numpy as np

from .configuration import config_dist
from .planning_env import State


def path_cost(path: List[State], batch_cost_fun: Callable, agent_slices=None) -> float:
    """Sum of the batch cost of consecutive states (util.py:51-67)."""
    if isinstance(path[0], State):
        pts = [s.q.state() for s in path]
        return float(np.sum(batch_cost_fun(pts, None, tmp_agent_slice=path[0].q._array_slice)))
    if isinstance(path[0], np.ndarray) and agent_slices is not None:
        return float(np.sum(batch_cost_fun(path, None, tmp_agent_slice=agent_slices)))
    raise ValueError("path_cost: a list of States, or arrays with agent_slices")


def interpolate_path(path: List[State], resolution: float = 0.1, kind: str = "max") -> List[State]:
    """util.py:70-101: each segment cut into N = max(1, int(dist / resolution)) steps of
    (q1 - q0) / N in the configuration metric `kind`; the states q0 + step * j (j < N) carry the
    segment's first mode, and the path's last state closes it."""
    out = []
    for a, b in zip(path[:-1], path[1:]):
        if kind == "max":  # compute_abs_max_reduction's rule (configuration.py:174-197), on the host
            d = np.abs(b.q.state() - a.q.state())
            dist = float(d[0]) if np.isnan(d[0]) else float(np.nanmax(d))
        else:
            dist = config_dist(a.q, b.q, kind)
        N = max(1, int(dist / resolution))
        q0, q1 = a.q.state(), b.q.state()
        step = (q1 - q0) / N
        for j in range(N):
            out.append(State(a.q.from_flat(q0 + step * j), a.mode))
    out.append(State(path[-1].q, path[-1].mode))
    return out


def _fixed_states(path: List[State]) -> np.ndarray:
    """Ends of the path and both states of every mode switch."""
    W = len(path)
    fixed = np.zeros(W, dtype=bool)
    fixed[0] = fixed[-1] = True
    for k in range(W - 1):
        if path[k].mode != path[k + 1].mode:
            fixed[k] = fixed[k + 1] = True
    return fixed


def gradient_shortcut(env, path: List[State], resolution: float = 0.1, steps: int = 60, lr: float = 0.02,
                      margin: float = 0.02, smooth: float = 1.0,
                      collision_resolution: Optional[float] = None) -> Tuple[List[State], dict]:
    """Returns (path, info): the interpolated path moved down clearance + smoothness when the exact
    checker accepts it (and, for a collision-free input, when its cost is lower), else the
    interpolated input.  info: accepted, valid_before, valid_after, cost_before, cost_after."""
    import torch

    from .clearance import clearance_cost

    new_path = interpolate_path(path, resolution)
    res = env.collision_resolution if collision_resolution is None else collision_resolution
    valid_before = env.is_path_collision_free(new_path, resolution=res)
    cost_before = path_cost(new_path, env.batch_config_cost)
    W = len(new_path)
    fixed = torch.from_numpy(_fixed_states(new_path)).cuda()
    lo = torch.tensor(env.limits[0], dtype=torch.float32, device="cuda")
    hi = torch.tensor(env.limits[1], dtype=torch.float32, device="cuda")
    Q0 = np.stack([s.q.state() for s in new_path]).astype(np.float32)
    q = torch.from_numpy(Q0).cuda()
    anchor = q.clone()
    q.requires_grad_(True)
    opt = torch.optim.Adam([q], lr=lr)
    for _ in range(steps):
        opt.zero_grad()
        c = clearance_cost(env.scene, q, margin).sum()
        s = ((q[1:] - q[:-1]) ** 2).sum()
        (c + smooth * s).backward()
        q.grad[fixed] = 0
        opt.step()
        with torch.no_grad():
            q.copy_(torch.maximum(torch.minimum(q, hi), lo))
            q[fixed] = anchor[fixed]
    moved = q.detach().double().cpu().numpy()
    cand = []
    for k, s in enumerate(new_path):
        v = s.q.state() if bool(fixed[k]) else moved[k]  # fixed states keep their binary64 values
        cand.append(State(s.q.from_flat(np.array(v, dtype=np.float64)), s.mode))
    valid_after = env.is_path_collision_free(cand, resolution=res)
    cost_after = path_cost(cand, env.batch_config_cost)
    accept = bool(valid_after and (not valid_before or cost_after < cost_before))
    info = {"accepted": accept, "valid_before": bool(valid_before), "valid_after": bool(valid_after),
            "cost_before": cost_before, "cost_after": cost_after, "states": W}
    return (cand if accept else new_path), info
