"""Batched forms of the planners' edge-check call sites.

The planners check candidate edges one call at a time and act on the first result that decides
(``rrtstar_base.py``, ``prm_graph.py``, ``planner_eitstar.py``).  An edge check is a pure function
of its inputs, so checking every candidate of one decision in one launch
(``MrccEnvironment.batch_is_edge_collision_free``) and then walking the results in the
reference's order gives the same decision; the planners' caches get the same entries, plus the
results of candidates the loop would not have reached.

Each helper keeps the reference's edge direction (q1 -> q2 interpolates from q1), its candidate
order and its conditions.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import numpy as np

from .configuration import Configuration


def _rows(qs: Sequence) -> np.ndarray:
    # np.array over the rows: a fresh C-contiguous float64 block (np.stack costs several times more)
    return np.array([q.state() if isinstance(q, Configuration) else np.asarray(q) for q in qs], dtype=np.float64)


def _repeat_row(q, n: int) -> np.ndarray:
    """n copies of one configuration as contiguous rows."""
    return np.repeat(_rows([q]), n, axis=0)


def find_parent_batched(env, mode, near_qs: Sequence, q_new, node_indices: np.ndarray, nearest_id,
                        batch_cost: np.ndarray, n_near_costs: np.ndarray) -> Tuple[int, float, float]:
    """The parent choice of RRT* ``find_parent`` (rrtstar_base.py:1339-1386).

    Candidates cheaper than the nearest node, in ascending cost (``np.argsort`` of the same array,
    as the reference), are accepted at the first collision-free edge candidate -> q_new.  All of
    them are checked in one launch.  Returns (index into node_indices of the parent, its cost
    c_min, its cost to the parent)."""
    idx = int(np.where(node_indices == nearest_id)[0][0])
    c_new_tensor = n_near_costs + batch_cost
    c_min, c_min_to_parent, chosen = c_new_tensor[idx], batch_cost[idx], idx
    valid_mask = c_new_tensor < c_min
    if np.any(valid_mask):
        order = np.where(valid_mask)[0][np.argsort(c_new_tensor[valid_mask])]
        A = _rows([near_qs[i] for i in order])
        B = _repeat_row(q_new, len(order))
        free = env.batch_is_edge_collision_free(A, B, modes=[mode] * len(order))
        for k, i in enumerate(order):
            if free[k]:
                c_min, c_min_to_parent, chosen = c_new_tensor[i], batch_cost[i], int(i)
                break
    return chosen, float(c_min), float(c_min_to_parent)


def rewire_batched(env, mode, q_new, near_qs: Sequence, eligible: np.ndarray) -> np.ndarray:
    """The edge checks of RRT* ``rewire`` (rrtstar_base.py:1388-1446): for each candidate i with
    ``eligible[i]`` (cheaper through q_new, not q_new's parent, finite cost, compatible mode), is
    the edge q_new -> near_qs[i] collision free?  One launch; False where not eligible."""
    eligible = np.asarray(eligible, dtype=bool)
    out = np.zeros(len(near_qs), dtype=bool)
    sel = np.nonzero(eligible)[0]
    if len(sel) == 0:
        return out
    A = _repeat_row(q_new, len(sel))
    B = _rows([near_qs[i] for i in sel])
    out[sel] = env.batch_is_edge_collision_free(A, B, modes=[mode] * len(sel))
    return out


def prevalidate_edges(env, edges: Sequence[Tuple[int, int]], qs: Dict[int, Configuration], modes: Dict[int, object],
                      resolution: Optional[float], whitelist: Dict[int, set], blacklist: Dict[int, set]) -> int:
    """Fills the PRM edge caches (prm_graph.py:663-684) for the next queued edges (n0, n1) in one
    launch: a collision-free edge adds each node to the other's whitelist, a colliding one to the
    blacklist, exactly what the per-edge loop records when it reaches the edge.  The loop checks a
    pair in the direction it reaches first and caches the result for both directions, so only the
    first occurrence of each pair is checked, in its own direction; pairs already cached are
    skipped.  Returns the number of edges checked."""
    todo, seen = [], set()
    for a, b in edges:
        key = (a, b) if a < b else (b, a)
        if key in seen or b in whitelist.get(a, ()) or b in blacklist.get(a, ()):
            continue
        seen.add(key)
        todo.append((a, b))
    if not todo:
        return 0
    A = _rows([qs[a] for a, _ in todo])
    B = _rows([qs[b] for _, b in todo])
    free = env.batch_is_edge_collision_free(A, B, modes=[modes[a] for a, _ in todo], resolution=resolution)
    for (a, b), f in zip(todo, free):
        cache = whitelist if f else blacklist
        cache.setdefault(a, set()).add(b)
        cache.setdefault(b, set()).add(a)
    return len(todo)


def sparse_checks_batched(env, starts: Sequence, ends: Sequence, modes: Sequence, N_start: int, N_max: int,
                          Ns: Sequence[int]) -> np.ndarray:
    """The sparse edge checks of EIT* (planner_eitstar.py:630-665): states [N_start, N_max) of the
    binary order of each edge, with the edge's own N, in one launch.  Edges with N_start > N are
    collision free there without a check (the reference's ``valid_check``)."""
    Ns = np.asarray(Ns, dtype=np.int64)
    out = np.ones(len(starts), dtype=bool)
    sel = np.nonzero(Ns >= N_start)[0]
    if len(sel) == 0:
        return out
    A = _rows([starts[i] for i in sel])
    B = _rows([ends[i] for i in sel])
    k = len(sel)
    out[sel] = env.batch_is_edge_collision_free(A, B, modes=[modes[i] for i in sel],
                                                N_start=np.full(k, N_start), N_max=np.full(k, N_max),
                                                N=Ns[sel])
    return out
