"""Autotuner of the batched edge check.

Times candidate execution shapes (states per pass G, queue slots per ticket, persistent waves per
CU, and the single launch without a queue: one workgroup per edge walking its states in rounds,
which wins on small batches of long edges that mostly collide early) on the caller's own batch
with HIP events and keeps the fastest.  Choices are cached in a JSON
file keyed by the GPU architecture and a workload signature (scene kind, dimension, sphere count,
edge count and mean state count, both rounded to powers of two), so a process pays the search
once per workload shape.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Optional

import numpy as np

from .backend import EdgePlan, make_tune

CACHE = os.environ.get("MRCC_AUTOTUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "mrcc_autotune.json"))

DIRECT = dict(persistent=0, sorted=0)  # the single launch: no queue, one workgroup per edge
LANE_GRID = [dict(slots_per_pass=g, slots_per_ticket=t, waves_per_cu=w)
             for g in (4, 8, 16, 32) for t in (1, 2, 4) for w in (8, 16)] + [DIRECT]
ARM_GRID = [dict(slots_per_pass=g, slots_per_ticket=t, waves_per_cu=w)
            for g in (1, 2, 4, 8) for t in (8, 16, 32) for w in (12, 16)] + [DIRECT]


def candidate_name(cand: dict) -> str:
    if not cand.get("persistent", 1):
        return "direct"
    return "G{slots_per_pass}/t{slots_per_ticket}/w{waves_per_cu}".format(**cand)


def _p2(x: float) -> int:
    return 1 << max(0, int(round(math.log2(max(x, 1.0)))))


def signature(scene, n_edges: int, mean_states: float) -> str:
    import torch

    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    S = getattr(scene.desc, "link_sphere_count", None)
    S = int(np.sum(S)) if S is not None and len(S) else 0
    return f"{arch}|k{scene.kind}|D{scene.D}|S{S}|n{_p2(n_edges)}|m{_p2(mean_states)}"


def _load() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(db: Dict[str, dict]) -> None:
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def autotune(scene, batch, mean_states: float, reps: int = 3, grid=None, use_cache: bool = True,
             verbose: bool = False):
    """Returns (Tune, timings) for `batch` (an EdgeBatch over device tensors)."""
    import torch

    key = signature(scene, batch.n_edges, mean_states)
    db = _load() if use_cache else {}
    if use_cache and key in db:
        return make_tune(**db[key]["tune"]), db[key].get("timings", {})
    if grid is None:
        grid = ARM_GRID if scene.kind == 3 else LANE_GRID
    timings = {}
    best = None
    for cand in grid:
        plan = EdgePlan(scene, batch, make_tune(use_graph=1, **cand))
        plan.run()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        name = candidate_name(cand)
        timings[name] = ms
        if verbose:
            print(f"[autotune] {name}: {ms:.3f} ms", flush=True)
        if best is None or ms < best[0]:
            best = (ms, cand)
        del plan
    if use_cache:
        db[key] = {"tune": best[1], "ms": best[0], "timings": timings}
        _store(db)
    return make_tune(use_graph=1, **best[1]), timings
