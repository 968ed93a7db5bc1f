"""The planners' collision API (mirror of problems/planning_env.py, collision part).

* ``generate_binary_search_indices`` — breadth-first midpoint order of edge states
  (planning_env.py:34-51).
* ``Mode`` / ``State`` — the types the collision calls receive (planning_env.py:102-193); a mode
  of an engine scene carries ``mrcc_mode_id``, the index of its scene graph on the device.
* ``BaseProblem`` — the abstract collision contract (planning_env.py:1531-1881); its default
  ``is_path_collision_free`` gives the reference's boolean from the per-call checks, so backends
  that only provide those behave like the reference.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from collections import deque
from functools import cache
from typing import Any, Dict, List, Optional

import numpy as np
from numpy.typing import NDArray

from .configuration import Configuration, edge_num_states


@cache
def generate_binary_search_indices(N: int):
    """BFS over intervals (s, e) emitting (s + e) // 2 (planning_env.py:34-51)."""
    sequence = [0] * N
    queue = deque([(0, N - 1)])
    idx = 0
    while queue:
        start, end = queue.popleft()
        mid = (start + end) // 2
        sequence[idx] = mid
        idx += 1
        if start <= mid - 1:
            queue.append((start, mid - 1))
        if mid + 1 <= end:
            queue.append((mid + 1, end))
    return tuple(sequence)


class Mode:
    """Discrete mode: task ids + scene graph (planning_env.py:102-172)."""

    id_counter = 0

    def __init__(self, task_list, entry_configuration, mrcc_mode_id: int = 0):
        self.task_ids = task_list
        self.entry_configuration = entry_configuration
        self.prev_mode = None
        self.sg: Dict[str, tuple] = {}
        self.next_modes: List["Mode"] = []
        self.id = Mode.id_counter
        Mode.id_counter += 1
        self.additional_hash_info = None
        self.mrcc_mode_id = int(mrcc_mode_id)

    def __repr__(self):
        return f"Tasks: {self.task_ids}, id: {self.id}"

    def __eq__(self, other):
        if not isinstance(other, Mode):
            return False
        if self.id == other.id:
            return True
        return self.task_ids == other.task_ids and self.mrcc_mode_id == other.mrcc_mode_id

    def __hash__(self):
        return hash((tuple(self.task_ids), self.mrcc_mode_id))


class State:
    """A configuration together with its mode (planning_env.py:175-193)."""

    def __init__(self, q: Configuration, m: Mode):
        self.q = q
        self.mode = m

    def to_dict(self):
        return {"q": self.q.state().tolist(), "mode": self.mode.task_ids}


class BaseProblem(ABC):
    """Collision contract of a planning problem (planning_env.py:1531-1881)."""

    robots: List[str]
    robot_dims: Dict[str, int]
    robot_idx: Dict[str, List[int]]
    start_pos: Configuration
    start_mode: Mode
    limits: NDArray
    collision_tolerance: float
    collision_resolution: float
    cost_metric: str = "euclidean"
    cost_reduction: str = "max"

    def __init__(self):
        self._uniform_sampler = self._make_uniform_sampler()

    def _make_uniform_sampler(self, batch_size=1000):
        """planning_env.py:1697-1705: batches of uniform samples in the joint limits."""
        while True:
            batch = np.random.uniform(low=self.limits[0, :], high=self.limits[1, :],
                                      size=(batch_size, self.limits.shape[1]))
            for i in range(batch_size):
                yield self.start_pos.from_flat(batch[i])

    def sample_config_uniform_in_limits(self) -> Configuration:
        return next(self._uniform_sampler)

    @abstractmethod
    def get_scenegraph_info_for_mode(self, mode: Mode, is_start_mode: bool = False) -> Dict: ...

    @abstractmethod
    def set_to_mode(self, mode: Mode) -> None: ...

    @abstractmethod
    def is_collision_free(self, q: Optional[Configuration], mode: Optional[Mode]) -> bool: ...

    def is_collision_free_for_robot(self, r: List[str] | str, q: NDArray, m: Mode,
                                    collision_tolerance: float = 0.01, set_mode: bool = True) -> bool:
        raise NotImplementedError

    @abstractmethod
    def is_edge_collision_free(self, q1: Configuration, q2: Configuration, m: Mode,
                               resolution: float | None = None, tolerance: float | None = None,
                               include_endpoints: bool = False, N_start: int = 0,
                               N_max: int | None = None, N: int | None = None) -> bool: ...

    def is_path_collision_free(self, path: List[State], binary_order: bool = True,
                               resolution: Optional[float] = None, tolerance: Optional[float] = None,
                               check_edges_in_order: bool = False,
                               check_start_and_end: bool = True) -> bool:
        """The boolean of planning_env.py:1765-1881 from the per-call checks.

        Both of the reference's schedules test every interior state of every edge; they differ in
        the order of the calls and in the vertices: the first vertex unless check_start_and_end is
        False, every inner vertex, and the last vertex when check_start_and_end or
        check_edges_in_order.  Here the vertices are checked first, then the edges in path order
        (one is_edge_collision_free call each); the first failure decides.  Backends override this
        with batched forms (MrccEnvironment issues one launch for the vertices and one for the
        edges)."""
        if resolution is None:
            resolution = self.collision_resolution
        if tolerance is None:
            tolerance = self.collision_tolerance
        K = len(path)
        if K < 2:
            raise IndexError("a path needs at least two states")
        verts = [i for i in range(K - 1) if i > 0 or check_start_and_end]
        if check_start_and_end or check_edges_in_order:
            verts.append(K - 1)
        for i in verts:
            if not self.is_collision_free(path[i].q, path[i].mode):
                return False
        for i in range(K - 1):
            if not self.is_edge_collision_free(path[i].q, path[i + 1].q, path[i].mode, resolution=resolution,
                                               tolerance=tolerance, include_endpoints=False):
                return False
        return True


# ---------------------------------------------------------------------------------------------
# registry (core/registry.py:8-38)
REGISTRY: Dict[str, Any] = {}


def register(names):
    """``names``: a string, or a list of (name, kwargs) pairs."""

    def decorator(cls):
        if isinstance(names, str):
            REGISTRY[names] = lambda: cls()
        else:
            for name, kwargs in names:
                if name in REGISTRY:
                    raise ValueError(f"Duplicate env name: {name}")
                REGISTRY[name] = lambda cls=cls, kwargs=kwargs: cls(**kwargs)
        return cls

    return decorator


def get_env_by_name(name: str) -> BaseProblem:
    try:
        return REGISTRY[name]()
    except KeyError:
        raise ValueError(f"Unknown environment: {name}")


def list_envs(prefix: str = ""):
    return sorted(k for k in REGISTRY if k.startswith(prefix))
