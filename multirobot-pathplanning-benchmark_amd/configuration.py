"""Configuration types of the planners' collision API (mirror of core/configuration.py).

``NpConfiguration`` keeps the reference's layout: one flat float64 vector ``q`` and, per robot, a
slice ``[s, e)`` (core/configuration.py:222-296).  ``config_dist(..., "max")`` is the metric that
sets the number of interpolation states of an edge (configuration.py:174-197, :328-339).
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import List, Sequence, Tuple, Union

import numpy as np
from numpy.typing import NDArray


class Configuration(ABC):
    """Base class of a multi-robot configuration (core/configuration.py:12-40)."""

    @abstractmethod
    def num_agents(self) -> int: ...

    def __getitem__(self, ind):
        return self.robot_state(ind)

    @abstractmethod
    def robot_state(self, ind: int) -> NDArray: ...

    @abstractmethod
    def state(self) -> NDArray: ...


class NpConfiguration(Configuration):
    """Flat float64 vector + per-robot slices (core/configuration.py:222-296)."""

    __slots__ = ("_array_slice", "q", "_num_agents")

    def __init__(self, q: NDArray, _slice: List[Tuple[int, int]] | NDArray):
        self._array_slice = np.array(_slice)
        self.q = np.asarray(q).astype(np.float64)
        self._num_agents = len(self._array_slice)

    def num_agents(self) -> int:
        return self._num_agents

    def __setitem__(self, ind, data):
        s, e = self._array_slice[ind]
        self.q[s:e] = data

    @classmethod
    def from_list(cls, q_list: Sequence[Union[List[float], NDArray]]) -> "NpConfiguration":
        slices = []
        s = 0
        for v in q_list:
            slices.append((s, s + len(v)))
            s += len(v)
        return cls(np.concatenate([np.asarray(v, dtype=np.float64) for v in q_list]), slices)

    def from_flat(self, q: NDArray) -> "NpConfiguration":
        assert q.shape == self.q.shape, "Shape mismatch"
        return NpConfiguration(q, self._array_slice.copy())

    @classmethod
    def from_numpy(cls, arr: NDArray) -> "NpConfiguration":
        return cls(arr, [(0, len(arr))])

    def robot_state(self, ind: int) -> NDArray:
        if self._num_agents == 1:
            return self.q
        s, e = self._array_slice[ind]
        return self.q[s:e]

    def state(self) -> NDArray:
        return self.q

    def as_list(self) -> List:
        return [self.robot_state(i) for i in range(self._num_agents)]

    @classmethod
    def _batch_dist(cls, pt: "NpConfiguration", batch_other, metric: str = "euclidean") -> NDArray:
        if isinstance(batch_other, np.ndarray):
            diff = pt.q - batch_other
        else:
            diff = pt.q - np.array([o.q for o in batch_other], dtype=np.float64)
        diff = np.atleast_2d(diff)
        if metric == "euclidean":
            return np.sqrt(np.sum(diff * diff, axis=1))
        if metric in ("sum_euclidean", "max_euclidean"):
            d = np.stack([np.sqrt(np.sum(diff[:, s:e] ** 2, axis=1)) for s, e in pt._array_slice])
            return d.sum(axis=0) if metric == "sum_euclidean" else d.max(axis=0)
        # "max": max_k |diff_k| over all coordinates (compute_abs_max_reduction)
        return np.max(np.abs(diff), axis=1)

    @classmethod
    def _dist(cls, pt, other, metric: str = "euclidean") -> float:
        return float(cls._batch_dist(pt, [other], metric)[0])


def config_dist(q_start: Configuration, q_end: Configuration, metric: str = "max") -> float:
    """Distance between two configurations; metric in euclidean, sum_euclidean, max_euclidean, max."""
    return type(q_start)._dist(q_start, q_end, metric)


def edge_num_states(q1: NDArray, q2: NDArray, resolution: float) -> int:
    """N = max(2, int(max_k |q1_k - q2_k| / resolution) + 1) (abstract_env.py:321-323)."""
    m = float(np.max(np.abs(np.asarray(q1, dtype=np.float64) - np.asarray(q2, dtype=np.float64))))
    return max(2, int(m / resolution) + 1)
