"""Configuration types and the planners' cost / distance functions (mirror of core/configuration.py).

``NpConfiguration`` keeps the reference's layout: one flat float64 vector ``q`` and, per robot, a
slice ``[s, e)`` (core/configuration.py:222-296).  ``batch_config_cost`` / ``config_cost``
(:437-510, :352-363) and ``batch_config_dist`` / ``config_dist`` / ``NpConfiguration._batch_dist``
(:303-349) run on the device (``mrcc_config_cost``, one kernel launch per call; numpy inputs are
copied to the device and the result back, device tensors stay on the device).  Their rounding is
the reference's numba loops': per-robot sums of squares in coordinate order with each step fused,
``max + w * sum`` fused.  ``edge_num_states`` is the host's exact N rule of an edge
(abstract_env.py:321-323), used by the path schedule.
"""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import List, Sequence, Tuple, Union

import ctypes

import numpy as np
from numpy.typing import NDArray

from . import _lib


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
    def _batch_dist(cls, pt: "NpConfiguration", batch_other, metric: str = "euclidean"):
        """Distances from pt to each row of batch_other (configuration.py:303-329):
        "euclidean" over all coordinates, "sum_euclidean" / "max_euclidean" over the robots'
        slices, otherwise the largest absolute coordinate difference (compute_abs_max_reduction)."""
        if isinstance(batch_other, np.ndarray) or _is_device_tensor(batch_other):
            other = batch_other
        else:
            other = np.array([o.q for o in batch_other], dtype=np.float64)
        if metric == "euclidean":
            return _config_cost(pt.q, other, [[0, int(pt._array_slice[-1][-1])]], _lib.COST_EUCLIDEAN,
                                _lib.REDUCE_MAX, 0.0)
        if metric in ("sum_euclidean", "max_euclidean"):
            red = _lib.REDUCE_SUM if metric == "sum_euclidean" else _lib.REDUCE_MAX
            return _config_cost(pt.q, other, pt._array_slice, _lib.COST_EUCLIDEAN, red, 0.0)
        return _config_cost(pt.q, other, None, _lib.COST_ABSMAX, _lib.REDUCE_MAX, 0.0)

    @classmethod
    def _dist(cls, pt, other, metric: str = "euclidean") -> float:
        return float(cls._batch_dist(pt, [other], metric)[0])


def _is_device_tensor(x) -> bool:
    return type(x).__module__.startswith("torch") and getattr(x, "is_cuda", False)


def _config_cost(P, Q, slices, metric: int, reduction: int, w: float):
    """One mrcc_config_cost launch: rows of Q against P (one configuration, broadcast, or one row
    per row of Q).  slices None: one slice over every coordinate of Q.  numpy in -> numpy out;
    device tensors in -> a device tensor out (on the current stream)."""
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("batch_config_cost / batch_config_dist need a HIP device")
    on_device = _is_device_tensor(Q)
    dev = torch.device("cuda", torch.cuda.current_device())

    def as_dev(x):
        if _is_device_tensor(x):
            t = x if x.dtype == torch.float64 else x.to(torch.float64)
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).to(dev)
        if t.dim() >= 1 and t.stride(-1) != 1:
            t = t.contiguous()
        return t

    Qd = as_dev(Q)
    if Qd.dim() == 1:
        Qd = Qd.reshape(1, -1)
    Pd = as_dev(P)
    n, D = Qd.shape
    if Pd.dim() == 1:
        if Pd.shape[0] != D:
            raise ValueError(f"configuration of dimension {Pd.shape[0]} against rows of dimension {D}")
        ps = 0
    else:
        if tuple(Pd.shape) != (n, D):
            raise ValueError(f"shape mismatch: {tuple(Pd.shape)} vs {(n, D)}")
        ps = Pd.stride(0)
    sl = np.ascontiguousarray(np.asarray([[0, D]] if slices is None else slices, dtype=np.int32).reshape(-1))
    out = torch.empty(n, dtype=torch.float64, device=Qd.device)
    if n:
        _lib.check(_lib.lib().mrcc_config_cost(
            ctypes.c_void_p(Pd.data_ptr()), ps, ctypes.c_void_p(Qd.data_ptr()), Qd.stride(0), n, D,
            sl.ctypes.data, len(sl) // 2, metric, reduction, float(w),
            ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
    return out if on_device else out.cpu().numpy()


def config_dist(q_start: Configuration, q_end: Configuration, metric: str = "max") -> float:
    """Distance between two configurations; metric in euclidean, sum_euclidean, max_euclidean, max
    (configuration.py:332-339)."""
    return type(q_start)._dist(q_start, q_end, metric)


def batch_config_dist(pt: Configuration, batch_pts, metric: str = "max"):
    """Distances from pt to each of batch_pts (configuration.py:342-349)."""
    return type(pt)._batch_dist(pt, batch_pts, metric)


def config_cost(q_start: Configuration, q_end: Configuration, metric: str = "max",
                reduction: str = "max") -> float:
    """Cost between two configurations (configuration.py:352-363): metric "max" or "euclidean"
    per robot, reduction "max" (max + 0.01 sum) or "sum"."""
    return float(batch_config_cost(q_start, q_end.state()[None, :], metric, reduction)[0])


def batch_config_cost(starts, batch_other, metric: str = "max", reduction: str = "max", w: float = 0.01,
                      tmp_agent_slice=None):
    """Costs of configuration pairs (configuration.py:437-510), with the reference's input forms:
    (Configuration, rows), (rows, rows) with tmp_agent_slice, (list of rows, None) with
    tmp_agent_slice for the consecutive pairs of a path, (list, list) of Configurations or of
    States.  Per robot slice: the largest absolute difference (metric "max", np.max) or the
    Euclidean norm ("euclidean"); reduced by max + w * sum ("max") or sum ("sum")."""
    if isinstance(starts, Configuration) and (isinstance(batch_other, np.ndarray) or _is_device_tensor(batch_other)):
        P, Q, slices = starts.state(), batch_other, starts._array_slice
    elif (isinstance(starts, np.ndarray) or _is_device_tensor(starts)) and \
            (isinstance(batch_other, np.ndarray) or _is_device_tensor(batch_other)):
        P, Q, slices = starts, batch_other, tmp_agent_slice
    elif batch_other is None and tmp_agent_slice is not None:
        if isinstance(starts[0], np.ndarray):
            arr = np.array(starts)
            P, Q, slices = arr[1:, :], arr[:-1, :], tmp_agent_slice
        elif _is_device_tensor(starts):
            P, Q, slices = starts[1:, :], starts[:-1, :], tmp_agent_slice
        else:
            raise ValueError
    elif isinstance(starts, list) and isinstance(batch_other, list):
        if isinstance(starts[0], Configuration):
            P = np.array([x.state() for x in starts])
            Q = np.array([x.state() for x in batch_other], dtype=np.float64)
            slices = starts[0]._array_slice
        else:
            P = np.array([x.q.state() for x in starts])
            Q = np.array([x.q.state() for x in batch_other], dtype=np.float64)
            slices = starts[0].q._array_slice
    else:
        raise ValueError
    if reduction not in ("max", "sum"):
        raise ValueError
    code = _lib.COST_EUCLIDEAN if metric == "euclidean" else _lib.COST_MAX
    red = _lib.REDUCE_MAX_SUM if reduction == "max" else _lib.REDUCE_SUM
    return _config_cost(P, Q, np.asarray(slices), code, red, w)


def edge_num_states(q1: NDArray, q2: NDArray, resolution: float) -> int:
    """N = max(2, int(config_dist(q1, q2, "max") / resolution) + 1) (abstract_env.py:321-323), the
    distance by compute_abs_max_reduction's rule (configuration.py:174-197): the largest |q1_k - q2_k|,
    a NaN counting only in coordinate 0.  Exact, so computed on the host."""
    d = np.abs(np.asarray(q1, dtype=np.float64).reshape(-1) - np.asarray(q2, dtype=np.float64).reshape(-1))
    m = float(d[0]) if np.isnan(d[0]) else float(np.nanmax(d))
    return max(2, int(m / resolution) + 1)
