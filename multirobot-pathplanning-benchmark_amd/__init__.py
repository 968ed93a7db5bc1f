"""MI355X (gfx950) batched collision-checking engine for multi-robot motion planning.

A new backend behind the collision API of vhartman/multirobot-pathplanning-benchmark
(problems/planning_env.py:1724-1881): configuration validity, interpolated-edge validity and the
forward kinematics feeding them run in hand-written HIP kernels (libmrcc.so, C ABI in
include/mrcc.h), and so do the planners' configuration costs and distances
(core/configuration.py); this package is the Python host side (ctypes), mirroring the reference's
``BaseProblem`` interface and configuration functions.

The directory name is not a Python identifier; import it with
``importlib.import_module("multirobot-pathplanning-benchmark_amd")``.
"""
from . import scenes
from ._lib import LIB_PATH, MrccError, lib
from .configuration import (Configuration, NpConfiguration, batch_config_cost, batch_config_dist, config_cost,
                            config_dist, edge_num_states)
from .planning_env import (BaseProblem, Mode, State, generate_binary_search_indices, get_env_by_name,
                           list_envs, register)

__all__ = [
    "LIB_PATH", "MrccError", "lib", "scenes", "Configuration", "NpConfiguration", "config_dist",
    "batch_config_dist", "config_cost", "batch_config_cost", "edge_num_states", "BaseProblem", "Mode", "State", "generate_binary_search_indices",
    "get_env_by_name", "list_envs", "register", "backend", "envs",
]


def __getattr__(name):
    # the backend (and the registered environments) load libmrcc lazily
    import importlib

    if name in ("backend", "envs", "clearance", "autotune", "callers"):
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(name)
