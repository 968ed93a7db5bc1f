"""A BaseProblem over the CPU oracle (test infrastructure): per-call checks only, the edge check as
the reference's per-state loop (abstract_env.py:301-354) through is_collision_free, and modes by
their device mode id (an int or a Mode's mrcc_mode_id)."""
import numpy as np

from helpers import O, pkg


class OracleEnv(pkg.BaseProblem):
    def __init__(self, desc, resolution=0.01):
        self.desc = desc
        self.o = O.Scene(desc)
        self.limits = desc.limits
        sl = desc.robot_slices()
        self.start_pos = pkg.NpConfiguration(np.zeros(desc.dim), sl)
        self.start_mode = pkg.Mode([0] * desc.n_robots, self.start_pos)
        self.collision_resolution = resolution
        self.collision_tolerance = 0.01
        pkg.BaseProblem.__init__(self)

    @staticmethod
    def _mode(m):
        if m is None:
            return 0
        return int(m) if isinstance(m, (int, np.integer)) else int(getattr(m, "mrcc_mode_id", 0))

    def get_scenegraph_info_for_mode(self, mode, is_start_mode=False):
        return {}

    def set_to_mode(self, mode):
        pass

    def is_collision_free(self, q, mode):
        return self.o.config_free(q.state(), self._mode(mode))

    def is_edge_collision_free(self, q1, q2, m, resolution=None, tolerance=None, include_endpoints=False,
                               N_start=0, N_max=None, N=None):
        if resolution is None:
            resolution = self.collision_resolution
        if N is None:
            N = pkg.edge_num_states(q1.state(), q2.state(), resolution)
        if N_start > N:
            assert False
        N_max = N if N_max is None else min(N, N_max)
        idx = pkg.generate_binary_search_indices(N)
        d = (q2.state() - q1.state()) / (N - 1)
        for i in idx[N_start:N_max]:
            if not include_endpoints and (i == 0 or i == N - 1):
                continue
            if not self.is_collision_free(q1.from_flat(q1.state() + d * i), m):
                return False
        return True
