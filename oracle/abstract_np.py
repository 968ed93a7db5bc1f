"""numpy restatement of the reference's abstract environment, for small cases.

TEST INFRASTRUCTURE ONLY (a second, independent CPU restatement used to pin the C oracle).
It follows src/multi_robot_multi_goal_planning/problems/abstract_env.py with the same numpy
operations, so its rounding is numpy's own:
  Sphere.collides_with_sphere      abstract_env.py:46-49    np.linalg.norm(pos - c) < r_o + r
  Rectangle.__init__ / collides     abstract_env.py:61-84    clip, (..)**2, np.sum, <= r**2
  AbstractEnvironment.is_collision_free   :255-276
  AbstractEnvironment.is_edge_collision_free   :301-354 (loop over the BFS order)
"""
from __future__ import annotations

from collections import deque

import numpy as np


def bfs(N):
    seq = [0] * N
    q = deque([(0, N - 1)])
    i = 0
    while q:
        s, e = q.popleft()
        m = (s + e) // 2
        seq[i] = m
        i += 1
        if s <= m - 1:
            q.append((s, m - 1))
        if m + 1 <= e:
            q.append((m + 1, e))
    return tuple(seq)


class AbstractNp:
    def __init__(self, desc):
        self.rd = desc.robot_dim
        self.radii = [float(r) for r in desc.robot_radius]
        self.spheres = [(np.asarray(c, dtype=np.float64), float(r))
                        for c, r in zip(desc.sphere_center, desc.sphere_radius)]
        self.rects = []
        for c, b in zip(desc.rect_center, desc.rect_size):
            c = np.asarray(c, dtype=np.float64)
            b = np.asarray(b, dtype=np.float64)
            self.rects.append((c - b / 2, c + b / 2))

    def robot(self, q, i):
        return q[i * self.rd:(i + 1) * self.rd]

    def is_collision_free(self, q, mask=0xFFFFFFFF):
        n = len(self.radii)
        for i in range(n):
            for j in range(i + 1, n):
                if not ((mask >> i) & 1 and (mask >> j) & 1):
                    continue
                if np.linalg.norm(self.robot(q, i) - self.robot(q, j)) < self.radii[i] + self.radii[j]:
                    return False
        for i in range(n):
            if not (mask >> i) & 1:
                continue
            pos = self.robot(q, i)
            for c, r in self.spheres:
                if np.linalg.norm(c - pos) < r + self.radii[i]:
                    return False
            for lo, hi in self.rects:
                center = np.array(pos, dtype=float)
                closest = np.clip(center, lo, hi)
                if np.sum((closest - center) ** 2) <= self.radii[i] ** 2:
                    return False
        return True

    def is_edge_collision_free(self, q1, q2, resolution, include_endpoints=False, N_start=0, N_max=None,
                               N=None):
        if N is None:
            N = int(np.max(np.abs(q1 - q2)) / resolution) + 1
            N = max(2, N)
        if N_start > N:
            assert False
        if N_max is None:
            N_max = N
        N_max = min(N, N_max)
        idx = bfs(N)
        d = (q2 - q1) / (N - 1)
        for i in idx[N_start:N_max]:
            if not include_endpoints and (i == 0 or i == N - 1):
                continue
            if not self.is_collision_free(q1 + d * i):
                return False
        return True
