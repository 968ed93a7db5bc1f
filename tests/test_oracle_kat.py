"""The oracle against the reference's own known answers (tests/test.py) and closed forms."""
import math

import numpy as np
import pytest

from helpers import O, pkg, scenes


@pytest.mark.parametrize("n, expected", [(1, (0,)), (2, (0, 1)), (3, (1, 0, 2)), (4, (1, 0, 2, 3)),
                                         (5, (2, 0, 3, 1, 4))])
def test_binary_indices(n, expected):
    """tests/test.py:36-47."""
    assert O.bfs_indices(n) == expected
    assert pkg.generate_binary_search_indices(n) == expected


def test_binary_indices_permutation():
    for n in range(1, 300):
        idx = O.bfs_indices(n)
        assert sorted(idx) == list(range(n))
        assert idx == pkg.generate_binary_search_indices(n)


def test_edge_checking_kat():
    """tests/test.py:50-58: edge [-1,0,1,1] -> [-1,1,1,0] is free on abstract.test."""
    o = O.Scene(scenes.abstract_test_scene())
    assert o.edge_free(np.array([-1, 0, 1, 1.0]), np.array([-1, 1, 1, 0.0]), 0.01)


@pytest.mark.parametrize("res, incl, calls", [(0.5, True, 3), (0.5, False, 1), (0.1, False, 9), (0.1, True, 11)])
def test_edge_checking_resolution_calls(res, incl, calls):
    """tests/test.py:61-90: number of is_collision_free calls."""
    o = O.Scene(scenes.abstract_test_scene())
    ok, n = o.edge_free(np.array([-1, 0, 1, 1.0]), np.array([-1, 1, 1, 0.0]), res, include_endpoints=incl,
                        count=True)
    assert ok and n == calls
    N = O.edge_N(np.array([-1, 0, 1, 1.0]), np.array([-1, 1, 1, 0.0]), res)
    assert O.edge_num_states(N, 0, None, incl) == calls
